// Python bindings: plainmp_amd._native with the submodules of the reference
// (bindings/python.cpp:15-21): kinematics, primitive_sdf, constraint, kdtree,
// ompl.  Method names and argument meaning follow the reference bindings;
// batched methods are added next to them.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "constraint.hpp"

namespace py = pybind11;
using namespace pmp_host;

namespace {

using DArr = py::array_t<double, py::array::c_style | py::array::forcecast>;

Vec3 vec3(const DArr& a) {
  if (a.size() != 3) throw std::invalid_argument("expected 3 values");
  return {a.data()[0], a.data()[1], a.data()[2]};
}
Mat3 mat3(const DArr& a) {
  if (a.size() != 9) throw std::invalid_argument("expected a 3x3 matrix");
  Mat3 m;
  std::copy(a.data(), a.data() + 9, m.begin());
  return m;
}
DArr arr3(const Vec3& v) {
  DArr a(3);
  std::copy(v.begin(), v.end(), a.mutable_data());
  return a;
}
DArr arr33(const Mat3& m) {
  DArr a({3, 3});
  std::copy(m.begin(), m.end(), a.mutable_data());
  return a;
}
std::vector<double> vecd(const DArr& a) { return std::vector<double>(a.data(), a.data() + a.size()); }
// (B, D) batch of configurations
const double* rows(const DArr& Q, size_t D, int64_t& B) {
  if (Q.ndim() != 2 || (size_t)Q.shape(1) != D)
    throw std::invalid_argument("expected a (B, " + std::to_string(D) + ") array of configurations");
  B = Q.shape(0);
  return Q.data();
}
// (3, N) points
const double* points3n(const DArr& P, int64_t& N) {
  if (P.ndim() != 2 || P.shape(0) != 3) throw std::invalid_argument("expected a (3, N) array of points");
  N = P.shape(1);
  return P.data();
}

}  // namespace

PYBIND11_MODULE(_native, m) {
  m.doc() = "MI355X (gfx950) FK + sphere-SDF collision path with plainmp's API";
  m.def("abi_version", &pmp_abi_version);

  // ------------------------------------------------------------ kinematics
  auto mk = m.def_submodule("kinematics");
  py::enum_<BaseType>(mk, "BaseType")
      .value("FIXED", FIXED)
      .value("PLANAR", PLANAR)
      .value("FLOATING", FLOATING);
  py::class_<KinematicModel, std::shared_ptr<KinematicModel>>(mk, "KinematicModel")
      .def(py::init<const std::string&>())
      .def("add_new_link",
           [](KinematicModel& k, const std::string& name, const std::string& parent, const DArr& pos, const DArr& rpy,
              bool consider_rotation) { return k.add_new_link(name, parent, vec3(pos), vec3(rpy), consider_rotation); })
      .def("debug_get_link_pose",
           [](std::shared_ptr<KinematicModel> k, const std::string& name) {
             // all joints from the model state, composed on the GPU in FP64
             size_t id = k->get_link_ids({name})[0];
             PackedRobot pr;
             k->pack({}, pr);
             pmp_robot_desc d{};
             d.n_nodes = (int32_t)pr.parent.size();
             d.node_parent = pr.parent.data();
             d.node_kind = pr.kind.data();
             d.node_q_col = pr.q_col.data();
             d.node_quat = pr.quat.data();
             d.node_trans = pr.trans.data();
             d.node_axis = pr.axis.data();
             d.node_identity = pr.identity.data();
             auto bp = k->get_base_pose();
             for (int i = 0; i < 7; ++i) d.base_pose[i] = bp[i];
             Device& dev = shared_device();
             check(pmp_upload_robot(dev.ctx(), &d));
             double* dq = (double*)dev.buf(5, 8);
             double* dp = (double*)dev.buf(6, sizeof(double) * 7 * d.n_nodes);
             check(pmp_link_poses_batch(dev.ctx(), dq, 1, 1, 1, dp));
             std::vector<double> h(7 * d.n_nodes);
             dev.d2h(h.data(), dp, h.size() * sizeof(double));
             return std::vector<double>(h.begin() + 7 * id, h.begin() + 7 * id + 7);
           })
      .def("set_joint_positions",
           [](KinematicModel& k, const std::vector<size_t>& ids, const DArr& q, bool accurate) {
             k.set_joint_angles(ids, vecd(q), accurate);
           },
           py::arg("joint_ids"), py::arg("positions"), py::arg("accurate") = true)
      .def("get_joint_positions", &KinematicModel::get_joint_angles)
      .def("set_base_pose",
           [](KinematicModel& k, const DArr& v) {
             if (v.size() != 7) throw std::invalid_argument("base pose is [x, y, z, qx, qy, qz, qw]");
             std::array<double, 7> p;
             std::copy(v.data(), v.data() + 7, p.begin());
             k.set_base_pose(p);
           })
      .def("get_base_pose",
           [](const KinematicModel& k) {
             auto p = k.get_base_pose();
             DArr a(7);
             std::copy(p.begin(), p.end(), a.mutable_data());
             return a;
           })
      .def("get_joint_position_limits", &KinematicModel::get_joint_position_limits)
      .def("get_link_ids", &KinematicModel::get_link_ids)
      .def("get_joint_ids", &KinematicModel::get_joint_ids)
      .def_property_readonly("n_links", &KinematicModel::n_links)
      .def_property_readonly("n_joints", &KinematicModel::n_joints);

  // --------------------------------------------------------- primitive_sdf
  auto ms = m.def_submodule("primitive_sdf");
  py::class_<Pose>(ms, "Pose")
      .def(py::init([](const DArr& p, const DArr& R) { return Pose(vec3(p), mat3(R)); }),
           py::arg("position") = [] {
             DArr z(3);
             std::fill(z.mutable_data(), z.mutable_data() + 3, 0.0);
             return z;
           }(),
           py::arg("rotation") = [] {
             DArr I({3, 3});
             auto r = mat_identity();
             std::copy(r.begin(), r.end(), I.mutable_data());
             return I;
           }())
      .def("__copy__", [](const Pose& p) { return Pose(p); })
      .def("__deepcopy__", [](const Pose& p, py::dict) { return Pose(p); })
      .def("translate", [](Pose& p, const DArr& t) { p.translate(vec3(t)); })
      .def("rotate_z", &Pose::rotate_z)
      .def_readonly("axis_aligned", &Pose::axis_aligned)
      .def_readonly("z_axis_aligned", &Pose::z_axis_aligned)
      .def_property_readonly("position", [](const Pose& p) { return arr3(p.position); })
      .def_property_readonly("rotation", [](const Pose& p) { return arr33(p.rot); });

  py::class_<SDFBase, std::shared_ptr<SDFBase>>(ms, "SDFBase")
      .def("clone", &SDFBase::clone)
      .def("translate", [](SDFBase& s, const DArr& t) { s.translate(vec3(t)); })
      .def("rotate_z", &SDFBase::rotate_z)
      .def("evaluate", [](SDFBase& s, const DArr& p) { return s.evaluate(vec3(p)); })
      .def("evaluate_batch",
           [](SDFBase& s, const DArr& P) {
             int64_t N;
             const double* d = points3n(P, N);
             std::vector<double> v;
             {
               py::gil_scoped_release rel;
               v = s.evaluate_batch(d, N);
             }
             DArr out(N);
             std::copy(v.begin(), v.end(), out.mutable_data());
             return out;
           })
      .def("is_outside", [](SDFBase& s, const DArr& p, double r) { return s.is_outside(vec3(p), r); })
      .def("is_outside_batch",
           [](SDFBase& s, const DArr& P, double r) {
             int64_t N;
             const double* d = points3n(P, N);
             std::vector<uint8_t> v;
             {
               py::gil_scoped_release rel;
               v = s.is_outside_batch(d, N, r);
             }
             py::array_t<bool> out(N);
             for (int64_t i = 0; i < N; ++i) out.mutable_data()[i] = v[i] != 0;
             return out;
           })
      .def_property_readonly("sdf_type", [](const SDFBase& s) { return (int)s.get_type(); });
  py::class_<PrimitiveSDFBase, SDFBase, std::shared_ptr<PrimitiveSDFBase>>(ms, "PrimitiveSDFBase")
      .def_property_readonly("lb", [](const PrimitiveSDFBase& s) { return arr3(s.lb); })
      .def_property_readonly("ub", [](const PrimitiveSDFBase& s) { return arr3(s.ub); });
  py::class_<UnionSDF, SDFBase, std::shared_ptr<UnionSDF>>(ms, "UnionSDF")
      .def(py::init<const std::vector<SDFBase::Ptr>&>())
      .def("merge", &UnionSDF::merge, py::arg("other"), py::arg("clone") = false)
      .def("add", &UnionSDF::add, py::arg("sdf"), py::arg("clone") = false)
      .def_property_readonly("sdfs", [](const UnionSDF& u) { return u.sdfs; });
  py::class_<GroundSDF, PrimitiveSDFBase, std::shared_ptr<GroundSDF>>(ms, "GroundSDF").def(py::init<double>());
  py::class_<BoxSDF, PrimitiveSDFBase, std::shared_ptr<BoxSDF>>(ms, "BoxSDF")
      .def(py::init([](const DArr& w, const Pose& p) { return std::make_shared<BoxSDF>(vec3(w), p); }))
      .def_readonly("pose", &BoxSDF::pose);
  py::class_<CylinderSDF, PrimitiveSDFBase, std::shared_ptr<CylinderSDF>>(ms, "CylinderSDF")
      .def(py::init<double, double, const Pose&>())
      .def_readonly("pose", &CylinderSDF::pose);
  py::class_<SphereSDF, PrimitiveSDFBase, std::shared_ptr<SphereSDF>>(ms, "SphereSDF")
      .def(py::init<double, const Pose&>())
      .def_readonly("pose", &SphereSDF::pose);
  py::class_<CloudSDF, PrimitiveSDFBase, std::shared_ptr<CloudSDF>>(ms, "CloudSDF")
      .def(py::init([](const DArr& pts, double r) {
        if (pts.ndim() != 2 || pts.shape(1) != 3) throw std::invalid_argument("expected (N, 3) points");
        return std::make_shared<CloudSDF>(vecd(pts), r);
      }));

  // ----------------------------------------------------------------- kdtree
  auto mt = m.def_submodule("kdtree");
  py::class_<KDTree, std::shared_ptr<KDTree>>(mt, "KDTree")
      .def(py::init([](const DArr& pts) {
        if (pts.ndim() != 2 || pts.shape(1) != 3) throw std::invalid_argument("expected (N, 3) points");
        return std::make_shared<KDTree>(vecd(pts));
      }))
      .def("query", [](const KDTree& t, const DArr& p) { return arr3(t.query(vec3(p))); })
      .def("sqdist", [](const KDTree& t, const DArr& p) { return t.sqdist(vec3(p)); })
      .def("query_batch", [](const KDTree& t, const DArr& Q) {  // (M, 3) -> (points (M, 3), sqdist (M,))
        if (Q.ndim() != 2 || Q.shape(1) != 3) throw std::invalid_argument("expected (M, 3) queries");
        const int64_t M = Q.shape(0);
        std::vector<double> soa(3 * M);
        for (int64_t i = 0; i < M; ++i)
          for (int k = 0; k < 3; ++k) soa[k * M + i] = Q.data()[3 * i + k];
        std::vector<double> sq;
        std::vector<int32_t> idx;
        {
          py::gil_scoped_release rel;
          t.nearest(soa.data(), M, sq, idx);
        }
        DArr P({M, (int64_t)3}), S(M);
        for (int64_t i = 0; i < M; ++i) {
          for (int k = 0; k < 3; ++k) P.mutable_data()[3 * i + k] = t.points()[3 * idx[i] + k];
          S.mutable_data()[i] = sq[i];
        }
        return std::make_pair(P, S);
      });
  // The cloud's tile tables built on the device (pmp_cloud_build_tiles).
  mt.def("cloud_tile_sizes", [](int64_t n) {
    int64_t nt = 0, ns = 0, scr = 0;
    check(pmp_cloud_tile_sizes(n, &nt, &ns, &scr));
    return py::make_tuple(nt, ns, scr);
  });
  mt.def("build_cloud_tiles", [](const DArr& pts) {  // (N, 3) -> dict of the tables as bytes
    if (pts.ndim() != 2 || pts.shape(1) != 3) throw std::invalid_argument("expected (N, 3) points");
    const int64_t n = pts.shape(0);
    int64_t nt = 0, ns = 0, scr = 0;
    check(pmp_cloud_tile_sizes(n, &nt, &ns, &scr));
    Device& dev = shared_device();
    const size_t sz[5] = {(size_t)nt * 48, (size_t)nt * 48, (size_t)nt * 256 * sizeof(float),
                          (size_t)nt * 256 * sizeof(double), (size_t)ns * 32};
    void* out[5];
    for (int i = 0; i < 5; ++i) out[i] = dev.buf(10 + i, sz[i]);
    void* dp = dev.buf(15, (size_t)n * 3 * sizeof(double));
    void* ds = dev.buf(16, (size_t)scr);
    std::vector<std::string> host(5);
    {
      py::gil_scoped_release rel;
      dev.h2d(dp, pts.data(), (size_t)n * 3 * sizeof(double));
      check(pmp_cloud_build_tiles(dev.ctx(), (const double*)dp, n, out[0], out[1], (float*)out[2], (double*)out[3],
                                  out[4], ds));
      check(pmp_synchronize(dev.ctx()));
      for (int i = 0; i < 5; ++i) {
        host[i].resize(sz[i]);
        dev.d2h(host[i].data(), out[i], sz[i]);
      }
    }
    const char* names[5] = {"tiles", "tiles64", "tile_b", "points64", "supers"};
    py::dict d;
    for (int i = 0; i < 5; ++i) d[names[i]] = py::bytes(host[i]);
    return d;
  });
  mt.def("_dev_cloud_build_tiles", [](uintptr_t pts, int64_t n, uintptr_t tiles, uintptr_t tiles64, uintptr_t tile_b,
                                      uintptr_t points64, uintptr_t supers, uintptr_t scratch, uintptr_t stream) {
    Device& dev = shared_device();
    check(pmp_set_stream(dev.ctx(), (void*)stream));
    const pmp_status s = pmp_cloud_build_tiles(dev.ctx(), (const double*)pts, n, (void*)tiles, (void*)tiles64,
                                               (float*)tile_b, (double*)points64, (void*)supers, (void*)scratch);
    check(pmp_set_stream(dev.ctx(), nullptr));
    check(s);
  });

  // ------------------------------------------------------------- constraint
  auto mc = m.def_submodule("constraint");
  py::class_<SphereAttachmentSpec>(mc, "SphereAttachmentSpec")
      .def(py::init([](const std::string& link, const DArr& rel, const DArr& radii, bool only_self) {
             // relative_positions is (3, k) as Eigen::Matrix3Xd in the reference
             if (rel.ndim() != 2 || rel.shape(0) != 3) throw std::invalid_argument("relative_positions must be (3, k)");
             const int64_t k = rel.shape(1);
             if (radii.size() != k) throw std::invalid_argument("radii must have k entries");
             SphereAttachmentSpec s;
             s.parent_link_name = link;
             for (int64_t j = 0; j < k; ++j)
               for (int a = 0; a < 3; ++a) s.relative_positions.push_back(rel.data()[a * k + j]);
             s.radii = vecd(radii);
             s.only_self_collision = only_self;
             return s;
           }),
           py::arg("parent_link_name"), py::arg("relative_positions"), py::arg("radii"),
           py::arg("only_self_collision") = false)
      .def_readonly("parent_link_name", &SphereAttachmentSpec::parent_link_name)
      .def_property_readonly("radii", [](const SphereAttachmentSpec& s) { return s.radii; })
      .def_property_readonly("relative_positions",
                             [](const SphereAttachmentSpec& s) {  // (3, k), as the reference exposes it
                               const py::ssize_t k = (py::ssize_t)s.radii.size();
                               DArr a({(py::ssize_t)3, k});
                               for (py::ssize_t j = 0; j < k; ++j)
                                 for (int c = 0; c < 3; ++c) a.mutable_data()[c * k + j] = s.relative_positions[3 * j + c];
                               return a;
                             })
      .def_readonly("only_self_collision", &SphereAttachmentSpec::only_self_collision);

  py::class_<SphereCollisionCst, std::shared_ptr<SphereCollisionCst>>(mc, "SphereCollisionCst")
      .def(py::init([](std::shared_ptr<KinematicModel> kin, const std::vector<std::string>& names, BaseType bt,
                       const std::vector<SphereAttachmentSpec>& specs,
                       const std::vector<std::pair<std::string, std::string>>& pairs,
                       std::optional<SDFBase::Ptr> fixed_sdf, bool reorder) {
             return std::make_shared<SphereCollisionCst>(kin, names, (int)bt, specs, pairs, fixed_sdf, reorder);
           }),
           py::arg("kin"), py::arg("control_joint_names"), py::arg("base_type"), py::arg("sphere_specs"),
           py::arg("selcol_pairs"), py::arg("fixed_sdf") = py::none(), py::arg("reorder_spheres") = true)
      .def("set_sdf", &SphereCollisionCst::set_sdf)
      .def("get_sdf", &SphereCollisionCst::get_sdf)
      .def("get_kin", &SphereCollisionCst::get_kin)
      .def("get_control_joint_names", &SphereCollisionCst::get_control_joint_names)
      .def("cst_dim", &SphereCollisionCst::cst_dim)
      .def("q_dim", &SphereCollisionCst::q_dim)
      .def("n_spheres", &SphereCollisionCst::n_spheres)
      .def("n_prims", &SphereCollisionCst::n_prims)
      .def("update_kintree",
           [](SphereCollisionCst& c, const DArr& q, bool acc) { c.update_kintree(vecd(q), acc); }, py::arg("q"),
           py::arg("high_accuracy") = true)
      .def("is_valid",
           [](SphereCollisionCst& c, const DArr& q) {
             auto v = vecd(q);
             py::gil_scoped_release rel;
             return c.is_valid(v);
           })
      .def("evaluate",
           [](SphereCollisionCst& c, const DArr& q) {
             auto qv = vecd(q);
             std::pair<std::vector<double>, std::vector<double>> r;
             {
               py::gil_scoped_release rel;
               r = c.evaluate(qv);
             }
             const int64_t C = (int64_t)c.cst_dim(), D = (int64_t)c.q_dim();
             DArr vals(C), jac({C, D});
             std::copy(r.first.begin(), r.first.end(), vals.mutable_data());
             std::copy(r.second.begin(), r.second.end(), jac.mutable_data());
             return std::make_pair(vals, jac);
           })
      .def("get_all_spheres",
           [](SphereCollisionCst& c) {
             py::list out;
             for (auto& s : c.get_all_spheres()) out.append(py::make_tuple(arr3(s.first), s.second));
             return out;
           })
      .def("get_group_spheres",
           [](SphereCollisionCst& c) {
             py::list out;
             for (auto& s : c.get_group_spheres()) out.append(py::make_tuple(arr3(s.first), s.second));
             return out;
           })
      // ---- batched entry points (host arrays)
      .def("is_valid_batch",
           [](SphereCollisionCst& c, const DArr& Q) {
             int64_t B;
             const double* q = rows(Q, c.q_dim(), B);
             int32_t n = 0;
             std::vector<uint8_t> v;
             {
               py::gil_scoped_release rel;
               v = c.is_valid_batch(q, B, &n);
             }
             py::array_t<bool> out(B);
             for (int64_t i = 0; i < B; ++i) out.mutable_data()[i] = v[i] != 0;
             return out;
           })
      .def("is_valid_batch_stats",
           [](SphereCollisionCst& c, const DArr& Q) {
             int64_t B;
             const double* q = rows(Q, c.q_dim(), B);
             int32_t n = 0;
             std::vector<uint8_t> v;
             {
               py::gil_scoped_release rel;
               v = c.is_valid_batch(q, B, &n);
             }
             py::array_t<bool> out(B);
             for (int64_t i = 0; i < B; ++i) out.mutable_data()[i] = v[i] != 0;
             return std::make_pair(out, n);
           })
      .def(
          "traj_cost_grad_batch",
          [](SphereCollisionCst& c, const DArr& X, const std::vector<int64_t>& offsets,
             const std::vector<double>& weights) {
            int64_t N;
            const double* x = rows(X, c.q_dim(), N);
            std::vector<double> vals, jac, cost, grad;
            {
              py::gil_scoped_release rel;
              c.traj_cost_grad_batch(x, N, offsets, weights, vals, jac, cost, grad);
            }
            const int64_t C = (int64_t)c.cst_dim(), D = (int64_t)c.q_dim(), T = (int64_t)cost.size();
            DArr v({N, C}), j({N, C, D}), co(T), g({N, D});
            std::copy(vals.begin(), vals.end(), v.mutable_data());
            std::copy(jac.begin(), jac.end(), j.mutable_data());
            std::copy(cost.begin(), cost.end(), co.mutable_data());
            std::copy(grad.begin(), grad.end(), g.mutable_data());
            return py::make_tuple(v, j, co, g);
          },
          py::arg("X"), py::arg("offsets"), py::arg("weights") = std::vector<double>{},
          "Per-waypoint evaluate rows, smoothness cost per trajectory and its gradient, one launch.")
      .def(
          "traj_sqp_ineq",
          [](SphereCollisionCst& c, const DArr& X, const std::vector<int64_t>& offsets,
             const std::vector<double>& msbox) {
            int64_t N;
            const double* x = rows(X, c.q_dim(), N);
            SphereCollisionCst::SqpPieces p;
            {
              py::gil_scoped_release rel;
              p = c.traj_sqp_ineq(x, N, offsets, msbox);
            }
            auto arr = [](const auto& v) {
              using T = typename std::decay_t<decltype(v)>::value_type;
              py::array_t<T> a((py::ssize_t)v.size());
              std::copy(v.begin(), v.end(), a.mutable_data());
              return a;
            };
            return py::make_tuple(arr(p.c), arr(p.colptr), arr(p.rowind), arr(p.values), arr(p.row_off),
                                  arr(p.nnz_off), arr(p.cp_off));
          },
          py::arg("X"), py::arg("offsets"), py::arg("msbox") = std::vector<double>{},
          "SQP inequality sequence of every trajectory: (c, colptr, rowind, values, row_off, nnz_off, colptr_off).")
      .def("traj_sqp_sizes",
           [](SphereCollisionCst& c, const std::vector<int64_t>& offsets, bool with_msbox) {
             const int32_t n = (int32_t)offsets.size() - 1;
             if (n < 0) throw std::invalid_argument("offsets must hold n_traj + 1 entries");
             std::vector<int64_t> r(n + 1), z(n + 1), p(n + 1);
             check(pmp_traj_sqp_sizes(c.device_context(nullptr), offsets.data(), n, with_msbox ? 1 : 0, r.data(),
                                      z.data(), p.data()));
             return py::make_tuple(r, z, p);
           })
      .def("_dev_traj_sqp_ineq",
           [](SphereCollisionCst& c, uintptr_t X, int64_t ldx, int64_t N, const std::vector<int64_t>& offsets,
              const std::vector<double>& msbox, uintptr_t cv, uintptr_t colptr, uintptr_t rowind, uintptr_t values,
              uintptr_t stream) {
             pmp_context* ctx = c.device_context((void*)stream);
             check(pmp_traj_sqp_ineq(ctx, (const double*)X, ldx, N, offsets.data(), (int32_t)offsets.size() - 1,
                                     msbox.empty() ? nullptr : msbox.data(), (double*)cv, (int32_t*)colptr,
                                     (int32_t*)rowind, (double*)values));
           })
      .def("_dev_traj_cost_grad",
           [](SphereCollisionCst& c, uintptr_t X, int64_t ldx, int64_t N, const std::vector<int64_t>& offsets,
              const std::vector<double>& weights, uintptr_t vals, uintptr_t jac, uintptr_t cost, uintptr_t grad,
              uintptr_t stream) {
             pmp_context* ctx = c.device_context((void*)stream);
             check(pmp_traj_cost_grad_batch(ctx, (const double*)X, ldx, N, offsets.data(), (int32_t)offsets.size() - 1,
                                            weights.empty() ? nullptr : weights.data(), (float*)vals, (float*)jac,
                                            (float*)cost, (float*)grad));
           })
      .def("evaluate_batch",
           [](SphereCollisionCst& c, const DArr& Q) {
             int64_t B;
             const double* q = rows(Q, c.q_dim(), B);
             std::vector<double> vals, jac;
             {
               py::gil_scoped_release rel;
               c.evaluate_batch(q, B, vals, jac);
             }
             const int64_t C = (int64_t)c.cst_dim(), D = (int64_t)c.q_dim();
             DArr v({B, C}), j({B, C, D});
             std::copy(vals.begin(), vals.end(), v.mutable_data());
             std::copy(jac.begin(), jac.end(), j.mutable_data());
             return std::make_pair(v, j);
           })
      .def("sphere_clearance_batch",
           [](SphereCollisionCst& c, const DArr& Q, bool accurate) {
             int64_t B;
             const double* q = rows(Q, c.q_dim(), B);
             std::vector<float> v;
             {
               py::gil_scoped_release rel;
               v = c.sphere_clearance_batch(q, B, accurate);
             }
             py::array_t<float> out({B, (int64_t)c.n_spheres()});
             std::copy(v.begin(), v.end(), out.mutable_data());
             return out;
           },
           py::arg("Q"), py::arg("accurate") = false)
      .def("sphere_centers_batch",
           [](SphereCollisionCst& c, const DArr& Q, bool accurate) {
             int64_t B;
             const double* q = rows(Q, c.q_dim(), B);
             std::vector<float> v;
             {
               py::gil_scoped_release rel;
               v = c.sphere_centers_batch(q, B, accurate);
             }
             py::array_t<float> out({B, (int64_t)c.n_spheres(), (int64_t)3});
             std::copy(v.begin(), v.end(), out.mutable_data());
             return out;
           },
           py::arg("Q"), py::arg("accurate") = false)
      .def("pair_clearance_batch",
           [](SphereCollisionCst& c, const DArr& Q, bool with_grad) {
             int64_t B;
             const double* q = rows(Q, c.q_dim(), B);
             std::vector<float> d, g;
             {
               py::gil_scoped_release rel;
               c.pair_clearance_batch(q, B, with_grad, d, g);
             }
             const int64_t S = c.n_spheres(), P = c.n_prims(), D = c.q_dim();
             py::array_t<float> dist({B, S, P});
             std::copy(d.begin(), d.end(), dist.mutable_data());
             py::object grad = py::none();
             if (with_grad) {
               py::array_t<float> ga({B, S, P, D});
               std::copy(g.begin(), g.end(), ga.mutable_data());
               grad = ga;
             }
             return py::make_tuple(dist, grad);
           },
           py::arg("Q"), py::arg("with_grad") = true)
      // ---- device-pointer entry points (torch tensors; pointers as integers)
      .def("_dev_pair_clearance",
           [](SphereCollisionCst& c, uintptr_t q, int64_t ldq, int64_t B, uintptr_t dist, uintptr_t grad,
              uintptr_t stream) {
             check(pmp_pair_clearance_batch(c.device_context((void*)stream), (const double*)q, ldq, B, (float*)dist,
                                            (float*)grad));
           })
      .def("_dev_pair_clearance_vjp",
           [](SphereCollisionCst& c, uintptr_t q, int64_t ldq, int64_t B, uintptr_t gd, uintptr_t gq,
              uintptr_t stream) {
             check(pmp_pair_clearance_vjp(c.device_context((void*)stream), (const double*)q, ldq, B,
                                          (const float*)gd, (float*)gq));
           })
      .def("_dev_is_valid",
           [](SphereCollisionCst& c, uintptr_t q, int64_t ldq, int64_t B, uintptr_t valid, uintptr_t nre,
              uintptr_t stream) {
             check(pmp_is_valid_batch(c.device_context((void*)stream), (const double*)q, ldq, B, (uint8_t*)valid,
                                      (int32_t*)nre));
           })
      .def("_dev_is_valid_noalloc",
           [](SphereCollisionCst& c, uintptr_t q, int64_t ldq, int64_t B, uintptr_t valid, uintptr_t nre,
              uintptr_t stream) {
             check(pmp_is_valid_batch_noalloc(c.device_context((void*)stream), (const double*)q, ldq, B,
                                              (uint8_t*)valid, (int32_t*)nre));
           })
      .def("_dev_sphere_clearance",
           [](SphereCollisionCst& c, uintptr_t q, int64_t ldq, int64_t B, int accurate, uintptr_t out,
              uintptr_t stream) {
             check(pmp_sphere_clearance_batch(c.device_context((void*)stream), (const double*)q, ldq, B, accurate,
                                              (float*)out));
           })
      .def(
          "link_jacobian_batch",
          [](SphereCollisionCst& c, const DArr& Q, const std::string& link, const std::vector<double>& point) {
            int64_t B;
            const double* q = rows(Q, c.q_dim(), B);
            const int32_t id = (int32_t)c.get_kin()->get_link_ids({link})[0];
            std::vector<float> j;
            {
              py::gil_scoped_release rel;
              j = c.link_jacobian_batch(q, B, id, point);
            }
            py::array_t<float> a({(py::ssize_t)B, (py::ssize_t)3, (py::ssize_t)c.q_dim()});
            std::copy(j.begin(), j.end(), a.mutable_data());
            return a;
          },
          py::arg("Q"), py::arg("link_name"), py::arg("point") = std::vector<double>{},
          "(B, 3, q_dim) Jacobian of a point fixed to a link, rotation ignored (exact FK).")
      .def("_dev_link_poses",
           [](SphereCollisionCst& c, uintptr_t q, int64_t ldq, int64_t B, int accurate, uintptr_t out,
              uintptr_t stream) {
             check(pmp_link_poses_batch(c.device_context((void*)stream), (const double*)q, ldq, B, accurate,
                                        (double*)out));
           })
      .def("_dev_link_jacobian",
           [](SphereCollisionCst& c, uintptr_t q, int64_t ldq, int64_t B, int link, const std::vector<double>& point,
              uintptr_t jac, uintptr_t stream) {
             if (!point.empty() && point.size() != 3) throw std::invalid_argument("point must have 3 entries");
             check(pmp_link_jacobian_batch(c.device_context((void*)stream), (const double*)q, ldq, B, link,
                                           point.empty() ? nullptr : point.data(), (float*)jac));
           })
      .def("_dev_ik_step",
           [](SphereCollisionCst& c, uintptr_t q, int64_t R, uintptr_t poses, int kind, int link, uintptr_t target,
              uintptr_t jac0, const std::vector<uintptr_t>& jac_u, uintptr_t vals, uintptr_t cjac, int n_cst,
              uintptr_t lo, uintptr_t hi, double damping, double max_step, double margin, double penalty,
              uintptr_t stream) {
             pmp_ik_step_desc d{};
             d.kind = kind;
             d.link = link;
             d.target = (const double*)target;
             d.jac0 = (const float*)jac0;
             for (size_t k = 0; k < 3 && k < jac_u.size(); ++k) d.jac_u[k] = (const float*)jac_u[k];
             d.vals = (const float*)vals;
             d.cjac = (const float*)cjac;
             d.n_cst = n_cst;
             d.lo = (const double*)lo;
             d.hi = (const double*)hi;
             d.damping = damping;
             d.max_step = max_step;
             d.margin = margin;
             d.penalty = penalty;
             check(pmp_ik_step_batch(c.device_context((void*)stream), (double*)q, R, (const double*)poses, &d));
           })
      .def("_dev_evaluate",
           [](SphereCollisionCst& c, uintptr_t q, int64_t ldq, int64_t B, uintptr_t vals, uintptr_t jac,
              uintptr_t stream) {
             check(pmp_evaluate_batch(c.device_context((void*)stream), (const double*)q, ldq, B, (float*)vals,
                                      (float*)jac));
           })
      .def("set_band",
           [](SphereCollisionCst& c, double band) { check(pmp_set_band(c.device_context(nullptr), band)); },
           "Width (m) of the band around zero clearance whose FP32 decisions are re-made in FP64 (default 1e-5).")
      .def(
          "_dev_set_schedule",
          [](SphereCollisionCst& c, int block, int waves_per_cu, int whole_chunks, int pair_shares) {
            pmp_schedule s{block, waves_per_cu, whole_chunks, pair_shares};
            check(pmp_set_schedule(c.device_context(nullptr), &s));
          },
          py::arg("block"), py::arg("waves_per_cu"), py::arg("whole_chunks"), py::arg("pair_shares") = 0);

  // -------------------------------------------------------------------- ompl
  auto mo = m.def_submodule("ompl");
  py::enum_<ValidatorConfig::Type>(mo, "ValidatorType")
      .value("BOX", ValidatorConfig::Type::BOX)
      .value("EUCLIDEAN", ValidatorConfig::Type::EUCLIDEAN)
      .export_values();
  py::class_<ValidatorConfig>(mo, "ValidatorConfig")
      .def(py::init<>())
      .def_readwrite("type", &ValidatorConfig::type)
      .def_readwrite("resolution", &ValidatorConfig::resolution)
      .def_readwrite("box_width", &ValidatorConfig::box_width);
  // A captured edge batch.  It keeps its constraint (and so the device
  // context) alive; launch() first brings the context's tables up to date
  // with the constraint (a set_sdf or a moved primitive re-uploads them), and
  // a graph captured before such a change then refuses to replay.
  struct Graph {
    pmp_graph* g = nullptr;
    std::shared_ptr<SphereCollisionCst> cst;
    ~Graph() {
      if (g) pmp_graph_destroy(g);
    }
  };
  py::class_<Graph, std::shared_ptr<Graph>>(mo, "EdgeGraph")
      .def("launch", [](Graph& g) {
        if (!g.g) throw std::invalid_argument("graph destroyed");
        g.cst->sync_batch_tables();
        check(pmp_graph_launch(g.g));
      })
      .def("destroy", [](Graph& g) {
        if (g.g) check(pmp_graph_destroy(g.g));
        g.g = nullptr;
      });
  // trajectories of uneven length (cost, gradient and evaluate rows of every
  // waypoint) captured once and replayed by launch()
  mo.def("_dev_capture_traj_graph",
         [](std::shared_ptr<SphereCollisionCst> c, uintptr_t X, int64_t ldx, int64_t N,
            const std::vector<int64_t>& offsets, const std::vector<double>& weights, uintptr_t vals, uintptr_t jac,
            uintptr_t cost, uintptr_t grad, uintptr_t stream) {
           pmp_context* ctx = c->device_context((void*)stream);
           auto g = std::make_shared<Graph>();
           g->cst = c;
           check(pmp_traj_cost_grad_graph(ctx, (const double*)X, ldx, N, offsets.data(), (int32_t)offsets.size() - 1,
                                          weights.empty() ? nullptr : weights.data(), (float*)vals, (float*)jac,
                                          (float*)cost, (float*)grad, &g->g));
           return g;
         });
  // n_iter iterations of the batched IK (poses, link Jacobians, collision
  // rows, step) captured once and replayed by launch()
  mo.def("_dev_capture_ik_graph",
         [](std::shared_ptr<SphereCollisionCst> c, uintptr_t q, int64_t R, int kind, int link,
            uintptr_t target, uintptr_t jac0, const std::vector<uintptr_t>& jac_u, uintptr_t vals, uintptr_t cjac,
            int n_cst, uintptr_t lo, uintptr_t hi, double damping, double max_step, double margin, double penalty,
            int n_iter, uintptr_t stream) {
           pmp_context* ctx = c->device_context((void*)stream);
           pmp_ik_step_desc d{};
           d.kind = kind;
           d.link = link;
           d.target = (const double*)target;
           d.jac0 = (const float*)jac0;
           for (size_t k = 0; k < 3 && k < jac_u.size(); ++k) d.jac_u[k] = (const float*)jac_u[k];
           d.vals = (const float*)vals;
           d.cjac = (const float*)cjac;
           d.n_cst = n_cst;
           d.lo = (const double*)lo;
           d.hi = (const double*)hi;
           d.damping = damping;
           d.max_step = max_step;
           d.margin = margin;
           d.penalty = penalty;
           auto g = std::make_shared<Graph>();
           g->cst = c;
           check(pmp_ik_graph(ctx, (double*)q, R, &d, n_iter, &g->g));
           return g;
         });
  py::class_<MotionValidator, std::shared_ptr<MotionValidator>>(mo, "MotionValidator")
      .def(py::init<std::shared_ptr<SphereCollisionCst>, const ValidatorConfig&>())
      .def("checkMotion", [](MotionValidator& v, const DArr& s1, const DArr& s2) {
        auto a = vecd(s1), b = vecd(s2);
        py::gil_scoped_release rel;
        return v.checkMotion(a, b);
      })
      .def("check_motions",
           [](MotionValidator& v, const DArr& S1, const DArr& S2) {
             int64_t E, E2;
             const double* a = rows(S1, v.constraint()->q_dim(), E);
             const double* b = rows(S2, v.constraint()->q_dim(), E2);
             if (E != E2) throw std::invalid_argument("S1 and S2 differ in length");
             std::vector<int32_t> n;
             std::vector<uint8_t> ok;
             {
               py::gil_scoped_release rel;
               ok = v.check_motions(a, b, E, &n);
             }
             py::array_t<bool> out(E);
             py::array_t<int32_t> nc(E);
             for (int64_t i = 0; i < E; ++i) {
               out.mutable_data()[i] = ok[i] != 0;
               nc.mutable_data()[i] = n[i];
             }
             return std::make_pair(out, nc);
           })
      .def("check_path",
           [](MotionValidator& v, const DArr& P) {
             int64_t n;
             const double* p = rows(P, v.constraint()->q_dim(), n);
             std::vector<uint8_t> ok;
             {
               py::gil_scoped_release rel;
               ok = v.check_path(p, n);
             }
             py::array_t<bool> out((py::ssize_t)ok.size());
             for (size_t i = 0; i < ok.size(); ++i) out.mutable_data()[i] = ok[i] != 0;
             return out;
           })
      .def(
          "shortcut_path",
          [](MotionValidator& v, const DArr& P, int candidates, int rounds, uint64_t seed) {
            const int64_t D = (int64_t)v.constraint()->q_dim();
            int64_t n, m = 0;
            const double* p = rows(P, D, n);
            std::vector<double> out;
            {
              py::gil_scoped_release rel;
              out = v.shortcut_path(p, n, candidates, rounds, seed, &m);
            }
            py::array_t<double> a({(py::ssize_t)m, (py::ssize_t)D});
            std::copy(out.begin(), out.end(), a.mutable_data());
            return a;
          },
          py::arg("path"), py::arg("candidates") = 256, py::arg("rounds") = 8, py::arg("seed") = 0)
      .def("_dev_check_motions",
           [](MotionValidator& v, uintptr_t s1, uintptr_t s2, int64_t lds, int64_t E, uintptr_t valid,
              uintptr_t nchk, uintptr_t stream) {
             pmp_context* ctx = v.constraint()->device_context((void*)stream);
             pmp_validator_desc d = v.desc();
             check(pmp_check_edges_batch(ctx, (const double*)s1, (const double*)s2, lds, E, &d, (uint8_t*)valid,
                                         (int32_t*)nchk));
           })
      .def("last_recheck_count", [](MotionValidator& v) {
        py::gil_scoped_release rel;
        return v.last_recheck_count();
      })
      .def("_dev_edge_kernel_ms",
           [](MotionValidator& v, uintptr_t stream) {
             pmp_context* ctx = v.constraint()->device_context((void*)stream);
             float ms = 0.f;
             {
               py::gil_scoped_release rel;
               check(pmp_edge_kernel_ms(ctx, &ms));
             }
             return ms;
           })
      .def("_dev_capture_graph",
           [](MotionValidator& v, uintptr_t s1, uintptr_t s2, int64_t lds, int64_t E, uintptr_t valid, uintptr_t nchk,
              uintptr_t stream) {
             pmp_context* ctx = v.constraint()->device_context((void*)stream);
             pmp_validator_desc d = v.desc();
             auto g = std::make_shared<Graph>();
             g->cst = v.constraint();
             check(pmp_check_edges_graph(ctx, (const double*)s1, (const double*)s2, lds, E, &d, (uint8_t*)valid,
                                         (int32_t*)nchk, &g->g));
             return g;
           });
}
