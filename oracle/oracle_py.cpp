// Python bindings of the FP64 oracle (TEST INFRASTRUCTURE ONLY): the classes
// of oracle.hpp plus thread-parallel batch loops.  Batch loops give every
// worker thread its own copy of the constraint, because the reference objects
// keep mutable caches and are not re-entrant (kinematics.hpp:89-91).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <thread>

#include "oracle.hpp"

namespace py = pybind11;
using namespace oracle;

namespace {

using Arr = py::array_t<double, py::array::c_style | py::array::forcecast>;

V3 v3(const Arr& a) {
  if (a.size() != 3) throw std::invalid_argument("expected 3 values");
  auto r = a.unchecked();
  const double* d = a.data();
  (void)r;
  return {d[0], d[1], d[2]};
}
M3 m3(const Arr& a) {
  if (a.size() != 9) throw std::invalid_argument("expected a 3x3 matrix");
  const double* d = a.data();
  M3 R;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) R.m[i][j] = d[i * 3 + j];
  return R;
}
Arr to_arr(const V3& v) {
  Arr a(3);
  auto m = a.mutable_unchecked<1>();
  m(0) = v.x;
  m(1) = v.y;
  m(2) = v.z;
  return a;
}
Arr to_arr(const M3& R) {
  Arr a({3, 3});
  auto m = a.mutable_unchecked<2>();
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m(i, j) = R.m[i][j];
  return a;
}
std::vector<double> row(const double* p, size_t n) { return std::vector<double>(p, p + n); }

template <class F>
void parallel_for(size_t n, int n_threads, F&& f) {
  if (n_threads <= 1 || n < 2) {
    f(0, n, 0);
    return;
  }
  std::vector<std::thread> th;
  size_t chunk = (n + n_threads - 1) / n_threads;
  for (int t = 0; t < n_threads; t++) {
    size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([&, b, e, t] { f(b, e, t); });
  }
  for (auto& x : th) x.join();
}

Arr take_rows(const Arr& Q, size_t& B, size_t& D) {
  if (Q.ndim() != 2) throw std::invalid_argument("expected a (B, dof) array");
  B = Q.shape(0);
  D = Q.shape(1);
  return Q;
}

// Per-sphere minimum signed clearance over all primitives, in group order,
// using FK in the given accuracy mode.  (Composition of get_all_spheres,
// primitive_sphere_collision.cpp:412-424, and PrimitiveSDF evaluate.)
void sphere_clearances(SphereCollisionCst& c, const std::vector<double>& q, bool accurate, double* out) {
  c.update_kintree(q, accurate);
  auto sp = c.get_all_spheres();
  for (size_t s = 0; s < sp.size(); s++) {
    double best = std::numeric_limits<double>::infinity();
    for (auto& p : c.all_sdfs_) best = std::min(best, p.evaluate(sp[s].first) - sp[s].second);
    out[s] = best;
  }
}

}  // namespace

PYBIND11_MODULE(_oracle, m) {
  m.doc() = "FP64 CPU restatement of plainmp's FK + sphere-SDF path (test oracle)";

  m.def("taylor_sincos", [](double x) {
    double s, c;
    taylor_sincos(x, s, c);
    return std::make_pair(s, c);
  });
  m.def("urdf_quat_from_rpy", [](double r, double p, double y) {
    Q4 q = urdf_quat_from_rpy(r, p, y);
    return std::vector<double>{q.x, q.y, q.z, q.w};
  });
  m.def("quat_from_rpy", [](double r, double p, double y) {
    Q4 q = quat_from_rpy(r, p, y);
    return std::vector<double>{q.x, q.y, q.z, q.w};
  });
  m.def("quat_to_rpy", [](double x, double y, double z, double w) {
    Q4 q;
    q.x = x;
    q.y = y;
    q.z = z;
    q.w = w;
    V3 r = quat_to_rpy(q);
    return std::vector<double>{r.x, r.y, r.z};
  });
  m.def("informative_order", [](const Arr& P) {  // (k, 3) local sphere positions -> visiting order
    if (P.ndim() != 2 || P.shape(1) != 3) throw std::invalid_argument("expected (k, 3) positions");
    std::vector<V3> pos;
    for (py::ssize_t i = 0; i < P.shape(0); i++) pos.push_back({P.at(i, 0), P.at(i, 1), P.at(i, 2)});
    return informative_order(pos);
  });
  m.def("farthest_sequence", &farthest_sequence);
  m.def("sequence_table", []() { return sequence_table(); });
  m.def("box_step_ratio", &box_step_ratio);
  m.def("euclid_step_ratio", &euclid_step_ratio);
  m.def("motion_travel_rates", &motion_travel_rates);
  // SequentialCst::evaluate of the SQP inequality sequence for one trajectory
  // x (T, dof): (c, colptr, rowind, values, (rows, cols))
  m.def(
      "seq_ineq_evaluate",
      [](SphereCollisionCst& cst, const Arr& X, std::optional<std::vector<double>> msbox) {
        if (X.ndim() != 2 || (size_t)X.shape(1) != cst.q_dim()) throw std::invalid_argument("expected (T, dof) waypoints");
        const size_t T = X.shape(0);
        std::vector<double> x(X.data(), X.data() + X.size());
        SeqCsc r = seq_ineq_evaluate(cst, x, T, msbox ? &*msbox : nullptr);
        return py::make_tuple(Arr(r.c.size(), r.c.data()), py::array_t<int32_t>(r.colptr.size(), r.colptr.data()),
                              py::array_t<int32_t>(r.rowind.size(), r.rowind.data()),
                              Arr(r.values.size(), r.values.data()), py::make_tuple(r.rows, r.cols));
      },
      py::arg("cst"), py::arg("X"), py::arg("msbox") = py::none());
  m.def("smooth_cost", [](const Arr& x, size_t T, size_t dof, const std::vector<double>& w) {
    std::vector<double> g;
    double c = smooth_cost(row(x.data(), x.size()), T, dof, w, &g);
    return std::make_pair(c, g);
  });

  py::enum_<BaseType>(m, "BaseType")
      .value("FIXED", BaseType::FIXED)
      .value("PLANAR", BaseType::PLANAR)
      .value("FLOATING", BaseType::FLOATING);

  py::class_<KinematicModel>(m, "KinematicModel")
      .def(py::init([](const std::vector<std::string>& links, const py::list& joints) {
        std::vector<UrdfJoint> js;
        for (auto item : joints) {
          auto t = item.cast<py::tuple>();
          UrdfJoint j;
          j.name = t[0].cast<std::string>();
          j.type = t[1].cast<int>();
          j.parent = t[2].cast<std::string>();
          j.child = t[3].cast<std::string>();
          auto xyz = t[4].cast<std::vector<double>>();
          auto qv = t[5].cast<std::vector<double>>();
          auto ax = t[6].cast<std::vector<double>>();
          j.origin_xyz = {xyz[0], xyz[1], xyz[2]};
          j.origin_quat = {qv[0], qv[1], qv[2], qv[3]};
          j.axis = {ax[0], ax[1], ax[2]};
          j.lower = t[7].cast<double>();
          j.upper = t[8].cast<double>();
          js.push_back(j);
        }
        return KinematicModel(links, js);
      }))
      .def("get_joint_ids", &KinematicModel::get_joint_ids)
      .def("get_link_ids", &KinematicModel::get_link_ids)
      .def("get_joint_position_limits", &KinematicModel::get_joint_position_limits)
      .def("set_joint_positions",
           [](KinematicModel& k, const std::vector<size_t>& ids, const std::vector<double>& q, bool acc) {
             k.set_joint_angles(ids, q, acc);
           },
           py::arg("joint_ids"), py::arg("positions"), py::arg("accurate") = true)
      .def("get_joint_positions", &KinematicModel::get_joint_angles)
      .def("get_link_pose",
           [](KinematicModel& k, size_t id) {
             const QT& t = k.get_link_pose(id);
             return std::vector<double>{t.trans.x, t.trans.y, t.trans.z, t.quat.x, t.quat.y, t.quat.z, t.quat.w};
           })
      .def("set_base_pose",
           [](KinematicModel& k, const std::vector<double>& v) {
             QT p;
             p.trans = {v[0], v[1], v[2]};
             p.quat = {v[3], v[4], v[5], v[6]};
             k.set_base_pose(p);
           })
      .def("get_attached_point_jacobian",
           [](KinematicModel& k, size_t link, const Arr& p, const std::vector<size_t>& ids, BaseType bt) {
             auto J = k.get_attached_point_jacobian(link, v3(p), ids, bt);
             Arr out({(size_t)3, J.size()});
             auto mo = out.mutable_unchecked<2>();
             for (size_t c = 0; c < J.size(); c++)
               for (int r = 0; r < 3; r++) mo(r, c) = J[c][r];
             return out;
           })
      .def("add_new_link",
           [](KinematicModel& k, const std::string& name, const std::string& parent, const Arr& pos,
              const Arr& rpy, bool consider_rotation) {
             size_t pid = k.get_link_ids({parent})[0];
             return k.add_new_link(pid, v3(pos), v3(rpy), consider_rotation, name);
           })
      .def_property_readonly("n_links", &KinematicModel::n_links)
      .def_property_readonly("n_joints", &KinematicModel::n_joints);

  py::class_<Prim>(m, "Prim")
      .def_static("box", [](const Arr& w, const Arr& p, const Arr& R) { return Prim::box(v3(w), Pose(v3(p), m3(R))); })
      .def_static("cylinder", [](double r, double h, const Arr& p, const Arr& R) {
        return Prim::cylinder(r, h, Pose(v3(p), m3(R)));
      })
      .def_static("sphere", [](double r, const Arr& p, const Arr& R) { return Prim::sphere(r, Pose(v3(p), m3(R))); })
      .def_static("ground", &Prim::ground)
      .def_static("cloud", [](const Arr& pts, double r) {
        if (pts.ndim() != 2 || pts.shape(1) != 3) throw std::invalid_argument("expected (N, 3) points");
        std::vector<V3> v;
        auto a = pts.unchecked<2>();
        for (py::ssize_t i = 0; i < pts.shape(0); i++) v.push_back({a(i, 0), a(i, 1), a(i, 2)});
        return Prim::cloud(v, r);
      })
      .def_readonly("type", &Prim::type)
      .def("evaluate", [](const Prim& s, const Arr& p) { return s.evaluate(v3(p)); })
      .def("evaluate_batch",
           [](const Prim& s, const Arr& P) {
             if (P.ndim() != 2 || P.shape(0) != 3) throw std::invalid_argument("expected (3, N) points");
             auto a = P.unchecked<2>();
             Arr out(P.shape(1));
             auto o = out.mutable_unchecked<1>();
             for (py::ssize_t i = 0; i < P.shape(1); i++) o(i) = s.evaluate({a(0, i), a(1, i), a(2, i)});
             return out;
           })
      .def("is_outside", [](const Prim& s, const Arr& p, double r) { return s.is_outside(v3(p), r); })
      .def("is_outside_aabb", [](const Prim& s, const Arr& p, double r) { return s.is_outside_aabb(v3(p), r); })
      .def("translate", [](Prim& s, const Arr& t) { s.translate(v3(t)); })
      .def("rotate_z", &Prim::rotate_z)
      .def_property_readonly("lb", [](const Prim& s) { return to_arr(s.lb); })
      .def_property_readonly("ub", [](const Prim& s) { return to_arr(s.ub); })
      .def_property_readonly("position", [](const Prim& s) { return to_arr(s.pose.position); })
      .def_property_readonly("rotation", [](const Prim& s) { return to_arr(s.pose.rot); })
      .def_property_readonly("axis_aligned", [](const Prim& s) { return s.pose.axis_aligned; })
      .def_property_readonly("z_axis_aligned", [](const Prim& s) { return s.pose.z_axis_aligned; });

  py::class_<KDTree>(m, "KDTree")
      .def(py::init([](const Arr& pts) {
        std::vector<V3> v;
        auto a = pts.unchecked<2>();
        for (py::ssize_t i = 0; i < pts.shape(0); i++) v.push_back({a(i, 0), a(i, 1), a(i, 2)});
        return KDTree(v);
      }))
      .def("query", [](const KDTree& t, const Arr& p) { return to_arr(t.query(v3(p))); })
      .def("sqdist", [](const KDTree& t, const Arr& p) { return t.sqdist(v3(p)); });

  m.def("cloud_tiles", [](const Arr& pts) {  // (N, 3) -> dict of the tables as bytes
    if (pts.ndim() != 2 || pts.shape(1) != 3) throw std::invalid_argument("expected (N, 3) points");
    std::vector<double> v(pts.data(), pts.data() + pts.size());
    const CloudTilesRef r = cloud_tiles(v);
    auto raw = [](const void* p, size_t n) { return py::bytes((const char*)p, n); };
    py::dict d;
    d["tiles"] = raw(r.t32.data(), r.t32.size() * sizeof(CloudTilesRef::T32));
    d["tiles64"] = raw(r.t64.data(), r.t64.size() * sizeof(CloudTilesRef::T64));
    d["tile_b"] = raw(r.tb.data(), r.tb.size() * sizeof(float));
    d["points64"] = raw(r.p64.data(), r.p64.size() * sizeof(double));
    d["supers"] = raw(r.sup.data(), r.sup.size() * sizeof(CloudTilesRef::Sup));
    return d;
  });

  py::class_<SphereSpec>(m, "SphereSpec")
      .def(py::init([](const std::string& link, const Arr& rel, const std::vector<double>& radii, bool only_self) {
        if (rel.ndim() != 2 || rel.shape(1) != 3) throw std::invalid_argument("expected (k, 3) positions");
        SphereSpec s;
        s.parent_link_name = link;
        auto a = rel.unchecked<2>();
        for (py::ssize_t i = 0; i < rel.shape(0); i++) s.relative_positions.push_back({a(i, 0), a(i, 1), a(i, 2)});
        s.radii = radii;
        s.only_self_collision = only_self;
        return s;
      }));

  py::class_<SphereCollisionCst>(m, "SphereCollisionCst")
      .def(py::init([](const KinematicModel& kin, const std::vector<std::string>& names, BaseType bt,
                       const std::vector<SphereSpec>& specs,
                       const std::vector<std::pair<std::string, std::string>>& pairs,
                       const std::vector<Prim>& fixed_prims, bool reorder) {
        return SphereCollisionCst(kin, names, bt, specs, pairs, fixed_prims, reorder);
      }))
      .def("set_sdf", &SphereCollisionCst::set_sdf)
      .def("q_dim", &SphereCollisionCst::q_dim)
      .def("cst_dim", &SphereCollisionCst::cst_dim)
      .def("is_valid", [](SphereCollisionCst& c, const Arr& q) { return c.is_valid(row(q.data(), q.size())); })
      .def("update_kintree",
           [](SphereCollisionCst& c, const Arr& q, bool acc) { c.update_kintree(row(q.data(), q.size()), acc); })
      .def("evaluate",
           [](SphereCollisionCst& c, const Arr& q) {
             auto r = c.evaluate(row(q.data(), q.size()));
             size_t m = r.first.size(), n = c.q_dim();
             Arr vals(m), jac({m, n});
             auto vv = vals.mutable_unchecked<1>();
             auto jj = jac.mutable_unchecked<2>();
             for (size_t i = 0; i < m; i++) {
               vv(i) = r.first[i];
               for (size_t j = 0; j < n; j++) jj(i, j) = r.second[i][j];
             }
             return std::make_pair(vals, jac);
           })
      .def("get_all_spheres",
           [](SphereCollisionCst& c) {
             auto sp = c.get_all_spheres();
             Arr C({sp.size(), (size_t)3}), R(sp.size());
             auto cc = C.mutable_unchecked<2>();
             auto rr = R.mutable_unchecked<1>();
             for (size_t i = 0; i < sp.size(); i++) {
               for (int k = 0; k < 3; k++) cc(i, k) = sp[i].first[k];
               rr(i) = sp[i].second;
             }
             return std::make_pair(C, R);
           })
      .def("get_group_spheres",
           [](SphereCollisionCst& c) {
             auto sp = c.get_group_spheres();
             Arr C({sp.size(), (size_t)3}), R(sp.size());
             auto cc = C.mutable_unchecked<2>();
             auto rr = R.mutable_unchecked<1>();
             for (size_t i = 0; i < sp.size(); i++) {
               for (int k = 0; k < 3; k++) cc(i, k) = sp[i].first[k];
               rr(i) = sp[i].second;
             }
             return std::make_pair(C, R);
           })
      .def("n_spheres",
           [](const SphereCollisionCst& c) {
             size_t n = 0;
             for (auto& g : c.groups_) n += g.radii.size();
             return n;
           })
      .def("n_prims", [](const SphereCollisionCst& c) { return c.all_sdfs_.size(); })
      // ---- batch loops (each worker owns a copy of the constraint)
      .def("is_valid_batch",
           [](const SphereCollisionCst& c, const Arr& Q, int n_threads) {
             size_t B, D;
             take_rows(Q, B, D);
             py::array_t<uint8_t> out(B);
             uint8_t* o = out.mutable_data();
             const double* q = Q.data();
             {
               py::gil_scoped_release rel;
               parallel_for(B, n_threads, [&](size_t b, size_t e, int) {
                 SphereCollisionCst local = c;
                 for (size_t i = b; i < e; i++) o[i] = local.is_valid(row(q + i * D, D)) ? 1 : 0;
               });
             }
             return out;
           },
           py::arg("Q"), py::arg("n_threads") = 1)
      .def("sphere_clearance_batch",
           [](const SphereCollisionCst& c, const Arr& Q, bool accurate, int n_threads) {
             size_t B, D;
             take_rows(Q, B, D);
             size_t S = 0;
             for (auto& g : c.groups_) S += g.radii.size();
             Arr out({B, S});
             double* o = out.mutable_data();
             const double* q = Q.data();
             {
               py::gil_scoped_release rel;
               parallel_for(B, n_threads, [&](size_t b, size_t e, int) {
                 SphereCollisionCst local = c;
                 for (size_t i = b; i < e; i++) sphere_clearances(local, row(q + i * D, D), accurate, o + i * S);
               });
             }
             return out;
           },
           py::arg("Q"), py::arg("accurate") = false, py::arg("n_threads") = 1)
      .def("sphere_centers_batch",
           [](const SphereCollisionCst& c, const Arr& Q, bool accurate) {
             size_t B, D;
             take_rows(Q, B, D);
             size_t S = 0;
             for (auto& g : c.groups_) S += g.radii.size();
             Arr out({B, S, (size_t)3});
             double* o = out.mutable_data();
             SphereCollisionCst local = c;
             for (size_t i = 0; i < B; i++) {
               local.update_kintree(row(Q.data() + i * D, D), accurate);
               auto sp = local.get_all_spheres();
               for (size_t s = 0; s < S; s++)
                 for (int k = 0; k < 3; k++) o[(i * S + s) * 3 + k] = sp[s].first[k];
             }
             return out;
           },
           py::arg("Q"), py::arg("accurate") = false)
      .def("evaluate_batch",
           [](const SphereCollisionCst& c, const Arr& Q, int n_threads) {
             size_t B, D;
             take_rows(Q, B, D);
             size_t m = c.cst_dim(), n = c.q_dim();
             Arr vals({B, m}), jac({B, m, n});
             double* vo = vals.mutable_data();
             double* jo = jac.mutable_data();
             const double* q = Q.data();
             {
               py::gil_scoped_release rel;
               parallel_for(B, n_threads, [&](size_t b, size_t e, int) {
                 SphereCollisionCst local = c;
                 for (size_t i = b; i < e; i++) {
                   auto r = local.evaluate(row(q + i * D, D));
                   for (size_t k = 0; k < m; k++) {
                     vo[i * m + k] = r.first[k];
                     for (size_t j = 0; j < n; j++) jo[(i * m + k) * n + j] = r.second[k][j];
                   }
                 }
               });
             }
             return std::make_pair(vals, jac);
           },
           py::arg("Q"), py::arg("n_threads") = 1)
      // signed distance of every (sphere, primitive) pair with exact FK:
      // d[b, s, p] = sdf_p(c_s) - r_s
      .def("pair_clearance_batch",
           [](const SphereCollisionCst& c, const Arr& Q, bool accurate, int n_threads) {
             size_t B, D;
             take_rows(Q, B, D);
             size_t S = 0;
             for (auto& g : c.groups_) S += g.radii.size();
             size_t P = c.all_sdfs_.size();
             Arr out({B, S, P});
             double* o = out.mutable_data();
             const double* q = Q.data();
             {
               py::gil_scoped_release rel;
               parallel_for(B, n_threads, [&](size_t b, size_t e, int) {
                 SphereCollisionCst local = c;
                 for (size_t i = b; i < e; i++) {
                   local.update_kintree(row(q + i * D, D), accurate);
                   auto sp = local.get_all_spheres();
                   for (size_t s = 0; s < S; s++)
                     for (size_t p = 0; p < P; p++)
                       o[(i * S + s) * P + p] = local.all_sdfs_[p].evaluate(sp[s].first) - sp[s].second;
                 }
               });
             }
             return out;
           },
           py::arg("Q"), py::arg("accurate") = true, py::arg("n_threads") = 1)
      .def("check_motions",
           [](const SphereCollisionCst& c, const Arr& S1, const Arr& S2, bool box, const std::vector<double>& width,
              double resolution, int n_threads) {
             size_t E, D;
             take_rows(S1, E, D);
             py::array_t<uint8_t> out(E);
             py::array_t<int32_t> nchk(E);
             uint8_t* o = out.mutable_data();
             int32_t* nc = nchk.mutable_data();
             const double* a = S1.data();
             const double* b2 = S2.data();
             {
               py::gil_scoped_release rel;
               parallel_for(E, n_threads, [&](size_t b, size_t e, int) {
                 SphereCollisionCst local = c;
                 auto iv = [&](const std::vector<double>& q) { return local.is_valid(q); };
                 for (size_t i = b; i < e; i++) {
                   auto s1 = row(a + i * D, D), s2 = row(b2 + i * D, D);
                   double ratio = box ? box_step_ratio(s1, s2, width) : euclid_step_ratio(s1, s2, resolution);
                   int n = 0;
                   o[i] = check_motion(s1, s2, ratio, iv, &n) ? 1 : 0;
                   nc[i] = n;
                 }
               });
             }
             return std::make_pair(out, nchk);
           },
           py::arg("S1"), py::arg("S2"), py::arg("box"), py::arg("width"), py::arg("resolution"),
           py::arg("n_threads") = 1);
}
