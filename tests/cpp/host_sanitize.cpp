// Drives the library's host code -- the URDF reader, the host mirror of the
// reference classes and the C ABI's table builders (robot, scene, cull
// tables, kept-pose slots, cloud tile sizing, trajectory and SQP offsets,
// graphs) -- on the CPU, for a run under AddressSanitizer and
// UndefinedBehaviorSanitizer against host stand-ins of the HIP runtime
// (hip_host_stub.cpp).  TEST INFRASTRUCTURE ONLY: kernels do not run, so
// results are not checked here (the GPU tests do that); the run checks that
// the host code touches only memory it owns and has no undefined behaviour.
//
// Input: a case file (tests/test_host_sanitize.py writes them), one record a
// line: urdf_file PATH | joints N NAME... | base T | spec LINK ONLY K (x y z r)*K
// | pair A B | fixed|prim box wx wy wz px py pz R(9) | ... sphere r p R |
// ... cylinder r h p R | ... ground h | cloud N RC (x y z)*N | q B (values)*B*dof
// | width (values)*dof
#include <fstream>
#include <iostream>
#include <sstream>
#include <unordered_set>

#include "../../plainmp_amd/csrc/host/constraint.hpp"

using namespace pmp_host;

namespace {

SDFBase::Ptr read_prim(std::istringstream& in) {
  std::string kind;
  in >> kind;
  auto pose = [&](void) {
    Vec3 p;
    Mat3 R;
    in >> p[0] >> p[1] >> p[2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) in >> R[3 * i + j];
    return Pose(p, R);
  };
  if (kind == "box") {
    Vec3 w;
    in >> w[0] >> w[1] >> w[2];
    return std::make_shared<BoxSDF>(w, pose());
  }
  if (kind == "sphere") {
    double r;
    in >> r;
    return std::make_shared<SphereSDF>(r, pose());
  }
  if (kind == "cylinder") {
    double r, h;
    in >> r >> h;
    return std::make_shared<CylinderSDF>(r, h, pose());
  }
  if (kind == "ground") {
    double h;
    in >> h;
    return std::make_shared<GroundSDF>(h);
  }
  throw std::runtime_error("unknown primitive " + kind);
}

// The reference's visiting order with libstdc++'s own unordered_set
// (primitive_sphere_collision.cpp:53-96), against the host mirror's model.
std::vector<size_t> stdset_order(const std::vector<double>& pos) {
  const size_t n = pos.size() / 3;
  std::unordered_set<size_t> open;
  open.reserve(n);
  for (size_t i = 0; i < n; ++i) open.insert(i);
  std::vector<size_t> order{*open.begin()};
  open.erase(order[0]);
  while (!open.empty()) {
    double far = 0.0;
    auto pick = open.begin();
    for (auto it = open.begin(); it != open.end(); ++it) {
      double near = std::numeric_limits<double>::max();
      for (size_t v : order) {
        const double dx = pos[3 * *it] - pos[3 * v], dy = pos[3 * *it + 1] - pos[3 * v + 1],
                     dz = pos[3 * *it + 2] - pos[3 * v + 2];
        near = std::min(near, std::sqrt(dx * dx + dy * dy + dz * dz));
      }
      if (near > far) {
        far = near;
        pick = it;
      }
    }
    order.push_back(*pick);
    open.erase(pick);
  }
  return order;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::cerr << "usage: host_sanitize CASE_FILE\n";
    return 2;
  }
  std::ifstream f(argv[1]);
  std::string line, urdf;
  std::vector<std::string> joints;
  int base = 0;
  std::vector<SphereAttachmentSpec> specs;
  std::vector<std::pair<std::string, std::string>> pairs;
  std::vector<SDFBase::Ptr> prims, fixed;
  std::vector<double> Q, width;
  int64_t B = 0;
  while (std::getline(f, line)) {
    std::istringstream in(line);
    std::string key;
    in >> key;
    if (key == "urdf_file") {
      std::string path;
      in >> path;
      std::ifstream u(path);
      std::stringstream ss;
      ss << u.rdbuf();
      urdf = ss.str();
    } else if (key == "joints") {
      int n;
      in >> n;
      joints.resize(n);
      for (auto& j : joints) in >> j;
    } else if (key == "base") {
      in >> base;
    } else if (key == "spec") {
      SphereAttachmentSpec s;
      int only, k;
      in >> s.parent_link_name >> only >> k;
      s.only_self_collision = only != 0;
      for (int i = 0; i < k; ++i) {
        double x, y, z, r;
        in >> x >> y >> z >> r;
        s.relative_positions.insert(s.relative_positions.end(), {x, y, z});
        s.radii.push_back(r);
      }
      specs.push_back(s);
    } else if (key == "pair") {
      std::string a, b;
      in >> a >> b;
      pairs.push_back({a, b});
    } else if (key == "prim") {
      prims.push_back(read_prim(in));
    } else if (key == "fixed") {
      fixed.push_back(read_prim(in));
    } else if (key == "cloud") {
      int64_t n;
      double rc;
      in >> n >> rc;
      std::vector<double> pts(3 * n);
      for (auto& v : pts) in >> v;
      prims.push_back(std::make_shared<CloudSDF>(pts, rc));
    } else if (key == "q") {
      in >> B;
      double v;
      while (in >> v) Q.push_back(v);
    } else if (key == "width") {
      double v;
      while (in >> v) width.push_back(v);
    }
  }
  auto kin = std::make_shared<KinematicModel>(urdf);
  std::optional<SDFBase::Ptr> fx;
  if (!fixed.empty()) fx = std::make_shared<UnionSDF>(fixed);
  auto cst = std::make_shared<SphereCollisionCst>(kin, joints, base, specs, pairs, fx, true);
  if (!prims.empty()) cst->set_sdf(std::make_shared<UnionSDF>(prims));
  const int64_t D = (int64_t)cst->q_dim();
  if ((int64_t)Q.size() != B * D) throw std::runtime_error("configuration count mismatch");
  for (const auto& s : specs) {  // the host mirror's reorder model against libstdc++
    if (informative_order(s.relative_positions) != stdset_order(s.relative_positions)) {
      std::cerr << "reorder mismatch on " << s.parent_link_name << "\n";
      return 1;
    }
  }
  auto valid = cst->is_valid_batch(Q.data(), B);
  std::vector<double> vals, jac;
  cst->evaluate_batch(Q.data(), B, vals, jac);
  auto clear = cst->sphere_clearance_batch(Q.data(), B, false);
  auto cen = cst->sphere_centers_batch(Q.data(), B, true);
  std::vector<float> dist, grad;
  cst->pair_clearance_batch(Q.data(), std::min<int64_t>(B, 16), true, dist, grad);
  auto lj = cst->link_jacobian_batch(Q.data(), B, (int32_t)kin->get_link_ids({specs.back().parent_link_name})[0],
                                     {0.01, 0.02, 0.03});
  const std::vector<int64_t> off{0, B / 3, B / 3, B};
  std::vector<double> tv, tj, tc, tg;
  cst->traj_cost_grad_batch(Q.data(), B, off, std::vector<double>(D, 1.5), tv, tj, tc, tg);
  auto sqp = cst->traj_sqp_ineq(Q.data(), B, off, std::vector<double>(D, 0.1));
  std::vector<double> q0(Q.begin(), Q.begin() + D);
  cst->is_valid(q0);
  cst->evaluate(q0);
  cst->update_kintree(q0, true);
  cst->get_all_spheres();
  cst->get_group_spheres();
  // validity without allocation, and a changed band (cull tables rebuilt)
  {
    pmp_context* ctx = cst->device_context(nullptr);
    check(pmp_set_band(ctx, 0.01));
    cst->is_valid_batch(Q.data(), B);
  }
  // motion validators, graphs of edges and trajectories
  ValidatorConfig vc;
  vc.box_width = width;
  MotionValidator mv(cst, vc);
  const int64_t E = B / 2;
  std::vector<int32_t> n_checked;
  mv.check_motions(Q.data(), Q.data() + E * D, E, &n_checked);
  mv.check_path(Q.data(), B);
  int64_t m = 0;
  mv.shortcut_path(Q.data(), B, 16, 2, 7, &m);
  ValidatorConfig ve;
  ve.type = ValidatorConfig::Type::EUCLIDEAN;
  ve.resolution = 0.05;
  MotionValidator(cst, ve).check_motions(Q.data(), Q.data() + E * D, E, nullptr);
  {
    pmp_context* ctx = cst->device_context(nullptr);
    std::vector<double> soa(B * D);
    for (int64_t b = 0; b < B; ++b)
      for (int64_t j = 0; j < D; ++j) soa[j * B + b] = Q[b * D + j];
    std::vector<uint8_t> v(B);
    std::vector<int32_t> nc(B);
    pmp_validator_desc d = mv.desc();
    pmp_graph* g = nullptr;
    check(pmp_check_edges_graph(ctx, soa.data(), soa.data() + E, B, E, &d, v.data(), nc.data(), &g));
    check(pmp_graph_launch(g));
    check(pmp_graph_destroy(g));
    check(pmp_is_valid_batch_noalloc(ctx, soa.data(), B, B, v.data(), nullptr));
    const size_t C = cst->cst_dim();
    std::vector<float> gv(C * B + 1), gj(C * D * B + 1), gc(3), gg(D * B);
    check(pmp_traj_cost_grad_graph(ctx, soa.data(), B, B, off.data(), 3, nullptr, gv.data(), gj.data(), gc.data(),
                                   gg.data(), &g));
    check(pmp_graph_launch(g));
    check(pmp_graph_destroy(g));
  }
  // argument errors come back as exceptions with the reference's types
  int errors = 0;
  try {
    SphereCollisionCst(kin, {"no_such_joint"}, base, specs, pairs, fx, true);
  } catch (const std::invalid_argument&) {
    ++errors;
  }
  try {
    SphereCollisionCst(kin, joints, base, specs, {{"no_such_link", specs[0].parent_link_name}}, fx, true);
  } catch (const std::runtime_error&) {
    ++errors;
  }
  if (errors != 2) {
    std::cerr << "expected two argument errors, got " << errors << "\n";
    return 1;
  }
  std::cout << "ok B=" << B << " valid=" << valid.size() << " rows=" << vals.size() << " sqp_nnz=" << sqp.values.size()
            << "\n";
  return 0;
}
