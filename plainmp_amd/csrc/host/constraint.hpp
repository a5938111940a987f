// SphereCollisionCst and the motion validators above the C ABI.
// Construction follows the reference ctor (constraints/primitive_sphere_collision.cpp:117-179):
// groups with mean centre and radius max(|p - c| + r), the optional
// informative sphere reorder, self pairs resolved by link name, and the SDF
// tree flattened at set_sdf time with primitives shared by pointer.
#pragma once

#include <random>

#include "host.hpp"

namespace pmp_host {

struct SphereAttachmentSpec {  // primitive_sphere_collision.hpp:16-21
  std::string parent_link_name;
  std::vector<double> relative_positions;  // (k, 3) row-major
  std::vector<double> radii;
  bool only_self_collision = false;
};

// reorder_spheres_for_informative_collision_check (primitive_sphere_collision.cpp:53-96):
// farthest-point order seeded from unordered_set<size_t>::begin().  For keys
// 0..n-1 inserted in increasing order after reserve(n), libstdc++ iterates
// them n-1 ... 0 and erase() keeps the order of the rest, so the seed is n-1
// and ties go to the first candidate in descending index order.
inline std::vector<size_t> informative_order(const std::vector<double>& pos) {
  const size_t n = pos.size() / 3;
  std::vector<size_t> unvisited;
  for (size_t i = 0; i < n; ++i) unvisited.push_back(n - 1 - i);
  std::vector<size_t> visited{unvisited.front()};
  unvisited.erase(unvisited.begin());
  auto dist = [&](size_t a, size_t b) {
    double dx = pos[3 * a] - pos[3 * b], dy = pos[3 * a + 1] - pos[3 * b + 1], dz = pos[3 * a + 2] - pos[3 * b + 2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
  };
  while (!unvisited.empty()) {
    double max_dist = 0.0;
    size_t best = 0;
    for (size_t k = 0; k < unvisited.size(); ++k) {
      double mind = std::numeric_limits<double>::max();
      for (size_t v : visited) mind = std::min(mind, dist(unvisited[k], v));
      if (mind > max_dist) {
        max_dist = mind;
        best = k;
      }
    }
    visited.push_back(unvisited[best]);
    unvisited.erase(unvisited.begin() + best);
  }
  return visited;
}

struct SphereGroup {
  std::string parent_link_name;
  size_t parent_link_id = 0;
  std::vector<double> rel;  // (k, 3)
  std::vector<double> radii;
  double group_radius = 0;
  Vec3 group_rel{0, 0, 0};
  bool only_self_collision = false;
};

class SphereCollisionCst {
 public:
  SphereCollisionCst(std::shared_ptr<KinematicModel> kin, const std::vector<std::string>& control_joint_names,
                     int base_type, const std::vector<SphereAttachmentSpec>& specs,
                     const std::vector<std::pair<std::string, std::string>>& selcol_pairs,
                     std::optional<SDFBase::Ptr> fixed_sdf, bool reorder_spheres = true)
      : kin_(std::move(kin)), control_names_(control_joint_names), base_type_(base_type) {
    if (!kin_) throw std::invalid_argument("null kinematic model");
    if (base_type < FIXED || base_type > FLOATING) throw std::invalid_argument("unknown base type");
    control_ids_ = kin_->get_joint_ids(control_joint_names);
    if (fixed_sdf && *fixed_sdf) fixed_sdf_ = *fixed_sdf;
    for (auto& spec : specs) {
      if (spec.relative_positions.size() != 3 * spec.radii.size() || spec.radii.empty())
        throw std::invalid_argument("sphere spec needs (k, 3) positions and k radii, k >= 1");
      SphereGroup g;
      g.parent_link_name = spec.parent_link_name;
      g.parent_link_id = kin_->get_link_ids({spec.parent_link_name})[0];
      const size_t k = spec.radii.size();
      Vec3 c{0, 0, 0};
      for (size_t j = 0; j < k; ++j)
        for (int a = 0; a < 3; ++a) c[a] += spec.relative_positions[3 * j + a];
      for (int a = 0; a < 3; ++a) c[a] /= (double)k;
      double max_dist = 0.0;
      for (size_t j = 0; j < k; ++j) {
        double dx = spec.relative_positions[3 * j] - c[0], dy = spec.relative_positions[3 * j + 1] - c[1],
               dz = spec.relative_positions[3 * j + 2] - c[2];
        max_dist = std::max(max_dist, std::sqrt(dx * dx + dy * dy + dz * dz) + spec.radii[j]);
      }
      g.rel = spec.relative_positions;
      g.radii = spec.radii;
      g.group_radius = max_dist;
      g.group_rel = c;
      g.only_self_collision = spec.only_self_collision;
      groups_.push_back(g);
      if (reorder_spheres)  // re-applied to every group on each push (:155-159)
        for (auto& gg : groups_) {
          auto ord = informative_order(gg.rel);
          std::vector<double> rp, rr;
          for (auto i : ord) {
            rp.insert(rp.end(), {gg.rel[3 * i], gg.rel[3 * i + 1], gg.rel[3 * i + 2]});
            rr.push_back(gg.radii[i]);
          }
          gg.rel = rp;
          gg.radii = rr;
        }
    }
    for (auto& pr : selcol_pairs) {
      int a = -1, b = -1;
      for (size_t i = 0; i < groups_.size(); ++i) {
        if (a < 0 && groups_[i].parent_link_name == pr.first) a = (int)i;
        if (b < 0 && groups_[i].parent_link_name == pr.second) b = (int)i;
      }
      if (a < 0 || b < 0) throw std::runtime_error("(cpp) Invalid pair of link names for self collision");
      self_pairs_.push_back({a, b});
    }
    set_all_sdfs();
  }

  void set_sdf(const SDFBase::Ptr& sdf) {
    sdf_ = sdf;
    set_all_sdfs();
  }
  SDFBase::Ptr get_sdf() const { return sdf_; }

  size_t q_dim() const { return control_ids_.size() + (base_type_ == FLOATING ? 6 : 0) + (base_type_ == PLANAR ? 3 : 0); }
  bool ext_enabled() const { return !all_sdfs_.empty(); }
  bool self_enabled() const { return !self_pairs_.empty(); }
  size_t cst_dim() const { return (ext_enabled() ? 1 : 0) + (self_enabled() ? 1 : 0); }
  size_t n_spheres() const {
    size_t n = 0;
    for (auto& g : groups_) n += g.radii.size();
    return n;
  }
  size_t n_prims() const { return all_sdfs_.size(); }
  std::shared_ptr<KinematicModel> get_kin() const { return kin_; }
  std::vector<std::string> get_control_joint_names() const { return control_names_; }
  int base_type() const { return base_type_; }
  const std::vector<SphereGroup>& groups() const { return groups_; }

  // ConstraintBase::update_kintree (constraints/primitive.hpp:38-62): the
  // model state is updated so that later state queries see q.
  void update_kintree(const std::vector<double>& q, bool high_accuracy) {
    check_q(q);
    const size_t n = control_ids_.size();
    if (base_type_ != FIXED) {
      kin_->set_joint_angles(control_ids_, std::vector<double>(q.begin(), q.begin() + n), true);
      Quat bq;
      std::array<double, 7> pose{};
      if (base_type_ == FLOATING) {
        bq = quat_from_rpy(q[n + 3], q[n + 4], q[n + 5]);
        pose = {q[n], q[n + 1], q[n + 2], bq.x, bq.y, bq.z, bq.w};
      } else {
        bq = quat_from_rpy(0, 0, q[n + 2]);
        pose = {q[n], q[n + 1], 0.0, bq.x, bq.y, bq.z, bq.w};
      }
      kin_->set_base_pose(pose);
    } else {
      kin_->set_joint_angles(control_ids_, q, high_accuracy);
    }
    state_q_ = q;
    state_accurate_ = base_type_ != FIXED || high_accuracy;
  }

  // IneqConstraintBase::is_valid (primitive.hpp:113-117): one configuration on the GPU.
  bool is_valid(const std::vector<double>& q) {
    update_kintree(q, false);
    auto v = is_valid_batch(q.data(), 1);
    return v[0] != 0;
  }

  // ConstraintBase::evaluate (primitive.hpp:74-79): (values, Jacobian rows).
  std::pair<std::vector<double>, std::vector<double>> evaluate(const std::vector<double>& q) {
    update_kintree(q, true);
    std::vector<double> vals, jac;
    evaluate_batch(q.data(), 1, vals, jac);
    return {vals, jac};
  }

  // ---------------------------------------------------------- batch (host arrays)
  std::vector<uint8_t> is_valid_batch(const double* Q, int64_t B, int32_t* n_recheck = nullptr) {
    sync_batch_tables();
    std::vector<uint8_t> out(B);
    if (B == 0) return out;
    const int D = (int)q_dim();
    double* dq = batch_.upload_soa(0, Q, B, D);
    uint8_t* dv = (uint8_t*)batch_.buf(1, B);
    int32_t* dn = (int32_t*)batch_.buf(2, 4);
    int32_t zero = 0;
    batch_.h2d(dn, &zero, 4);
    check(pmp_is_valid_batch(batch_.ctx(), dq, B, B, dv, dn));
    batch_.d2h(out.data(), dv, B);
    if (n_recheck) batch_.d2h(n_recheck, dn, 4);
    return out;
  }
  // vals (B, cst_dim), jac (B, cst_dim, q_dim), row-major
  void evaluate_batch(const double* Q, int64_t B, std::vector<double>& vals, std::vector<double>& jac) {
    sync_batch_tables();
    const int D = (int)q_dim(), C = (int)cst_dim();
    vals.assign((size_t)B * C, 0.0);
    jac.assign((size_t)B * C * D, 0.0);
    if (B == 0 || C == 0) return;
    double* dq = batch_.upload_soa(0, Q, B, D);
    float* dv = (float*)batch_.buf(1, sizeof(float) * B * C);
    float* dj = (float*)batch_.buf(2, sizeof(float) * B * C * D);
    check(pmp_evaluate_batch(batch_.ctx(), dq, B, B, dv, dj));
    std::vector<float> hv((size_t)B * C), hj((size_t)B * C * D);
    batch_.d2h(hv.data(), dv, hv.size() * sizeof(float));
    batch_.d2h(hj.data(), dj, hj.size() * sizeof(float));
    for (int64_t b = 0; b < B; ++b)
      for (int k = 0; k < C; ++k) {
        vals[b * C + k] = hv[(size_t)k * B + b];
        for (int j = 0; j < D; ++j) jac[((size_t)b * C + k) * D + j] = hj[((size_t)k * D + j) * B + b];
      }
  }
  // Trajectories back to back in X (N, q_dim) row-major, trajectory i =
  // rows [offsets[i], offsets[i+1]): per-waypoint evaluate rows, smoothness
  // cost per trajectory and its gradient (N, q_dim).
  void traj_cost_grad_batch(const double* X, int64_t N, const std::vector<int64_t>& offsets,
                            const std::vector<double>& weights, std::vector<double>& vals, std::vector<double>& jac,
                            std::vector<double>& cost, std::vector<double>& grad) {
    sync_batch_tables();
    const int D = (int)q_dim(), C = (int)cst_dim();
    if (offsets.empty()) throw std::invalid_argument("offsets must hold n_traj + 1 entries");
    if (!weights.empty() && (int)weights.size() != D) throw std::invalid_argument("weights must have q_dim entries");
    const int32_t n_traj = (int32_t)offsets.size() - 1;
    vals.assign((size_t)N * C, 0.0);
    jac.assign((size_t)N * C * D, 0.0);
    cost.assign(n_traj, 0.0);
    grad.assign((size_t)N * D, 0.0);
    double* dx = N ? batch_.upload_soa(0, X, N, D) : nullptr;
    float* dv = (float*)batch_.buf(1, sizeof(float) * std::max<int64_t>(N * C, 1));
    float* dj = (float*)batch_.buf(2, sizeof(float) * std::max<int64_t>(N * C * D, 1));
    float* dc = (float*)batch_.buf(3, sizeof(float) * std::max<int32_t>(n_traj, 1));
    float* dg = (float*)batch_.buf(4, sizeof(float) * std::max<int64_t>(N * D, 1));
    check(pmp_traj_cost_grad_batch(batch_.ctx(), dx, N, N, offsets.data(), n_traj,
                                   weights.empty() ? nullptr : weights.data(), dv, dj, dc, dg));
    if (n_traj == 0) return;
    std::vector<float> hv((size_t)N * C), hj((size_t)N * C * D), hc(n_traj), hg((size_t)N * D);
    batch_.d2h(hc.data(), dc, hc.size() * sizeof(float));
    for (int32_t i = 0; i < n_traj; ++i) cost[i] = hc[i];
    if (N == 0) return;
    batch_.d2h(hv.data(), dv, hv.size() * sizeof(float));
    batch_.d2h(hj.data(), dj, hj.size() * sizeof(float));
    batch_.d2h(hg.data(), dg, hg.size() * sizeof(float));
    for (int64_t n = 0; n < N; ++n) {
      for (int k = 0; k < C; ++k) {
        vals[n * C + k] = hv[(size_t)k * N + n];
        for (int j = 0; j < D; ++j) jac[((size_t)n * C + k) * D + j] = hj[((size_t)k * D + j) * N + n];
      }
      for (int j = 0; j < D; ++j) grad[(size_t)n * D + j] = hg[(size_t)j * N + n];
    }
  }
  // The SQP inequality sequence of every trajectory as CSC pieces
  // (pmp_traj_sqp_ineq); offsets of each piece in row_off / nnz_off / cp_off.
  struct SqpPieces {
    std::vector<double> c, values;
    std::vector<int32_t> colptr, rowind;
    std::vector<int64_t> row_off, nnz_off, cp_off;
  };
  SqpPieces traj_sqp_ineq(const double* X, int64_t N, const std::vector<int64_t>& offsets,
                          const std::vector<double>& msbox) {
    sync_batch_tables();
    const int D = (int)q_dim();
    if (offsets.empty()) throw std::invalid_argument("offsets must hold n_traj + 1 entries");
    if (!msbox.empty() && (int)msbox.size() != D) throw std::invalid_argument("msbox must have q_dim entries");
    const int32_t n_traj = (int32_t)offsets.size() - 1;
    SqpPieces p;
    p.row_off.resize(n_traj + 1);
    p.nnz_off.resize(n_traj + 1);
    p.cp_off.resize(n_traj + 1);
    check(pmp_traj_sqp_sizes(batch_.ctx(), offsets.data(), n_traj, msbox.empty() ? 0 : 1, p.row_off.data(),
                             p.nnz_off.data(), p.cp_off.data()));
    const int64_t R = p.row_off[n_traj], Z = p.nnz_off[n_traj], P = p.cp_off[n_traj];
    double* dx = N ? batch_.upload_soa(0, X, N, D) : nullptr;
    double* dc = (double*)batch_.buf(1, sizeof(double) * std::max<int64_t>(R, 1));
    int32_t* dp = (int32_t*)batch_.buf(2, sizeof(int32_t) * std::max<int64_t>(P, 1));
    int32_t* dr = (int32_t*)batch_.buf(3, sizeof(int32_t) * std::max<int64_t>(Z, 1));
    double* dv = (double*)batch_.buf(4, sizeof(double) * std::max<int64_t>(Z, 1));
    check(pmp_traj_sqp_ineq(batch_.ctx(), dx, N, N, offsets.data(), n_traj, msbox.empty() ? nullptr : msbox.data(),
                            dc, dp, dr, dv));
    p.c.resize(R);
    p.colptr.resize(P);
    p.rowind.resize(Z);
    p.values.resize(Z);
    if (R) batch_.d2h(p.c.data(), dc, sizeof(double) * R);
    if (P) batch_.d2h(p.colptr.data(), dp, sizeof(int32_t) * P);
    if (Z) {
      batch_.d2h(p.rowind.data(), dr, sizeof(int32_t) * Z);
      batch_.d2h(p.values.data(), dv, sizeof(double) * Z);
    }
    return p;
  }
  // (B, 3, q_dim) Jacobian of a point fixed to a link (rotation ignored)
  std::vector<float> link_jacobian_batch(const double* Q, int64_t B, int32_t link, const std::vector<double>& point) {
    sync_batch_tables();
    const int D = (int)q_dim();
    if (!point.empty() && point.size() != 3) throw std::invalid_argument("point must have 3 entries");
    std::vector<float> out((size_t)B * 3 * D);
    if (B == 0) return out;
    double* dq = batch_.upload_soa(0, Q, B, D);
    float* dj = (float*)batch_.buf(1, sizeof(float) * B * 3 * D);
    check(pmp_link_jacobian_batch(batch_.ctx(), dq, B, B, link, point.empty() ? nullptr : point.data(), dj));
    std::vector<float> h((size_t)B * 3 * D);
    batch_.d2h(h.data(), dj, h.size() * sizeof(float));
    for (int64_t b = 0; b < B; ++b)
      for (int k = 0; k < 3; ++k)
        for (int j = 0; j < D; ++j) out[((size_t)b * 3 + k) * D + j] = h[((size_t)k * D + j) * B + b];
    return out;
  }
  // (B, S) per-sphere minimum clearance
  std::vector<float> sphere_clearance_batch(const double* Q, int64_t B, bool accurate) {
    sync_batch_tables();
    const int D = (int)q_dim();
    const size_t S = n_spheres();
    std::vector<float> out((size_t)B * S);
    if (B == 0 || S == 0) return out;
    double* dq = batch_.upload_soa(0, Q, B, D);
    float* dc = (float*)batch_.buf(1, sizeof(float) * B * S);
    check(pmp_sphere_clearance_batch(batch_.ctx(), dq, B, B, accurate ? 1 : 0, dc));
    std::vector<float> h((size_t)B * S);
    batch_.d2h(h.data(), dc, h.size() * sizeof(float));
    for (int64_t b = 0; b < B; ++b)
      for (size_t s = 0; s < S; ++s) out[b * S + s] = h[s * B + b];
    return out;
  }
  // (B, S, 3) sphere centres
  std::vector<float> sphere_centers_batch(const double* Q, int64_t B, bool accurate) {
    sync_batch_tables();
    const int D = (int)q_dim();
    const size_t S = n_spheres();
    std::vector<float> out((size_t)B * S * 3);
    if (B == 0 || S == 0) return out;
    double* dq = batch_.upload_soa(0, Q, B, D);
    float* dc = (float*)batch_.buf(1, sizeof(float) * B * S * 3);
    check(pmp_fk_spheres_batch(batch_.ctx(), dq, B, B, accurate ? 1 : 0, dc));
    std::vector<float> h((size_t)B * S * 3);
    batch_.d2h(h.data(), dc, h.size() * sizeof(float));
    for (int64_t b = 0; b < B; ++b)
      for (size_t s = 0; s < S; ++s)
        for (int k = 0; k < 3; ++k) out[(b * S + s) * 3 + k] = h[(s * 3 + k) * B + b];
    return out;
  }
  // dist (B, S, P), grad (B, S, P, q_dim) or empty
  void pair_clearance_batch(const double* Q, int64_t B, bool with_grad, std::vector<float>& dist,
                            std::vector<float>& grad) {
    sync_batch_tables();
    const int D = (int)q_dim();
    const size_t S = n_spheres(), P = n_prims();
    dist.assign((size_t)B * S * P, 0.f);
    grad.assign(with_grad ? (size_t)B * S * P * D : 0, 0.f);
    if (B == 0 || S == 0 || P == 0) return;
    double* dq = batch_.upload_soa(0, Q, B, D);
    float* dd = (float*)batch_.buf(1, sizeof(float) * B * S * P);
    float* dg = with_grad ? (float*)batch_.buf(2, sizeof(float) * B * S * P * D) : nullptr;
    check(pmp_pair_clearance_batch(batch_.ctx(), dq, B, B, dd, dg));
    std::vector<float> h((size_t)B * S * P);
    batch_.d2h(h.data(), dd, h.size() * sizeof(float));
    for (int64_t b = 0; b < B; ++b)
      for (size_t r = 0; r < S * P; ++r) dist[b * S * P + r] = h[r * B + b];
    if (with_grad) {
      std::vector<float> g((size_t)B * S * P * D);
      batch_.d2h(g.data(), dg, g.size() * sizeof(float));
      for (int64_t b = 0; b < B; ++b)
        for (size_t r = 0; r < S * P * D; ++r) grad[b * S * P * D + r] = g[r * B + b];
    }
  }

  // ------------------------------------------------- batch (device pointers)
  pmp_context* device_context(void* stream) {
    sync_batch_tables();
    check(pmp_set_stream(batch_.ctx(), stream));
    return batch_.ctx();
  }

  // ----------------------------------------------------------- state queries
  // get_all_spheres / get_group_spheres on the model's current state
  // (constraints/primitive_sphere_collision.cpp:401-424): the control joints
  // take the configuration of the last update_kintree; FK runs on the GPU.
  std::vector<std::pair<Vec3, double>> get_all_spheres() { return state_spheres(false); }
  std::vector<std::pair<Vec3, double>> get_group_spheres() { return state_spheres(true); }

  // Tables for a batched call (also used by validators).  Host-array calls
  // run on the null stream, ordered with their synchronous copies; calls on
  // device pointers set the caller's stream afterwards (device_context).
  void sync_batch_tables() {
    check(pmp_set_stream(batch_.ctx(), nullptr));
    sync_robot(batch_, control_ids_, base_type_, false, batch_robot_version_);
    sync_scene(batch_, batch_scene_sig_);
  }

 private:
  void check_q(const std::vector<double>& q) const {
    if (q.size() != q_dim())
      throw std::invalid_argument("configuration has " + std::to_string(q.size()) + " entries, expected " +
                                  std::to_string(q_dim()));
  }
  void set_all_sdfs() {
    all_sdfs_.clear();
    if (fixed_sdf_) fixed_sdf_->flatten(all_sdfs_);
    if (sdf_) sdf_->flatten(all_sdfs_);
    batch_scene_sig_.clear();
    state_scene_sig_.clear();
  }
  // Robot tables: node tree from the model, spheres from the groups.  With
  // `group_spheres` each group contributes one sphere (centre, group radius).
  void sync_robot(Device& dev, const std::vector<size_t>& q_joints, int base_type, bool group_spheres,
                  uint64_t& cached_version) {
    (void)cached_version;
    std::vector<double> sig = kin_->pack_signature(q_joints, base_type == FIXED);
    sig.push_back(group_spheres ? 1.0 : 0.0);
    auto it = robot_sig_.find(&dev);
    if (it != robot_sig_.end() && it->second == sig && !q_joints_changed(dev, q_joints)) return;
    PackedRobot pr;
    kin_->pack(q_joints, pr);
    std::vector<int32_t> s_node, s_group;
    std::vector<double> s_local, s_radius;
    std::vector<uint8_t> only_self;
    for (size_t g = 0; g < groups_.size(); ++g) {
      const auto& gr = groups_[g];
      only_self.push_back(gr.only_self_collision ? 1 : 0);
      if (group_spheres) {
        s_node.push_back((int32_t)gr.parent_link_id);
        s_group.push_back((int32_t)g);
        s_local.insert(s_local.end(), gr.group_rel.begin(), gr.group_rel.end());
        s_radius.push_back(gr.group_radius);
      } else {
        for (size_t i = 0; i < gr.radii.size(); ++i) {
          s_node.push_back((int32_t)gr.parent_link_id);
          s_group.push_back((int32_t)g);
          s_local.insert(s_local.end(), {gr.rel[3 * i], gr.rel[3 * i + 1], gr.rel[3 * i + 2]});
          s_radius.push_back(gr.radii[i]);
        }
      }
    }
    std::vector<int32_t> pairs;
    for (auto& p : self_pairs_) pairs.insert(pairs.end(), {p.first, p.second});
    pmp_robot_desc d{};
    d.n_nodes = (int32_t)pr.parent.size();
    d.node_parent = pr.parent.data();
    d.node_kind = pr.kind.data();
    d.node_q_col = pr.q_col.data();
    d.node_quat = pr.quat.data();
    d.node_trans = pr.trans.data();
    d.node_axis = pr.axis.data();
    d.node_identity = pr.identity.data();
    d.n_q_cols = (int32_t)q_joints.size();
    d.base_type = base_type;
    auto bp = kin_->get_base_pose();
    for (int k = 0; k < 7; ++k) d.base_pose[k] = bp[k];
    d.n_spheres = (int32_t)s_radius.size();
    d.sphere_node = s_node.data();
    d.sphere_local = s_local.data();
    d.sphere_radius = s_radius.data();
    d.sphere_group = s_group.data();
    d.n_groups = (int32_t)groups_.size();
    d.group_only_self = only_self.data();
    d.n_self_pairs = group_spheres ? 0 : (int32_t)self_pairs_.size();
    d.self_pairs = pairs.data();
    check(pmp_upload_robot(dev.ctx(), &d));
    robot_sig_[&dev] = sig;
    uploaded_q_[&dev] = q_joints;
  }
  bool q_joints_changed(Device& dev, const std::vector<size_t>& q) {
    auto it = uploaded_q_.find(&dev);
    return it == uploaded_q_.end() || it->second != q;
  }
  void sync_scene(Device& dev, std::vector<uint64_t>& sig) {
    std::vector<uint64_t> now;
    for (auto& p : all_sdfs_) {
      now.push_back((uint64_t)(uintptr_t)p.get());
      now.push_back(p->version);
    }
    now.push_back(all_sdfs_.size());
    if (now == sig) return;
    std::vector<pmp_prim_desc> d(all_sdfs_.size());
    for (size_t i = 0; i < all_sdfs_.size(); ++i) all_sdfs_[i]->describe(d[i]);
    pmp_scene_desc sc{(int32_t)d.size(), d.data()};
    check(pmp_upload_scene(dev.ctx(), &sc));
    sig = now;
  }
  std::vector<std::pair<Vec3, double>> state_spheres(bool groups) {
    // q columns: the control joints (and base) as last set through this constraint
    std::vector<double> q = state_q_;
    if (q.size() != q_dim()) {  // never updated: current model angles, fixed base pose
      q = kin_->get_joint_angles(control_ids_);
      q.resize(q_dim(), 0.0);
    }
    Device& dev = groups ? state_groups_ : state_;
    uint64_t& ver = groups ? state_groups_version_ : state_version_;
    sync_robot(dev, control_ids_, base_type_, groups, ver);
    const int D = (int)q_dim();
    const size_t S = groups ? groups_.size() : n_spheres();
    std::vector<std::pair<Vec3, double>> out;
    if (S == 0) return out;
    double* dq = dev.upload_soa(0, q.data(), 1, D);
    float* dc = (float*)dev.buf(1, sizeof(float) * S * 3);
    check(pmp_fk_spheres_batch(dev.ctx(), dq, 1, 1, state_accurate_ ? 1 : 0, dc));
    std::vector<float> h(S * 3);
    dev.d2h(h.data(), dc, h.size() * sizeof(float));
    size_t s = 0;
    for (auto& g : groups_) {
      if (groups) {
        out.push_back({{h[3 * s], h[3 * s + 1], h[3 * s + 2]}, g.group_radius});
        ++s;
      } else {
        for (double r : g.radii) {
          out.push_back({{h[3 * s], h[3 * s + 1], h[3 * s + 2]}, r});
          ++s;
        }
      }
    }
    return out;
  }

  std::shared_ptr<KinematicModel> kin_;
  std::vector<std::string> control_names_;
  std::vector<size_t> control_ids_;
  int base_type_;
  std::vector<SphereGroup> groups_;
  std::vector<std::pair<int, int>> self_pairs_;
  SDFBase::Ptr fixed_sdf_, sdf_;
  std::vector<std::shared_ptr<PrimitiveSDFBase>> all_sdfs_;
  Device batch_, state_, state_groups_;
  uint64_t batch_robot_version_ = 0, state_version_ = 0, state_groups_version_ = 0;
  std::vector<uint64_t> batch_scene_sig_, state_scene_sig_;
  std::unordered_map<const Device*, std::vector<size_t>> uploaded_q_;
  std::unordered_map<const Device*, std::vector<double>> robot_sig_;
  std::vector<double> state_q_;
  bool state_accurate_ = true;
};

// ValidatorConfig (ompl/ompl_thin_wrap.hpp:50-59)
struct ValidatorConfig {
  enum class Type { BOX = 0, EUCLIDEAN = 1 };
  Type type = Type::BOX;
  double resolution = 0.0;
  std::vector<double> box_width;
};

// Box / euclidean motion validators over a SphereCollisionCst
// (ompl/motion_validator.cpp:17-126): checkMotion for one edge, and
// check_motions for a batch in one persistent launch.
class MotionValidator {
 public:
  MotionValidator(std::shared_ptr<SphereCollisionCst> cst, const ValidatorConfig& vc) : cst_(std::move(cst)), vc_(vc) {
    if (!cst_) throw std::runtime_error("ineq_cst is nullptr. You should provide a constraint");
    if (vc.type == ValidatorConfig::Type::BOX) {
      if (vc.box_width.size() != cst_->q_dim()) throw std::runtime_error("box dimension and space dimension mismatch");
    } else if (vc.type != ValidatorConfig::Type::EUCLIDEAN) {
      throw std::runtime_error("unknown validator type");
    }
  }
  bool checkMotion(const std::vector<double>& s1, const std::vector<double>& s2) {
    if (s1.size() != cst_->q_dim() || s2.size() != cst_->q_dim())
      throw std::invalid_argument("state dimension mismatch");
    return check_motions(s1.data(), s2.data(), 1, nullptr)[0] != 0;
  }
  // S1, S2 row-major (E, q_dim)
  std::vector<uint8_t> check_motions(const double* S1, const double* S2, int64_t E, std::vector<int32_t>* n_checked) {
    cst_->sync_batch_tables();
    pmp_context* ctx = cst_->device_context(nullptr);
    const int D = (int)cst_->q_dim();
    std::vector<uint8_t> out(E);
    if (E == 0) return out;
    double* d1 = dev_.upload_soa(0, S1, E, D);
    double* d2 = dev_.upload_soa(1, S2, E, D);
    uint8_t* dv = (uint8_t*)dev_.buf(2, E);
    int32_t* dn = (int32_t*)dev_.buf(3, sizeof(int32_t) * E);
    pmp_validator_desc v = desc();
    check(pmp_check_edges_batch(ctx, d1, d2, E, E, &v, dv, dn));
    dev_.d2h(out.data(), dv, E);
    if (n_checked) {
      n_checked->resize(E);
      dev_.d2h(n_checked->data(), dn, sizeof(int32_t) * E);
    }
    return out;
  }
  // Validity of every segment of a path (row-major (n, q_dim) states), all
  // n - 1 segments in one batched launch: the check OMPL's path utilities
  // make one checkMotion at a time (ompl_thin_wrap.hpp:130-170).
  std::vector<uint8_t> check_path(const double* P, int64_t n) {
    if (n < 2) return {};
    return check_motions(P, P + cst_->q_dim(), n - 1, nullptr);
  }

  // Batched shortcutting, the planner's path simplification
  // (PathSimplifier::shortcutPath / partialShortcutPath behind simplify(),
  // ompl_thin_wrap.hpp:130-170, 235-248) with every candidate of a round
  // checked in one launch.  A round draws `candidates` state pairs (i, j),
  // j >= i + 2, uniformly from the current path (seeded, so the result is
  // reproducible), checks the direct motions i -> j, and splices in the valid
  // ones that do not overlap, the widest index span first (ties: smaller i).
  // Rounds stop after `rounds` or when a round splices nothing.  The result
  // keeps both end states, is a subsequence of the input, every segment it
  // adds is a valid motion, and it is no longer than the input (each splice
  // replaces a chain by its straight segment).
  std::vector<double> shortcut_path(const double* P, int64_t n, int candidates, int rounds, uint64_t seed,
                                    int64_t* n_out) {
    const int64_t D = (int64_t)cst_->q_dim();
    if (candidates < 0 || rounds < 0) throw std::invalid_argument("candidates and rounds must be non-negative");
    std::vector<double> cur(P, P + n * D);
    std::mt19937_64 rng(seed);
    for (int r = 0; r < rounds; ++r) {
      const int64_t m = (int64_t)cur.size() / D;
      if (m < 3) break;
      std::vector<std::pair<int64_t, int64_t>> cand;
      cand.reserve(candidates);
      for (int c = 0; c < candidates; ++c) {
        const int64_t i = std::uniform_int_distribution<int64_t>(0, m - 3)(rng);
        const int64_t j = std::uniform_int_distribution<int64_t>(i + 2, m - 1)(rng);
        cand.push_back({i, j});
      }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      const int64_t E = (int64_t)cand.size();
      std::vector<double> S1(E * D), S2(E * D);
      for (int64_t e = 0; e < E; ++e) {
        std::copy(&cur[cand[e].first * D], &cur[cand[e].first * D] + D, &S1[e * D]);
        std::copy(&cur[cand[e].second * D], &cur[cand[e].second * D] + D, &S2[e * D]);
      }
      const std::vector<uint8_t> ok = check_motions(S1.data(), S2.data(), E, nullptr);
      std::vector<std::pair<int64_t, int64_t>> valid;
      for (int64_t e = 0; e < E; ++e)
        if (ok[e]) valid.push_back(cand[e]);
      std::stable_sort(valid.begin(), valid.end(), [](const auto& a, const auto& b) {
        return a.second - a.first > b.second - b.first;
      });
      std::vector<std::pair<int64_t, int64_t>> taken;
      for (const auto& v : valid) {
        bool clash = false;
        for (const auto& t : taken) clash = clash || (v.first < t.second && t.first < v.second);
        if (!clash) taken.push_back(v);
      }
      if (taken.empty()) break;
      std::vector<uint8_t> drop(m, 0);
      for (const auto& t : taken)
        for (int64_t k = t.first + 1; k < t.second; ++k) drop[k] = 1;
      std::vector<double> next;
      next.reserve(cur.size());
      for (int64_t k = 0; k < m; ++k)
        if (!drop[k]) next.insert(next.end(), &cur[k * D], &cur[k * D] + D);
      cur.swap(next);
    }
    *n_out = (int64_t)cur.size() / D;
    return cur;
  }

  pmp_validator_desc desc() const {
    pmp_validator_desc v{};
    v.type = vc_.type == ValidatorConfig::Type::BOX ? PMP_VALIDATOR_BOX : PMP_VALIDATOR_EUCLIDEAN;
    v.resolution = vc_.resolution;
    v.box_width = vc_.box_width.empty() ? nullptr : vc_.box_width.data();
    return v;
  }
  std::shared_ptr<SphereCollisionCst> constraint() const { return cst_; }
  // edges of the last batch re-decided in FP64
  int32_t last_recheck_count() {
    int32_t n = 0;
    check(pmp_edge_recheck_count(cst_->device_context(nullptr), &n));
    return n;
  }

 private:
  std::shared_ptr<SphereCollisionCst> cst_;
  ValidatorConfig vc_;
  Device dev_;  // edge staging buffers (the tables live in the constraint's context)
};

}  // namespace pmp_host
