// Host-side mirror of plainmp's hot-path classes above the C ABI.
//
// Names, argument meaning and error types follow the reference:
//   KinematicModel        kinematics/kinematics.hpp:51-169, bindings/kinematics.cpp:30-45
//   Pose, *SDF            collision/primitive_sdf.hpp:30-612, bindings/primitive.cpp:19-111
//   KDTree                collision/kdtree.hpp:28-47
//   SphereCollisionCst    constraints/primitive_sphere_collision.hpp:55-116
//   Box/Euclidean validators  ompl/motion_validator.hpp:42-63, ompl/ompl_thin_wrap.hpp:50-91
// Every computation on a configuration (FK, sphere centres, SDF values,
// validity, Jacobians, nearest points) runs in the gfx950 kernels through the
// C ABI; the host only builds and uploads tables.  A missing or failing GPU
// raises; there is no CPU fallback.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../../include/plainmp_amd.h"
#include "urdf.hpp"

namespace pmp_host {

using Vec3 = std::array<double, 3>;

inline void check(pmp_status s) {
  if (s == PMP_OK) return;
  std::string m = pmp_last_error();
  if (s == PMP_ERR_INVALID) throw std::invalid_argument(m);
  throw std::runtime_error(m);
}
inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// One C-ABI context plus growing device buffers for host-array calls.
class Device {
 public:
  Device() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    check(pmp_create(dev, &ctx_));
  }
  ~Device() {
    for (auto& b : bufs_)
      if (b.first) (void)hipFree(b.first);
    if (ctx_) pmp_destroy(ctx_);
  }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;
  pmp_context* ctx() const { return ctx_; }
  void* buf(int slot, size_t bytes) {
    if ((int)bufs_.size() <= slot) bufs_.resize(slot + 1, {nullptr, 0});
    auto& b = bufs_[slot];
    bytes = std::max<size_t>(bytes, 64);
    if (b.second < bytes) {
      if (b.first) hip_check(hipFree(b.first), "hipFree");
      b.first = nullptr;
      hip_check(hipMalloc(&b.first, bytes), "hipMalloc");
      b.second = bytes;
    }
    return b.first;
  }
  void h2d(void* d, const void* h, size_t n) {
    if (n) hip_check(hipMemcpy(d, h, n, hipMemcpyHostToDevice), "hipMemcpy H2D");
  }
  void d2h(void* h, const void* d, size_t n) {
    if (n) hip_check(hipMemcpy(h, d, n, hipMemcpyDeviceToHost), "hipMemcpy D2H");
  }
  // Column-major upload of a row-major (B, D) host array: dst[d * B + b].
  double* upload_soa(int slot, const double* rows, int64_t B, int D) {
    std::vector<double> t((size_t)B * D);
    for (int64_t b = 0; b < B; ++b)
      for (int d = 0; d < D; ++d) t[(size_t)d * B + b] = rows[(size_t)b * D + d];
    double* p = (double*)buf(slot, t.size() * sizeof(double));
    h2d(p, t.data(), t.size() * sizeof(double));
    return p;
  }

 private:
  pmp_context* ctx_ = nullptr;
  std::vector<std::pair<void*, size_t>> bufs_;
};

// Process-wide context for the free-standing SDF / KD-tree objects.
inline Device& shared_device() {
  static Device* d = new Device();  // lives until process end
  return *d;
}

// ------------------------------------------------------------------ math
struct Quat {
  double x = 0, y = 0, z = 0, w = 1;
};
inline Quat qmul(const Quat& a, const Quat& b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
          a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
// QuatTrans::setQuaternionFromRPY (kinematics/transform.hpp:122-130), no normalisation.
inline Quat quat_from_rpy(double r, double p, double y) {
  double phi = r / 2.0, the = p / 2.0, psi = y / 2.0;
  return {std::sin(phi) * std::cos(the) * std::cos(psi) - std::cos(phi) * std::sin(the) * std::sin(psi),
          std::cos(phi) * std::sin(the) * std::cos(psi) + std::sin(phi) * std::cos(the) * std::sin(psi),
          std::cos(phi) * std::cos(the) * std::sin(psi) - std::sin(phi) * std::sin(the) * std::cos(psi),
          std::cos(phi) * std::cos(the) * std::cos(psi) + std::sin(phi) * std::sin(the) * std::sin(psi)};
}
// Half-angle Taylor sin/cos of the collision check (kinematics/algorithm.cpp:49-86),
// used when a joint that is not a q column was last set without accuracy.
inline void taylor_sincos(double x, double& s, double& c) {
  const double c2 = 1.0 / 2.0, c3 = 1.0 / 6.0, c4 = 1.0 / 24.0, c5 = 1.0 / 120.0, c6 = 1.0 / 720.0,
               c7 = 1.0 / 5040.0, c8 = 1.0 / 40320.0, c9 = 1.0 / 362880.0;
  double cs = 1.0;
  if (x > M_PI * 0.5 || x < -M_PI * 0.5) {
    if (x > M_PI || x < -M_PI) x = x - 2 * M_PI * std::floor(x * (1.0 / (2 * M_PI)) + 0.5);
    if (x < -M_PI * 0.5) {
      x = -x - M_PI;
      cs = -1.0;
    } else if (x > M_PI * 0.5) {
      x = -x + M_PI;
      cs = -1.0;
    }
  }
  double xx = x * x, x4 = xx * xx, x6 = x4 * xx, x8 = x4 * x4;
  s = x * (1 - xx * c3 + x4 * c5 - x6 * c7 + x8 * c9);
  c = cs * (1 - xx * c2 + x4 * c4 - x6 * c6 + x8 * c8);
}

// ----------------------------------------------------------- kinematics
enum BaseType { FIXED = 0, PLANAR = 1, FLOATING = 2 };

struct PackedRobot {
  std::vector<int32_t> parent, kind, q_col;
  std::vector<double> quat, trans, axis;
  std::vector<uint8_t> identity;
};

class KinematicModel {
 public:
  explicit KinematicModel(const std::string& xml) {
    UModel m = parse_urdf(xml);
    if (m.links.empty()) throw std::runtime_error("URDF without links");
    std::map<std::string, UJoint> joints;  // std::map: children in joint-name order (urdfdom initTree)
    for (auto& j : m.joints) joints[j.name] = j;
    std::map<std::string, std::vector<std::string>> child_links, child_joints;
    std::map<std::string, std::string> parent_joint_of;
    for (auto& kv : joints) {
      child_links[kv.second.parent].push_back(kv.second.child);
      child_joints[kv.second.parent].push_back(kv.first);
      if (parent_joint_of.count(kv.second.child))
        throw std::runtime_error("link " + kv.second.child + " has two parent joints");
      parent_joint_of[kv.second.child] = kv.first;
    }
    std::string root;
    std::map<std::string, int> seen;
    for (auto& l : m.links) seen[l] = 0;
    for (auto& kv : joints) {
      if (!seen.count(kv.second.child)) throw std::runtime_error("child link [" + kv.second.child + "] not found");
      if (!seen.count(kv.second.parent)) throw std::runtime_error("parent link [" + kv.second.parent + "] not found");
    }
    for (auto& l : seen)
      if (!parent_joint_of.count(l.first)) {
        if (!root.empty()) throw std::runtime_error("URDF has more than one root link");
        root = l.first;
      }
    if (root.empty()) throw std::runtime_error("URDF has no root link");
    // DFS link numbering (kinematics.cpp:192-206)
    std::vector<std::string> stack{root};
    while (!stack.empty()) {
      std::string n = stack.back();
      stack.pop_back();
      link_id_[n] = (int)link_names_.size();
      link_names_.push_back(n);
      for (auto& c : child_links[n]) stack.push_back(c);
    }
    const size_t N = link_names_.size();
    parent_.assign(N, -1);
    children_.assign(N, {});
    consider_rotation_.assign(N, true);
    origin_quat_.assign(N, Quat{});
    origin_trans_.assign(N, Vec3{0, 0, 0});
    origin_identity_.assign(N, true);
    link_joint_.assign(N, -1);
    for (size_t i = 0; i < N; ++i) {
      auto it = parent_joint_of.find(link_names_[i]);
      if (it == parent_joint_of.end()) continue;
      const UJoint& j = joints[it->second];
      parent_[i] = link_id_[j.parent];
      children_[parent_[i]].push_back((int)i);
      origin_quat_[i] = {j.quat[0], j.quat[1], j.quat[2], j.quat[3]};
      origin_trans_[i] = {j.xyz[0], j.xyz[1], j.xyz[2]};
      origin_identity_[i] = std::abs(j.quat[3] - 1.0) < 1e-6;  // kinematics.cpp:176-178
    }
    // joint numbering: movable joints in DFS order (kinematics.cpp:66-145)
    std::vector<std::string> jstack(child_joints[root].begin(), child_joints[root].end());
    while (!jstack.empty()) {
      std::string jn = jstack.back();
      jstack.pop_back();
      const UJoint& j = joints[jn];
      int type = j.type == "revolute" ? 1 : j.type == "continuous" ? 2 : j.type == "prismatic" ? 3 : 0;
      if (type) {
        int id = (int)joint_type_.size();
        joint_id_[jn] = id;
        joint_type_.push_back(type);
        joint_axis_.push_back({j.axis[0], j.axis[1], j.axis[2]});
        joint_child_.push_back(link_id_[j.child]);
        link_joint_[link_id_[j.child]] = id;
        if (type == 2)
          limits_.push_back({-std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity()});
        else
          limits_.push_back({j.lower, j.upper});
      }
      for (auto& cj : child_joints[j.child]) jstack.push_back(cj);
    }
    all_identity_ = true;
    for (size_t j = 0; j < joint_type_.size(); ++j)
      all_identity_ = all_identity_ && origin_identity_[joint_child_[j]];
    joint_angles_.assign(joint_type_.size(), 0.0);
    joint_mode_.assign(joint_type_.size(), -1);
    base_ = {0, 0, 0, 0, 0, 0, 1};
  }

  size_t n_links() const { return link_names_.size(); }
  size_t n_joints() const { return joint_type_.size(); }
  uint64_t version() const { return version_; }
  const std::string& link_name(size_t i) const { return link_names_[i]; }

  std::vector<size_t> get_joint_ids(const std::vector<std::string>& names) const {
    std::vector<size_t> out;
    for (auto& n : names) {
      auto it = joint_id_.find(n);
      if (it == joint_id_.end()) throw std::invalid_argument("no joint named " + n);
      out.push_back(it->second);
    }
    return out;
  }
  std::vector<size_t> get_link_ids(const std::vector<std::string>& names) const {
    std::vector<size_t> out;
    for (auto& n : names) {
      auto it = link_id_.find(n);
      if (it == link_id_.end()) throw std::invalid_argument("no link named " + n);
      out.push_back(it->second);
    }
    return out;
  }
  std::vector<std::pair<double, double>> get_joint_position_limits(const std::vector<size_t>& ids) const {
    std::vector<std::pair<double, double>> out;
    for (auto i : ids) out.push_back(limits_.at(i));
    return out;
  }
  // set_joint_angles: records angle and sin/cos mode; the transform is built on the GPU.
  void set_joint_angles(const std::vector<size_t>& ids, const std::vector<double>& q, bool accurate) {
    if (q.size() != ids.size()) throw std::invalid_argument("joint ids and positions differ in length");
    for (size_t i = 0; i < ids.size(); ++i) {
      size_t j = ids[i];
      if (j >= joint_angles_.size()) throw std::invalid_argument("joint id out of range");
      joint_angles_[j] = q[i];
      joint_mode_[j] = accurate ? 1 : 0;
    }
    last_mode_ = accurate ? 1 : 0;
    ++version_;
  }
  std::vector<double> get_joint_angles(const std::vector<size_t>& ids) const {
    std::vector<double> out;
    for (auto i : ids) out.push_back(joint_angles_.at(i));
    return out;
  }
  void set_base_pose(const std::array<double, 7>& p) {
    base_ = p;
    ++version_;
  }
  std::array<double, 7> get_base_pose() const { return base_; }

  // add_new_link (kinematics.cpp:278-338), named form.
  size_t add_new_link(const std::string& name, const std::string& parent, const Vec3& pos, const Vec3& rpy,
                      bool consider_rotation) {
    size_t pid = get_link_ids({parent})[0];
    if (link_id_.count(name)) throw std::runtime_error("link name : " + name + " already exists");
    size_t id = link_names_.size();
    link_id_[name] = (int)id;
    link_names_.push_back(name);
    parent_.push_back((int)pid);
    children_[pid].push_back((int)id);
    children_.push_back({});
    consider_rotation_.push_back(consider_rotation);
    origin_quat_.push_back(quat_from_rpy(rpy[0], rpy[1], rpy[2]));
    origin_trans_.push_back(pos);
    origin_identity_.push_back(false);  // QuatTrans default flag
    link_joint_.push_back(-1);
    ++version_;
    return id;
  }

  // Node tables for pmp_upload_robot.  Joints listed in q_joints become q
  // columns in that order; every other movable joint is packed as a static
  // transform holding its current angle in the mode it was last set with,
  // or its origin transform if it was never set, as the reference's
  // tf_plink_to_hlink_cache_ holds it (kinematics.cpp:147-181, algorithm.cpp:109-158).
  void pack(const std::vector<size_t>& q_joints, PackedRobot& out) const {
    const size_t N = link_names_.size();
    std::vector<int> col_of(joint_type_.size(), -1);
    for (size_t c = 0; c < q_joints.size(); ++c) col_of.at(q_joints[c]) = (int)c;
    out = PackedRobot{};
    for (size_t i = 0; i < N; ++i) {
      int kind = PMP_NODE_FIXED, qc = -1;
      Quat q = origin_quat_[i];
      Vec3 t = origin_trans_[i];
      Vec3 ax{0, 0, 0};
      bool ident = origin_identity_[i];
      const int j = link_joint_[i];
      if (i == 0) {
        kind = PMP_NODE_ROOT;
      } else if (!consider_rotation_[i]) {
        kind = PMP_NODE_TRANSLATION;
      } else if (j >= 0) {
        ax = joint_axis_[j];
        const bool prismatic = joint_type_[j] == 3;
        if (col_of[j] >= 0) {
          kind = prismatic ? PMP_NODE_PRISMATIC : PMP_NODE_REVOLUTE;
          qc = col_of[j];
          ident = all_identity_ || origin_identity_[i];
        } else if (joint_mode_[j] >= 0) {
          const double a = joint_angles_[j];
          if (prismatic) {
            t = {t[0] + ax[0] * a, t[1] + ax[1] * a, t[2] + ax[2] * a};
            q = Quat{};
            ident = true;
          } else {
            double s, c;
            if (joint_mode_[j])
              s = std::sin(a * 0.5), c = std::cos(a * 0.5);
            else
              taylor_sincos(a * 0.5, s, c);
            Quat jq{s * ax[0], s * ax[1], s * ax[2], c};
            q = (all_identity_ || origin_identity_[i]) ? jq : qmul(origin_quat_[i], jq);
            ident = false;
          }
        }
      }
      out.parent.push_back(parent_[i]);
      out.kind.push_back(kind);
      out.q_col.push_back(qc);
      out.quat.insert(out.quat.end(), {q.x, q.y, q.z, q.w});
      out.trans.insert(out.trans.end(), t.begin(), t.end());
      out.axis.insert(out.axis.end(), ax.begin(), ax.end());
      out.identity.push_back(ident ? 1 : 0);
    }
  }
  // What pack() reads besides the static tree: the non-column joints' angles
  // and modes, the base pose (fixed bases) and the link count.
  std::vector<double> pack_signature(const std::vector<size_t>& q_joints, bool with_base) const {
    std::vector<bool> is_col(joint_type_.size(), false);
    for (auto j : q_joints) is_col.at(j) = true;
    std::vector<double> sig{(double)link_names_.size()};
    for (size_t j = 0; j < joint_type_.size(); ++j)
      if (!is_col[j]) {
        sig.push_back(joint_angles_[j]);
        sig.push_back(joint_mode_[j]);
      }
    if (with_base) sig.insert(sig.end(), base_.begin(), base_.end());
    return sig;
  }
  int last_mode() const { return last_mode_; }
  const std::vector<int8_t>& joint_modes() const { return joint_mode_; }

 private:
  std::vector<std::string> link_names_;
  std::unordered_map<std::string, int> link_id_, joint_id_;
  std::vector<int> parent_;
  std::vector<std::vector<int>> children_;
  std::vector<bool> consider_rotation_;
  std::vector<Quat> origin_quat_;
  std::vector<Vec3> origin_trans_;
  std::vector<bool> origin_identity_;
  std::vector<int> link_joint_;
  std::vector<int> joint_type_;
  std::vector<Vec3> joint_axis_;
  std::vector<int> joint_child_;
  std::vector<std::pair<double, double>> limits_;
  bool all_identity_ = true;
  std::vector<double> joint_angles_;
  std::vector<int8_t> joint_mode_;
  int last_mode_ = 1;
  std::array<double, 7> base_;
  uint64_t version_ = 0;
};

// ------------------------------------------------------------------- SDFs
using Mat3 = std::array<double, 9>;  // row-major

inline Mat3 mat_identity() { return {1, 0, 0, 0, 1, 0, 0, 0, 1}; }
inline Mat3 mat_mul(const Mat3& A, const Mat3& B) {
  Mat3 C{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
  return C;
}
inline Mat3 mat_inverse(const Mat3& A) {
  auto m = [&](int i, int j) { return A[i * 3 + j]; };
  auto cof = [&](int i, int j) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return m(i1, j1) * m(i2, j2) - m(i1, j2) * m(i2, j1);
  };
  double det = cof(0, 0) * m(0, 0) + cof(1, 0) * m(1, 0) + cof(2, 0) * m(2, 0);
  double inv = 1.0 / det;
  Mat3 R;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i * 3 + j] = cof(j, i) * inv;
  return R;
}

// Pose (collision/primitive_sdf.hpp:30-75)
struct Pose {
  Vec3 position{0, 0, 0};
  Mat3 rot = mat_identity(), rot_inv = mat_identity();
  bool axis_aligned = true, z_axis_aligned = true;
  Pose() = default;
  Pose(const Vec3& p, const Mat3& R) : position(p), rot(R), rot_inv(mat_inverse(R)) { update_flags(); }
  void translate(const Vec3& t) {
    for (int k = 0; k < 3; ++k) position[k] += t[k];
  }
  void rotate_z(double a) {  // orientation only; position unchanged
    double c = std::cos(a), s = std::sin(a);
    Mat3 rz{c, -s, 0, s, c, 0, 0, 0, 1};
    rot = mat_mul(rz, rot);
    rot_inv = mat_inverse(rot);
    update_flags();
  }
  void update_flags() {
    double d2 = 0, n2 = 0;  // Eigen isApprox(Identity), precision 1e-12
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double e = rot[i * 3 + j] - (i == j ? 1.0 : 0.0);
        d2 += e * e;
        n2 += rot[i * 3 + j] * rot[i * 3 + j];
      }
    axis_aligned = d2 <= 1e-24 * std::min(n2, 3.0);
    if (axis_aligned) {
      z_axis_aligned = true;
    } else {
      const double tol = 1e-6;
      z_axis_aligned = std::abs(rot[8] - 1.0) < tol && std::abs(rot[2]) < tol && std::abs(rot[5]) < tol &&
                       std::abs(rot[6]) < tol && std::abs(rot[7]) < tol;
    }
  }
};

enum class SDFType { UNION = 0, BOX = 1, CYLINDER = 2, SPHERE = 3, GROUND = 4, CLOUD = 5 };

struct PrimitiveSDFBase;

struct SDFBase {
  using Ptr = std::shared_ptr<SDFBase>;
  virtual ~SDFBase() = default;
  virtual SDFType get_type() const = 0;
  virtual Ptr clone() const = 0;
  virtual void translate(const Vec3& t) = 0;
  virtual void rotate_z(double a) = 0;
  // Leaf primitives in union order (set_all_sdfs_inner, primitive_sphere_collision.cpp:436-448).
  virtual void flatten(std::vector<std::shared_ptr<PrimitiveSDFBase>>& out) = 0;
  std::vector<double> evaluate_batch(const double* pts3xN, int64_t N);   // points[k * N + i]
  std::vector<uint8_t> is_outside_batch(const double* pts3xN, int64_t N, double r);
  double evaluate(const Vec3& p) { return evaluate_batch(col3(p).data(), 1)[0]; }
  bool is_outside(const Vec3& p, double r) { return is_outside_batch(col3(p).data(), 1, r)[0] != 0; }

 private:
  static std::array<double, 3> col3(const Vec3& p) { return {p[0], p[1], p[2]}; }
};

struct PrimitiveSDFBase : SDFBase, std::enable_shared_from_this<PrimitiveSDFBase> {
  Vec3 lb{0, 0, 0}, ub{0, 0, 0};
  uint64_t version = 0;  // bumped by translate / rotate_z so that constraints re-upload
  virtual void describe(pmp_prim_desc& d) const = 0;
  void flatten(std::vector<std::shared_ptr<PrimitiveSDFBase>>& out) override { out.push_back(shared_from_this()); }
};

struct TransformableSDF : PrimitiveSDFBase {
  Pose pose;
  explicit TransformableSDF(const Pose& p) : pose(p) {}
  void translate(const Vec3& t) override {
    pose.translate(t);
    update_aabb();
    ++version;
  }
  void rotate_z(double a) override {
    pose.rotate_z(a);
    update_aabb();
    ++version;
  }
  virtual void update_aabb() = 0;
  void describe_pose(pmp_prim_desc& d) const {
    d.axis_aligned = pose.axis_aligned;
    d.z_axis_aligned = pose.z_axis_aligned;
    for (int k = 0; k < 3; ++k) {
      d.position[k] = pose.position[k];
      d.lb[k] = lb[k];
      d.ub[k] = ub[k];
    }
    for (int k = 0; k < 9; ++k) {
      d.rotation[k] = pose.rot[k];
      d.rotation_inv[k] = pose.rot_inv[k];
    }
  }
  // AABB of a local box of half extents e (update_aabb, primitive_sdf.hpp:400-421, 500-521)
  void aabb_of_local_box(const Vec3& e) {
    double inf = std::numeric_limits<double>::infinity();
    lb = {inf, inf, inf};
    ub = {-inf, -inf, -inf};
    for (int k = 0; k < 8; ++k) {
      Vec3 v{(k & 1) ? e[0] : -e[0], (k & 2) ? e[1] : -e[1], (k & 4) ? e[2] : -e[2]};
      for (int i = 0; i < 3; ++i) {
        double w = pose.rot[i * 3] * v[0] + pose.rot[i * 3 + 1] * v[1] + pose.rot[i * 3 + 2] * v[2] + pose.position[i];
        lb[i] = std::min(lb[i], w);
        ub[i] = std::max(ub[i], w);
      }
    }
  }
};

struct BoxSDF : TransformableSDF {
  Vec3 width;
  BoxSDF(const Vec3& w, const Pose& p) : TransformableSDF(p), width(w) { update_aabb(); }
  SDFType get_type() const override { return SDFType::BOX; }
  Ptr clone() const override { return std::make_shared<BoxSDF>(width, pose); }
  void update_aabb() override { aabb_of_local_box({width[0] * 0.5, width[1] * 0.5, width[2] * 0.5}); }
  void describe(pmp_prim_desc& d) const override {
    d = pmp_prim_desc{};
    d.type = PMP_PRIM_BOX;
    describe_pose(d);
    for (int k = 0; k < 3; ++k) d.size[k] = width[k];
  }
};

struct CylinderSDF : TransformableSDF {
  double radius, height;
  CylinderSDF(double r, double h, const Pose& p) : TransformableSDF(p), radius(r), height(h) { update_aabb(); }
  SDFType get_type() const override { return SDFType::CYLINDER; }
  Ptr clone() const override { return std::make_shared<CylinderSDF>(radius, height, pose); }
  void update_aabb() override { aabb_of_local_box({radius, radius, 0.5 * height}); }
  void describe(pmp_prim_desc& d) const override {
    d = pmp_prim_desc{};
    d.type = PMP_PRIM_CYLINDER;
    describe_pose(d);
    d.size[0] = radius;
    d.size[1] = height;
  }
};

struct SphereSDF : TransformableSDF {
  double radius;
  SphereSDF(double r, const Pose& p) : TransformableSDF(p), radius(r) { update_aabb(); }
  SDFType get_type() const override { return SDFType::SPHERE; }
  Ptr clone() const override { return std::make_shared<SphereSDF>(radius, pose); }
  void update_aabb() override {  // primitive_sdf.hpp:558-561
    for (int k = 0; k < 3; ++k) {
      lb[k] = pose.position[k] - radius;
      ub[k] = pose.position[k] + radius;
    }
  }
  void describe(pmp_prim_desc& d) const override {
    d = pmp_prim_desc{};
    d.type = PMP_PRIM_SPHERE;
    describe_pose(d);
    d.size[0] = radius;
  }
};

struct GroundSDF : PrimitiveSDFBase {
  double height;
  explicit GroundSDF(double h) : height(h) {  // AABB z in [0, 0] (primitive_sdf.hpp:250-255)
    double inf = std::numeric_limits<double>::infinity();
    lb = {-inf, -inf, 0.0};
    ub = {inf, inf, 0.0};
  }
  SDFType get_type() const override { return SDFType::GROUND; }
  Ptr clone() const override { return std::make_shared<GroundSDF>(height); }
  void translate(const Vec3&) override { throw std::runtime_error("GroundSDF does not support translation"); }
  void rotate_z(double) override { throw std::runtime_error("GroundSDF does not support rotation"); }
  void describe(pmp_prim_desc& d) const override {
    d = pmp_prim_desc{};
    d.type = PMP_PRIM_GROUND;
    Mat3 I = mat_identity();
    for (int k = 0; k < 9; ++k) d.rotation[k] = d.rotation_inv[k] = I[k];
    d.axis_aligned = d.z_axis_aligned = 1;
    d.size[0] = height;
    for (int k = 0; k < 3; ++k) {
      d.lb[k] = lb[k];
      d.ub[k] = ub[k];
    }
  }
};

struct CloudSDF : PrimitiveSDFBase {
  std::vector<double> points;  // xyz interleaved
  double radius;
  CloudSDF(const std::vector<double>& pts, double r) : points(pts), radius(r) {  // primitive_sdf.hpp:571-583
    if (points.empty() || points.size() % 3) throw std::invalid_argument("CloudSDF needs (N, 3) points, N >= 1");
    double inf = std::numeric_limits<double>::infinity();
    lb = {inf, inf, inf};
    ub = {-inf, -inf, -inf};
    for (size_t i = 0; i < points.size(); i += 3)
      for (int k = 0; k < 3; ++k) {
        lb[k] = std::min(lb[k], points[i + k]);
        ub[k] = std::max(ub[k], points[i + k]);
      }
    for (int k = 0; k < 3; ++k) {
      lb[k] -= r;
      ub[k] += r;
    }
  }
  SDFType get_type() const override { return SDFType::CLOUD; }
  Ptr clone() const override { throw std::runtime_error("TODO: Not implemented yet"); }
  void translate(const Vec3&) override { throw std::runtime_error("TODO: Not implemented yet"); }
  void rotate_z(double) override { throw std::runtime_error("TODO: Not implemented yet"); }
  void describe(pmp_prim_desc& d) const override {
    d = pmp_prim_desc{};
    d.type = PMP_PRIM_CLOUD;
    Mat3 I = mat_identity();
    for (int k = 0; k < 9; ++k) d.rotation[k] = d.rotation_inv[k] = I[k];
    d.axis_aligned = d.z_axis_aligned = 1;
    d.size[0] = radius;
    for (int k = 0; k < 3; ++k) {
      d.lb[k] = lb[k];
      d.ub[k] = ub[k];
    }
    d.points = points.data();
    d.n_points = (int64_t)(points.size() / 3);
  }
};

struct UnionSDF : SDFBase {
  std::vector<Ptr> sdfs;
  explicit UnionSDF(const std::vector<Ptr>& s) : sdfs(s) {}
  SDFType get_type() const override { return SDFType::UNION; }
  Ptr clone() const override {  // deep copy (primitive_sdf.hpp:104-109)
    std::vector<Ptr> c;
    for (auto& s : sdfs) c.push_back(s->clone());
    return std::make_shared<UnionSDF>(c);
  }
  void merge(const UnionSDF& other, bool clone) {
    for (auto& s : other.sdfs) sdfs.push_back(clone ? s->clone() : s);
  }
  void add(const Ptr& s, bool clone) { sdfs.push_back(clone ? s->clone() : s); }
  void translate(const Vec3& t) override {
    for (auto& s : sdfs) s->translate(t);
  }
  void rotate_z(double a) override {
    for (auto& s : sdfs) s->rotate_z(a);
  }
  void flatten(std::vector<std::shared_ptr<PrimitiveSDFBase>>& out) override {
    for (auto& s : sdfs) s->flatten(out);
  }
};

inline std::vector<double> run_prims(SDFBase& sdf, const double* pts, int64_t N, int mode, double r) {
  std::vector<std::shared_ptr<PrimitiveSDFBase>> leaves;
  sdf.flatten(leaves);
  std::vector<double> out(N, 0.0);
  if (N == 0) return out;
  if (leaves.empty()) {  // UnionSDF over nothing: evaluate = DBL_MAX, is_outside = true
    std::fill(out.begin(), out.end(), mode == 0 ? std::numeric_limits<double>::max() : 1.0);
    return out;
  }
  std::vector<pmp_prim_desc> d(leaves.size());
  for (size_t i = 0; i < leaves.size(); ++i) leaves[i]->describe(d[i]);
  pmp_scene_desc sc{(int32_t)d.size(), d.data()};
  Device& dev = shared_device();
  double* dp = (double*)dev.buf(0, sizeof(double) * 3 * N);
  double* dout = (double*)dev.buf(1, sizeof(double) * N);
  dev.h2d(dp, pts, sizeof(double) * 3 * N);
  check(pmp_prims_evaluate(dev.ctx(), &sc, dp, N, mode, r, dout));
  dev.d2h(out.data(), dout, sizeof(double) * N);
  return out;
}
inline std::vector<double> SDFBase::evaluate_batch(const double* pts, int64_t N) { return run_prims(*this, pts, N, 0, 0.0); }
inline std::vector<uint8_t> SDFBase::is_outside_batch(const double* pts, int64_t N, double r) {
  auto v = run_prims(*this, pts, N, 1, r);
  std::vector<uint8_t> o(N);
  for (int64_t i = 0; i < N; ++i) o[i] = v[i] != 0.0;
  return o;
}

// KDTree (collision/kdtree.hpp:28-47): exact nearest neighbour by GPU brute force.
class KDTree {
 public:
  explicit KDTree(const std::vector<double>& pts) : n_((int64_t)(pts.size() / 3)) {
    if (n_ == 0 || pts.size() % 3) throw std::invalid_argument("KDTree needs (N, 3) points, N >= 1");
    std::vector<double> soa(3 * n_);
    for (int64_t i = 0; i < n_; ++i)
      for (int k = 0; k < 3; ++k) soa[k * n_ + i] = pts[3 * i + k];
    pts_ = pts;
    hip_check(hipMalloc(&dpts_, sizeof(double) * soa.size()), "hipMalloc");
    hip_check(hipMemcpy(dpts_, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice), "hipMemcpy");
  }
  ~KDTree() {
    if (dpts_) (void)hipFree(dpts_);
  }
  KDTree(const KDTree&) = delete;
  // queries[k * M + i] -> (sqdist, index)
  void nearest(const double* qsoa, int64_t M, std::vector<double>& sq, std::vector<int32_t>& idx) const {
    Device& dev = shared_device();
    double* dq = (double*)dev.buf(2, sizeof(double) * 3 * std::max<int64_t>(M, 1));
    double* ds = (double*)dev.buf(3, sizeof(double) * std::max<int64_t>(M, 1));
    int32_t* di = (int32_t*)dev.buf(4, sizeof(int32_t) * std::max<int64_t>(M, 1));
    dev.h2d(dq, qsoa, sizeof(double) * 3 * M);
    check(pmp_nearest_batch(dev.ctx(), dpts_, n_, dq, M, ds, di));
    sq.resize(M);
    idx.resize(M);
    dev.d2h(sq.data(), ds, sizeof(double) * M);
    dev.d2h(idx.data(), di, sizeof(int32_t) * M);
  }
  Vec3 query(const Vec3& p) const {
    std::vector<double> sq;
    std::vector<int32_t> idx;
    nearest(p.data(), 1, sq, idx);
    return {pts_[3 * idx[0]], pts_[3 * idx[0] + 1], pts_[3 * idx[0] + 2]};
  }
  double sqdist(const Vec3& p) const {
    std::vector<double> sq;
    std::vector<int32_t> idx;
    nearest(p.data(), 1, sq, idx);
    return sq[0];
  }
  const std::vector<double>& points() const { return pts_; }

 private:
  int64_t n_;
  std::vector<double> pts_;
  double* dpts_ = nullptr;
};

}  // namespace pmp_host
