// Geometry shared by the gfx950 kernels: rigid transforms as (quaternion,
// translation), the half-angle Taylor sin/cos of the collision check, and the
// primitive signed-distance tests.  Everything is templated on the scalar so
// that the same code runs in FP32 (the throughput path) and FP64 (the in-band
// re-decision path).
//
// Arithmetic order follows the reference where the result is observable:
//   quaternion product / q*v / toRotationMatrix   kinematics/transform.hpp:56-74
//                                                 (Eigen Quaternion semantics,
//                                                  no normalisation)
//   Taylor sin/cos                                kinematics/algorithm.cpp:49-86
//   box / cylinder / sphere / ground evaluate and is_outside
//                                                 collision/primitive_sdf.hpp:245-566
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "tables.h"

#define PMP_HD __host__ __device__ __forceinline__

namespace pmp {
// Read-only table access.  The robot, scene and edge tables are never written
// while a kernel runs, so they are read through the constant address space
// (by value, so the address space is kept up to the load): with a
// wave-uniform index the compiler then issues scalar loads into SGPRs instead
// of 64 identical vector loads, even in kernels that store to other buffers.
template <class T>
__device__ __forceinline__ T cload(const T* p, int64_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return ((const __attribute__((address_space(4))) T*)p)[i];
#else
  return p[i];  // host pass: never executed, only type-checked
#endif
}
}  // namespace pmp

namespace pmp {

template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
struct Q4 {
  T x, y, z, w;
};
template <class T>
struct Pose {
  Q4<T> q;
  V3<T> t;
};
template <class T>
struct M3 {  // row-major
  T a00, a01, a02, a10, a11, a12, a20, a21, a22;
};

template <class T>
PMP_HD V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
PMP_HD V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
PMP_HD V3<T> scale(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T>
PMP_HD T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T>
PMP_HD V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Hamilton product, Eigen order.
template <class T>
PMP_HD Q4<T> qmul(Q4<T> a, Q4<T> b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
          a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x,
          a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
// q * v as Eigen evaluates it: t = 2 (u x v); v + w t + u x t.
template <class T>
PMP_HD V3<T> qrot(Q4<T> q, V3<T> v) {
  V3<T> u{q.x, q.y, q.z};
  V3<T> t = cross(u, v);
  t = add(t, t);
  V3<T> c = cross(u, t);
  return {v.x + q.w * t.x + c.x, v.y + q.w * t.y + c.y, v.z + q.w * t.z + c.z};
}
// Quaternion::toRotationMatrix without normalisation.
template <class T>
PMP_HD M3<T> qmat(Q4<T> q) {
  T tx = T(2) * q.x, ty = T(2) * q.y, tz = T(2) * q.z;
  T twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
  T txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
  T tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
  return {T(1) - (tyy + tzz), txy - twz, txz + twy,
          txy + twz, T(1) - (txx + tzz), tyz - twx,
          txz - twy, tyz + twx, T(1) - (txx + tyy)};
}
template <class T>
PMP_HD V3<T> mv(const M3<T>& R, V3<T> v) {
  return {R.a00 * v.x + R.a01 * v.y + R.a02 * v.z,
          R.a10 * v.x + R.a11 * v.y + R.a12 * v.z,
          R.a20 * v.x + R.a21 * v.y + R.a22 * v.z};
}
// Transposed product R^T v.
template <class T>
PMP_HD V3<T> mtv(const M3<T>& R, V3<T> v) {
  return {R.a00 * v.x + R.a10 * v.y + R.a20 * v.z,
          R.a01 * v.x + R.a11 * v.y + R.a21 * v.z,
          R.a02 * v.x + R.a12 * v.y + R.a22 * v.z};
}

// Composition parent * local with the identity-sensitive rule of
// QuatTrans::quat_identity_sensitive_mult_and_assign.
template <class T>
PMP_HD Pose<T> compose(const Pose<T>& a, Q4<T> lq, V3<T> lt, bool local_identity) {
  Pose<T> r;
  r.q = local_identity ? a.q : qmul(a.q, lq);
  r.t = add(a.t, qrot(a.q, lt));
  return r;
}

// Half-angle sin/cos by the Taylor series of the collision check.
template <class T>
PMP_HD void taylor_sincos(T x, T& s, T& c) {
  const T c2 = T(1.0 / (1.0 * 2.0));
  const T c3 = T(1.0 / (1.0 * 2.0 * 3.0));
  const T c4 = T(1.0 / (1.0 * 2.0 * 3.0 * 4.0));
  const T c5 = T(1.0 / (1.0 * 2.0 * 3.0 * 4.0 * 5.0));
  const T c6 = T(1.0 / (1.0 * 2.0 * 3.0 * 4.0 * 5.0 * 6.0));
  const T c7 = T(1.0 / (1.0 * 2.0 * 3.0 * 4.0 * 5.0 * 6.0 * 7.0));
  const T c8 = T(1.0 / (1.0 * 2.0 * 3.0 * 4.0 * 5.0 * 6.0 * 7.0 * 8.0));
  const T c9 = T(1.0 / (1.0 * 2.0 * 3.0 * 4.0 * 5.0 * 6.0 * 7.0 * 8.0 * 9.0));
  const T half_pi = T(M_PI * 0.5);
  const T pi = T(M_PI);
  const T inv_2pi = T(1.0 / (2 * M_PI));
  T cos_sign = T(1);
  if (x > half_pi || x < -half_pi) {
    if (x > pi || x < -pi) x = x - T(2 * M_PI) * floor(x * inv_2pi + T(0.5));
    if (x < -half_pi) {
      x = -x - pi;
      cos_sign = T(-1);
    } else if (x > half_pi) {
      x = -x + pi;
      cos_sign = T(-1);
    }
  }
  T xx = x * x, x4 = xx * xx, x6 = x4 * xx, x8 = x4 * x4;
  s = x * (T(1) - xx * c3 + x4 * c5 - x6 * c7 + x8 * c9);
  c = cos_sign * (T(1) - xx * c2 + x4 * c4 - x6 * c6 + x8 * c8);
}

PMP_HD void exact_sincos(float x, float& s, float& c) { sincosf(x, &s, &c); }
PMP_HD void exact_sincos(double x, double& s, double& c) {
  s = sin(x);
  c = cos(x);
}

// QuatTrans::setQuaternionFromRPY (no normalisation).
template <class T>
PMP_HD Q4<T> quat_from_rpy(T r, T p, T y) {
  T sr, cr, sp, cp, sy, cy;
  exact_sincos(r / T(2), sr, cr);
  exact_sincos(p / T(2), sp, cp);
  exact_sincos(y / T(2), sy, cy);
  return {sr * cp * cy - cr * sp * sy, cr * sp * cy + sr * cp * sy, cr * cp * sy - sr * sp * cy,
          cr * cp * cy + sr * sp * sy};
}

template <class T>
PMP_HD bool is_nan3(V3<T> c) {
  return c.x != c.x || c.y != c.y || c.z != c.z;
}

// ---------------------------------------------------------------- primitives
// Device primitive record.  p0..p2 hold the type-specific sizes:
//   box      half widths
//   cylinder radius, half height, radius^2
//   sphere   radius, radius^2
//   ground   height
//   cloud    radius
template <class T>
struct alignas(16) DevPrim {
  int32_t type;
  int32_t flags;        // bit 0 axis_aligned, bit 1 z_axis_aligned
  int32_t cloud_begin;  // first point in the cloud table
  int32_t cloud_count;
  T pos[3];
  T p0, p1, p2;
  T R[9];     // world <- local (rows)
  T Rinv[9];  // general inverse of R, kept like the reference Pose
  T lb[3], ub[3];
};

template <class T>
PMP_HD M3<T> load_m3(const T* r) {
  return {r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8]};
}

template <class T>
PMP_HD bool outside_aabb(const DevPrim<T>& p, V3<T> c, T r) {
  return c.x < p.lb[0] - r || c.x > p.ub[0] + r || c.y < p.lb[1] - r || c.y > p.ub[1] + r ||
         c.z < p.lb[2] - r || c.z > p.ub[2] + r;
}

// Signed distance of a point (PrimitiveSDF evaluate), clouds excluded.
template <class T>
PMP_HD T sdf_eval(const DevPrim<T>& p, V3<T> c) {
  V3<T> pos{p.pos[0], p.pos[1], p.pos[2]};
  switch (p.type) {
    case PMP_PRIM_BOX: {
      V3<T> d = sub(c, pos);
      if (!(p.flags & 1)) d = mv(load_m3(p.Rinv), d);
      T sx = fabs(d.x) - p.p0, sy = fabs(d.y) - p.p1, sz = fabs(d.z) - p.p2;
      T mx = sx > T(0) ? sx : T(0), my = sy > T(0) ? sy : T(0), mz = sz > T(0) ? sz : T(0);
      T outside = sqrt(mx * mx + my * my + mz * mz);
      T ix = sx < T(0) ? sx : T(0), iy = sy < T(0) ? sy : T(0), iz = sz < T(0) ? sz : T(0);
      T inside = ix > iy ? ix : iy;
      inside = inside > iz ? inside : iz;
      return outside + inside;
    }
    case PMP_PRIM_CYLINDER: {
      T zs, xa, ya;
      if (p.flags & 2) {
        zs = fabs(c.z - pos.z) - p.p1;
        xa = fabs(c.x - pos.x);
        ya = fabs(c.y - pos.y);
      } else {
        V3<T> pc = sub(c, pos);
        zs = fabs(pc.x * p.R[2] + pc.y * p.R[5] + pc.z * p.R[8]) - p.p1;
        xa = fabs(pc.x * p.R[0] + pc.y * p.R[3] + pc.z * p.R[6]);
        ya = fabs(pc.x * p.R[1] + pc.y * p.R[4] + pc.z * p.R[7]);
      }
      T rs = sqrt(xa * xa + ya * ya) - p.p0;
      T o0 = rs > T(0) ? rs : T(0), o1 = zs > T(0) ? zs : T(0);
      T outside = sqrt(o0 * o0 + o1 * o1);
      T i0 = rs < T(0) ? rs : T(0), i1 = zs < T(0) ? zs : T(0);
      return outside + (i0 > i1 ? i0 : i1);
    }
    case PMP_PRIM_SPHERE: {
      V3<T> d = sub(c, pos);
      return sqrt(dot(d, d)) - p.p0;
    }
    case PMP_PRIM_GROUND:
      return c.z + p.p0;
  }
  return T(0);
}

// sdf_eval for a primitive of kind K (kKind* of prim_kind): the same
// arithmetic with the type branches resolved at compile time.
template <int K, class T>
PMP_HD T sdf_eval_kind(const DevPrim<T>& p, V3<T> c) {
  if constexpr (K == kKindBoxAA || K == kKindBoxRot) {
    V3<T> d = sub(c, V3<T>{p.pos[0], p.pos[1], p.pos[2]});
    if constexpr (K == kKindBoxRot) d = mv(load_m3(p.Rinv), d);
    T sx = fabs(d.x) - p.p0, sy = fabs(d.y) - p.p1, sz = fabs(d.z) - p.p2;
    T mx = sx > T(0) ? sx : T(0), my = sy > T(0) ? sy : T(0), mz = sz > T(0) ? sz : T(0);
    T outside = sqrt(mx * mx + my * my + mz * mz);
    T ix = sx < T(0) ? sx : T(0), iy = sy < T(0) ? sy : T(0), iz = sz < T(0) ? sz : T(0);
    T inside = ix > iy ? ix : iy;
    inside = inside > iz ? inside : iz;
    return outside + inside;
  } else if constexpr (K == kKindCylZ || K == kKindCylRot) {
    T zs, xa, ya;
    if constexpr (K == kKindCylZ) {
      zs = fabs(c.z - p.pos[2]) - p.p1;
      xa = fabs(c.x - p.pos[0]);
      ya = fabs(c.y - p.pos[1]);
    } else {
      V3<T> pc = sub(c, V3<T>{p.pos[0], p.pos[1], p.pos[2]});
      zs = fabs(pc.x * p.R[2] + pc.y * p.R[5] + pc.z * p.R[8]) - p.p1;
      xa = fabs(pc.x * p.R[0] + pc.y * p.R[3] + pc.z * p.R[6]);
      ya = fabs(pc.x * p.R[1] + pc.y * p.R[4] + pc.z * p.R[7]);
    }
    T rs = sqrt(xa * xa + ya * ya) - p.p0;
    T o0 = rs > T(0) ? rs : T(0), o1 = zs > T(0) ? zs : T(0);
    T outside = sqrt(o0 * o0 + o1 * o1);
    T i0 = rs < T(0) ? rs : T(0), i1 = zs < T(0) ? zs : T(0);
    return outside + (i0 > i1 ? i0 : i1);
  } else if constexpr (K == kKindSphere) {
    V3<T> d = sub(c, V3<T>{p.pos[0], p.pos[1], p.pos[2]});
    return sqrt(dot(d, d)) - p.p0;
  } else {  // ground
    return c.z + p.p0;
  }
}

// Gradient of sdf_eval with respect to the point (clouds excluded).  At kinks
// the branch taken follows the evaluate() formula's active terms.
template <class T>
PMP_HD V3<T> sdf_grad(const DevPrim<T>& p, V3<T> c) {
  V3<T> pos{p.pos[0], p.pos[1], p.pos[2]};
  switch (p.type) {
    case PMP_PRIM_BOX: {
      M3<T> R = load_m3(p.R);
      V3<T> d = sub(c, pos);
      if (!(p.flags & 1)) d = mv(load_m3(p.Rinv), d);
      T sx = fabs(d.x) - p.p0, sy = fabs(d.y) - p.p1, sz = fabs(d.z) - p.p2;
      T gx = d.x < T(0) ? T(-1) : T(1), gy = d.y < T(0) ? T(-1) : T(1), gz = d.z < T(0) ? T(-1) : T(1);
      V3<T> gl;
      if (sx > T(0) || sy > T(0) || sz > T(0)) {
        T mx = sx > T(0) ? sx : T(0), my = sy > T(0) ? sy : T(0), mz = sz > T(0) ? sz : T(0);
        T inv = T(1) / sqrt(mx * mx + my * my + mz * mz);
        gl = {gx * mx * inv, gy * my * inv, gz * mz * inv};
      } else if (sx >= sy && sx >= sz) {
        gl = {gx, T(0), T(0)};
      } else if (sy >= sz) {
        gl = {T(0), gy, T(0)};
      } else {
        gl = {T(0), T(0), gz};
      }
      return (p.flags & 1) ? gl : mv(R, gl);
    }
    case PMP_PRIM_CYLINDER: {
      V3<T> pc = sub(c, pos);
      V3<T> e0, e1, e2;
      if (p.flags & 2) {
        e0 = {T(1), T(0), T(0)};
        e1 = {T(0), T(1), T(0)};
        e2 = {T(0), T(0), T(1)};
      } else {
        e0 = {p.R[0], p.R[3], p.R[6]};
        e1 = {p.R[1], p.R[4], p.R[7]};
        e2 = {p.R[2], p.R[5], p.R[8]};
      }
      T x = dot(pc, e0), y = dot(pc, e1), z = dot(pc, e2);
      T rho = sqrt(x * x + y * y);
      T rs = rho - p.p0, zs = fabs(z) - p.p1;
      V3<T> drho = rho > T(0) ? scale(T(1) / rho, add(scale(x, e0), scale(y, e1))) : V3<T>{T(0), T(0), T(0)};
      V3<T> dz = scale(z < T(0) ? T(-1) : T(1), e2);
      T wr, wz;
      if (rs > T(0) || zs > T(0)) {
        T o0 = rs > T(0) ? rs : T(0), o1 = zs > T(0) ? zs : T(0);
        T inv = T(1) / sqrt(o0 * o0 + o1 * o1);
        wr = o0 * inv;
        wz = o1 * inv;
      } else if (rs > zs) {
        wr = T(1);
        wz = T(0);
      } else {
        wr = T(0);
        wz = T(1);
      }
      return add(scale(wr, drho), scale(wz, dz));
    }
    case PMP_PRIM_SPHERE: {
      V3<T> d = sub(c, pos);
      T n = sqrt(dot(d, d));
      return n > T(0) ? scale(T(1) / n, d) : V3<T>{T(0), T(0), T(0)};
    }
    case PMP_PRIM_GROUND:
      return {T(0), T(0), T(1)};
  }
  return {T(0), T(0), T(0)};
}

// Reference is_outside (primitive_sdf.hpp:313-398, 464-498, 546-552, 246-248)
// in its own branch structure; used in FP64 for the in-band re-decision.
// K >= 0 fixes the primitive kind at compile time (kind-sorted loops).
template <class T, int K = -1>
PMP_HD bool is_outside_ref(const DevPrim<T>& p, V3<T> c, T r) {
  V3<T> pos{p.pos[0], p.pos[1], p.pos[2]};
  const int type = K < 0 ? p.type
                         : (K <= kKindBoxRot ? PMP_PRIM_BOX
                                             : (K <= kKindCylRot ? PMP_PRIM_CYLINDER
                                                                 : (K == kKindSphere ? PMP_PRIM_SPHERE : PMP_PRIM_GROUND)));
  const int flags = K == kKindBoxAA    ? 1
                    : K == kKindBoxRot ? (p.flags & ~1)
                    : K == kKindCylZ   ? 2
                    : K == kKindCylRot ? (p.flags & ~2)
                                       : p.flags;
  switch (type) {
    case PMP_PRIM_BOX: {
      T xs, ys, zs;
      if (flags & 1) {
        xs = fabs(c.x - pos.x) - p.p0;
        if (xs > r) return true;
        ys = fabs(c.y - pos.y) - p.p1;
        if (ys > r) return true;
        zs = fabs(c.z - pos.z) - p.p2;
        if (zs > r) return true;
      } else if (flags & 2) {
        zs = fabs(c.z - pos.z) - p.p2;
        if (zs > r) return true;
        V3<T> pc = sub(c, pos);
        xs = fabs(pc.x * p.R[0] + pc.y * p.R[3] + pc.z * p.R[6]) - p.p0;
        if (xs > r) return true;
        ys = fabs(pc.x * p.R[1] + pc.y * p.R[4] + pc.z * p.R[7]) - p.p1;
        if (ys > r) return true;
      } else {
        V3<T> pc = sub(c, pos);
        xs = fabs(pc.x * p.R[0] + pc.y * p.R[3] + pc.z * p.R[6]) - p.p0;
        if (xs > r) return true;
        ys = fabs(pc.x * p.R[1] + pc.y * p.R[4] + pc.z * p.R[7]) - p.p1;
        if (ys > r) return true;
        zs = fabs(pc.x * p.R[2] + pc.y * p.R[5] + pc.z * p.R[8]) - p.p2;
        if (zs > r) return true;
      }
      if (r < T(1e-6)) return false;
      bool xo = xs > T(0), yo = ys > T(0), zo = zs > T(0);
      int cnt = (int)xo + (int)yo + (int)zo;
      if (cnt < 2) return false;
      T rr = r * r;
      if (cnt == 3) return xs * xs + ys * ys + zs * zs > rr;
      if (!xo) return ys * ys + zs * zs > rr;
      if (!yo) return xs * xs + zs * zs > rr;
      return xs * xs + ys * ys > rr;
    }
    case PMP_PRIM_CYLINDER: {
      T zs, xa, ya;
      if (flags & 2) {
        zs = fabs(c.z - pos.z) - p.p1;
        if (zs > r) return true;
        xa = fabs(c.x - pos.x);
        ya = fabs(c.y - pos.y);
      } else {
        V3<T> pc = sub(c, pos);
        zs = fabs(pc.x * p.R[2] + pc.y * p.R[5] + pc.z * p.R[8]) - p.p1;
        if (zs > r) return true;
        xa = fabs(pc.x * p.R[0] + pc.y * p.R[3] + pc.z * p.R[6]);
        ya = fabs(pc.x * p.R[1] + pc.y * p.R[4] + pc.z * p.R[7]);
      }
      T dsq = xa * xa + ya * ya;
      if (r < T(1e-6)) return dsq > p.p2;
      if (dsq > (p.p0 + r) * (p.p0 + r)) return true;
      bool h_out = zs > T(0), r_out = dsq > p.p2;
      if (h_out && r_out) {
        T rs = sqrt(dsq) - p.p0;
        return zs * zs + rs * rs > r * r;
      }
      return false;
    }
    case PMP_PRIM_SPHERE: {
      V3<T> d = sub(c, pos);
      if (r < T(1e-6)) return dot(d, d) > p.p1;
      return dot(d, d) > (p.p0 + r) * (p.p0 + r);
    }
    case PMP_PRIM_GROUND:
      return c.z + p.p0 > r;
  }
  return false;
}

// FP32 collision test of a sphere (c, r) against a non-cloud primitive of a
// known kind: returns true on collision (the reference's !is_outside_aabb &&
// !is_outside) and raises `inband` when the decision lies within `band`
// metres of flipping.  For box, cylinder and sphere the AABB test is implied
// by the exact test; for the ground it is part of the decision (its AABB is
// z in [0, 0]).  NaN centres are the caller's business (see collide_fast).
template <int K>
PMP_HD bool collide_kind(const DevPrim<float>& p, V3<float> c, float r, float band, bool& inband) {
  const float dx = c.x - p.pos[0], dy = c.y - p.pos[1], dz = c.z - p.pos[2];
  if constexpr (K == kKindBoxAA || K == kKindBoxRot) {
    V3<float> d = {dx, dy, dz};
    if constexpr (K == kKindBoxRot) d = mv(load_m3(p.Rinv), d);
    const float sx = fmaxf(fabsf(d.x) - p.p0, 0.0f);
    const float sy = fmaxf(fabsf(d.y) - p.p1, 0.0f);
    const float sz = fmaxf(fabsf(d.z) - p.p2, 0.0f);
    const float m = sx * sx + sy * sy + sz * sz - r * r;
    inband |= fabsf(m) <= band * (2.0f * r + band);
    return m <= 0.0f;
  } else if constexpr (K == kKindCylZ || K == kKindCylRot) {
    float zs, xa, ya;
    if constexpr (K == kKindCylZ) {
      zs = fabsf(dz) - p.p1;
      xa = dx;
      ya = dy;
    } else {
      zs = fabsf(dx * p.R[2] + dy * p.R[5] + dz * p.R[8]) - p.p1;
      xa = dx * p.R[0] + dy * p.R[3] + dz * p.R[6];
      ya = dx * p.R[1] + dy * p.R[4] + dz * p.R[7];
    }
    const float rs = fmaxf(sqrtf(xa * xa + ya * ya) - p.p0, 0.0f);
    zs = fmaxf(zs, 0.0f);
    const float m = rs * rs + zs * zs - r * r;
    inband |= fabsf(m) <= band * (2.0f * r + band);
    return m <= 0.0f;
  } else if constexpr (K == kKindSphere) {
    const float R = p.p0 + r;
    const float m = dx * dx + dy * dy + dz * dz - R * R;
    inband |= fabsf(m) <= band * (2.0f * R + band);
    return m <= 0.0f;
  } else {  // ground
    const float lo = c.z - (p.lb[2] - r), hi = (p.ub[2] + r) - c.z;  // inside the AABB z-slab iff both >= 0
    const float m = c.z + p.p0 - r;                                    // outside iff m > 0
    inband |= fabsf(lo) <= band || fabsf(hi) <= band || fabsf(m) <= band;
    return lo >= 0.0f && hi >= 0.0f && m <= 0.0f;
  }
}

template <class T>
PMP_HD int prim_kind(const DevPrim<T>& p) {
  switch (p.type) {
    case PMP_PRIM_BOX:
      return (p.flags & 1) ? kKindBoxAA : kKindBoxRot;
    case PMP_PRIM_CYLINDER:
      return (p.flags & 2) ? kKindCylZ : kKindCylRot;
    case PMP_PRIM_SPHERE:
      return kKindSphere;
    case PMP_PRIM_GROUND:
      return kKindGround;
  }
  return kKindCloud;
}

// Any non-cloud primitive.  A NaN centre collides, as every comparison of the
// reference is false for NaN.
PMP_HD bool collide_fast(const DevPrim<float>& p, V3<float> c, float r, float band, bool& inband) {
  if (is_nan3(c)) return true;
  switch (prim_kind(p)) {
    case kKindBoxAA:
      return collide_kind<kKindBoxAA>(p, c, r, band, inband);
    case kKindBoxRot:
      return collide_kind<kKindBoxRot>(p, c, r, band, inband);
    case kKindCylZ:
      return collide_kind<kKindCylZ>(p, c, r, band, inband);
    case kKindCylRot:
      return collide_kind<kKindCylRot>(p, c, r, band, inband);
    case kKindSphere:
      return collide_kind<kKindSphere>(p, c, r, band, inband);
    case kKindGround:
      return collide_kind<kKindGround>(p, c, r, band, inband);
  }
  return false;
}

}  // namespace pmp
