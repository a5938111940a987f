// One damped Gauss-Newton step of the batched multi-start IK, a lane per
// restart (the update of plainmp_amd.ik.solve_ik_batched; the reference's
// solve_ik, src/plainmp/ik.py:42-156, runs SLSQP restart after restart).
//
// Per restart the residual rows are the LinkPoseCst ones
// (constraints/primitive_link_pose.cpp:32-57): position error with the
// point-Jacobian rows of the link origin; with a 6-vector target the RPY
// error getRPY() - rpy (kinematics/transform.hpp:93-120) with
// rpy_derivative rows, with a 7-vector target the quaternion-coefficient
// error with q_derivative rows of the inverse quaternion
// (kinematics/algorithm.cpp:21-47).  The world axis w_j of each joint comes
// from the point Jacobians at the link's three unit axes u_k:
// w_j = 1/2 sum_k u_k x (J(p + u_k) - J(p))_j.  Collision values below the
// margin add penalty rows -sqrt(p) d(value)/dq with residual
// sqrt(p) (margin - value).  The step solves (A^T A + damping I) s = -A^T r
// by Cholesky in FP64, is scaled down to max_step in the max norm and the
// result is projected on [lo, hi].  Everything stays in registers: the
// normal matrix is accumulated row by row.
#include "ik.h"

namespace pmp {

template <int MAXD>
__global__ void __launch_bounds__(256) ik_step_kernel(IkArgs a) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.R) return;
  const int64_t R = a.R;
  const int D = a.D;
  double M[MAXD * (MAXD + 1) / 2];  // lower triangle, row-major
  double g[MAXD], row[MAXD];
#pragma unroll
  for (int i = 0; i < MAXD * (MAXD + 1) / 2; ++i) M[i] = 0.0;
#pragma unroll
  for (int j = 0; j < MAXD; ++j) g[j] = 0.0;
  auto add_row = [&](double res) {
#pragma unroll
    for (int i = 0; i < MAXD; ++i) {
      if (i >= D) break;
      g[i] += row[i] * res;
#pragma unroll
      for (int j = 0; j <= i; ++j) M[i * (i + 1) / 2 + j] += row[i] * row[j];
    }
  };
  double pv[7];
  if (a.pose32) {
#pragma unroll
    for (int k = 0; k < 7; ++k) pv[k] = (double)a.pose32[(int64_t)k * R + r];
  } else {
    const double* P = a.poses + (int64_t)a.link * 7 * R + r;
#pragma unroll
    for (int k = 0; k < 7; ++k) pv[k] = P[(int64_t)k * R];
  }
  const double pos[3] = {pv[0], pv[1], pv[2]};
  const double qx = pv[3], qy = pv[4], qz = pv[5], qw = pv[6];
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int j = 0; j < MAXD; ++j) row[j] = j < D ? (double)a.J0[((int64_t)k * D + j) * R + r] : 0.0;
    add_row(pos[k] - a.tgt[k]);
  }
  if (a.kind > 3) {
    // world axes of the link (columns of its rotation) and of each joint
    const double U[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                            {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                            {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
    double wx[MAXD], wy[MAXD], wz[MAXD];
#pragma unroll
    for (int j = 0; j < MAXD; ++j) {
      wx[j] = wy[j] = wz[j] = 0.0;
      if (j >= D) continue;
      for (int k = 0; k < 3; ++k) {
        const float* Jk = a.Ju[k];
        const double dx = (double)Jk[(0 * (int64_t)D + j) * R + r] - (double)a.J0[(0 * (int64_t)D + j) * R + r];
        const double dy = (double)Jk[(1 * (int64_t)D + j) * R + r] - (double)a.J0[(1 * (int64_t)D + j) * R + r];
        const double dz = (double)Jk[(2 * (int64_t)D + j) * R + r] - (double)a.J0[(2 * (int64_t)D + j) * R + r];
        const double ux = U[0][k], uy = U[1][k], uz = U[2][k];
        wx[j] += 0.5 * (uy * dz - uz * dy);
        wy[j] += 0.5 * (uz * dx - ux * dz);
        wz[j] += 0.5 * (ux * dy - uy * dx);
      }
    }
    if (a.kind == 6) {
      // getRPY (transform.hpp:93-120)
      const double sarg = -2 * (qx * qz - qw * qy);
      const double half_pi = 1.57079632679489661923;
      double roll, pitch, yaw;
      if (sarg <= -0.99999) {
        pitch = -half_pi;
        roll = 0.0;
        yaw = -2 * atan2(qx, qy);
      } else if (sarg >= 0.99999) {
        pitch = half_pi;
        roll = 0.0;
        yaw = 2 * atan2(qx, qy);
      } else {
        pitch = asin(sarg);
        roll = atan2(2 * (qy * qz + qw * qx), qw * qw - qx * qx - qy * qy + qz * qz);
        yaw = atan2(2 * (qx * qy + qw * qz), qw * qw + qx * qx - qy * qy - qz * qz);
      }
      const double rpy[3] = {roll, pitch, yaw};
      double s2, c2, s3, c3;
      sincos(-pitch, &s2, &c2);
      sincos(-yaw, &s3, &c3);
      for (int k = 0; k < 3; ++k) {  // rpy_derivative rows (algorithm.cpp:21-32)
#pragma unroll
        for (int j = 0; j < MAXD; ++j) {
          const double x = wx[j], y = wy[j], z = wz[j];
          row[j] = k == 0 ? c3 / c2 * x - s3 / c2 * y
                          : (k == 1 ? s3 * x + c3 * y : -c3 * s2 / c2 * x + s3 * s2 / c2 * y + z);
        }
        add_row(rpy[k] - a.tgt[3 + k]);
      }
    } else {
      // q_derivative of the inverse quaternion (algorithm.cpp:34-47), coefficients x, y, z, w
      const double quat[4] = {qx, qy, qz, qw};
      const double x = -qx, y = -qy, z = -qz, w = qw;
      for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int j = 0; j < MAXD; ++j) {
          const double ox = wx[j], oy = wy[j], oz = wz[j];
          row[j] = k == 0 ? 0.5 * (oz * y - oy * z + ox * w)
                          : (k == 1 ? 0.5 * (-oz * x + ox * z + oy * w)
                                    : (k == 2 ? 0.5 * (oy * x - ox * y + oz * w) : -0.5 * (-ox * x - oy * y - oz * z)));
        }
        add_row(quat[k] - a.tgt[3 + k]);
      }
    }
  }
  for (int c = 0; c < a.C; ++c) {
    const double v = (double)a.vals[(int64_t)c * R + r];
    if (!(v < a.margin)) continue;
#pragma unroll
    for (int j = 0; j < MAXD; ++j) row[j] = j < D ? -a.wpen * (double)a.cjac[((int64_t)c * D + j) * R + r] : 0.0;
    add_row(a.wpen * (a.margin - v));
  }
  // Cholesky of M + damping I (positive definite), then s = -(L L^T)^-1 g
#pragma unroll
  for (int i = 0; i < MAXD; ++i) M[i * (i + 1) / 2 + i] += a.damping;
#pragma unroll
  for (int i = 0; i < MAXD; ++i) {
    if (i >= D) break;
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      double s = M[i * (i + 1) / 2 + j];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= M[i * (i + 1) / 2 + k] * M[j * (j + 1) / 2 + k];
      M[i * (i + 1) / 2 + j] = i == j ? sqrt(s) : s / M[j * (j + 1) / 2 + j];
    }
  }
  double y[MAXD];
#pragma unroll
  for (int i = 0; i < MAXD; ++i) {  // L y = -g
    if (i >= D) {
      y[i] = 0.0;
      continue;
    }
    double s = -g[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= M[i * (i + 1) / 2 + k] * y[k];
    y[i] = s / M[i * (i + 1) / 2 + i];
  }
  double norm = 0.0;
#pragma unroll
  for (int i = MAXD - 1; i >= 0; --i) {  // L^T s = y
    if (i >= D) continue;
    double s = y[i];
#pragma unroll
    for (int k = i + 1; k < MAXD; ++k)
      if (k < D) s -= M[k * (k + 1) / 2 + i] * y[k];
    y[i] = s / M[i * (i + 1) / 2 + i];
    norm = fmax(norm, fabs(y[i]));
  }
  const double scale = fmin(a.max_step / fmax(norm, 1e-300), 1.0);
#pragma unroll
  for (int j = 0; j < MAXD; ++j) {
    if (j >= D) break;
    double* qj = a.q + (int64_t)j * R + r;
    *qj = fmin(fmax(*qj + y[j] * scale, a.lo[j]), a.hi[j]);
  }
}

hipError_t launch_ik_step(const IkArgs& a, hipStream_t st) {
  if (a.R <= 0) return hipSuccess;
  const dim3 grid((unsigned)((a.R + 255) / 256)), block(256);
  if (a.D <= 8)
    hipLaunchKernelGGL(ik_step_kernel<8>, grid, block, 0, st, a);
  else if (a.D <= 16)
    hipLaunchKernelGGL(ik_step_kernel<16>, grid, block, 0, st, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace pmp
