// Arguments of the batched IK step (ik.hip).  Internal to the library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmp {

struct IkArgs {
  int64_t R;             // restarts, one lane each
  int32_t D, C, kind;    // q columns (<= 16), collision rows, target length (3, 6 or 7)
  int32_t link;          // node index of the target link
  const double* poses;   // [(node * 7 + k) * R + r] link poses (x, y, z, qx, qy, qz, qw), FP64
  const float* pose32;   // or, when set, [k * R + r] the target link's pose in FP32
  const double* tgt;     // [kind] target
  const float* J0;       // [(k * D + j) * R + r] point Jacobian at the link origin
  const float* Ju[3];    // the same at the link's unit axes (kind > 3)
  const float* vals;     // [c * R + r] collision values (C rows), or null
  const float* cjac;     // [(c * D + j) * R + r] their rows
  const double* lo;      // [D] joint bounds
  const double* hi;
  double damping, max_step, margin, wpen;  // wpen = sqrt(penalty)
  double* q;             // [j * R + r] configurations, updated in place
};

hipError_t launch_ik_step(const IkArgs& a, hipStream_t st);

}  // namespace pmp
