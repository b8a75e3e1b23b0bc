// Wavefront-wide sums with DPP lane moves (no LDS traffic), shared by the
// perturbation and line-of-sight kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace abcmb {

// v from another lane by a DPP control (rows outside ROWMASK read 0)
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_get(double v) {
  int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xf, false);
  int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xf, false);
  return __hiloint2double(hi, lo);
}

template <int CTRL, int ROWMASK, int NV>
__device__ __forceinline__ void dpp_step(double (&v)[NV]) {
  double t[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) t[q] = dpp_get<CTRL, ROWMASK>(v[q]);
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] += t[q];
}

// sums over the 64 lanes of a wavefront with DPP moves (no LDS traffic): quad
// sums, 8-lane and 16-lane sums by mirroring within rows, then rows 0+1 and
// 2+3 by row broadcasts and the total in lane 63, read out as a uniform value.
// Step-major over the NV values, so independent moves fill each other's
// wait states.
template <int NV>
__device__ __forceinline__ void wave_sum(double (&v)[NV]) {
  dpp_step<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  dpp_step<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  dpp_step<0x141, 0xf>(v);  // row_half_mirror
  dpp_step<0x140, 0xf>(v);  // row_mirror
  dpp_step<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  dpp_step<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v[q]), 63);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v[q]), 63);
    v[q] = __hiloint2double(hi, lo);
  }
}

}  // namespace abcmb
