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

// v from another lane of the same 16-lane row, for controls under which every
// lane has a source (quad permutations, row mirrors): no old value to keep
template <int CTRL>
__device__ __forceinline__ double dpp_row_get(double v) {
  int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, false);
  int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

template <int CTRL, int NV>
__device__ __forceinline__ void dpp_row_step(double (&v)[NV]) {
  double t[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) t[q] = dpp_row_get<CTRL>(v[q]);
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] += t[q];
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

// gfx950 half-wave exchanges.  fold32(a, b): lanes 0-31 get a(l) + a(l + 32),
// lanes 32-63 get b(l - 32) + b(l) (v_permlane32_swap exchanges the upper half
// of a with the lower half of b).  fold16(a, b) does the same one level down
// with v_permlane16_swap, which exchanges the odd rows of a with the even rows
// of b: row 0 gets a's rows 0 + 1, row 1 b's rows 0 + 1, row 2 a's rows 2 + 3,
// row 3 b's rows 2 + 3 (a row is 16 lanes).
__device__ __forceinline__ double fold32(double a, double b) {
  auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(a), __double2loint(b), false, false);
  auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(a), __double2hiint(b), false, false);
  return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}

__device__ __forceinline__ double fold16(double a, double b) {
  auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(a), __double2loint(b), false, false);
  auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(a), __double2hiint(b), false, false);
  return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}

// Transposed wavefront sums of NV values.  Two folds halve the number of
// registers twice (pairs v[j], v[j + H1], then w1[q], w1[q + H2]), so the
// lane moves and adds of the first two levels are spent once per pair
// instead of once per value; the last four levels are DPP moves within
// rows.  On return every lane of row rho of w[q] holds the wavefront total of
// v[q + (rho & 1) * H2 + (rho >> 1) * H1] when q + (rho & 1) * H2 < H1 and
// that index is below NV, and 0 otherwise (row_sum_index).
template <int NV>
struct RowSums {
  static constexpr int H1 = (NV + 1) / 2;
  static constexpr int H2 = (H1 + 1) / 2;
};

// index of the value held by row rho of w[q], or -1 for a padding row
template <int NV>
__device__ __forceinline__ int row_sum_index(int q, int rho) {
  constexpr int H1 = RowSums<NV>::H1, H2 = RowSums<NV>::H2;
  const int j = q + ((rho & 1) ? H2 : 0);
  const int i = j + ((rho & 2) ? H1 : 0);
  return (j < H1 && i < NV) ? i : -1;
}

template <int NV>
__device__ __forceinline__ void wave_sum_rows(const double (&v)[NV], double (&w)[RowSums<NV>::H2]) {
  constexpr int H1 = RowSums<NV>::H1, H2 = RowSums<NV>::H2;
  double w1[H1];
#pragma unroll
  for (int j = 0; j < H1; ++j) w1[j] = fold32(v[j], j + H1 < NV ? v[j + H1] : 0.0);
#pragma unroll
  for (int q = 0; q < H2; ++q) w[q] = fold16(w1[q], q + H2 < H1 ? w1[q + H2] : 0.0);
  dpp_row_step<0xb1>(w);   // quad_perm [1,0,3,2]
  dpp_row_step<0x4e>(w);   // quad_perm [2,3,0,1]
  dpp_row_step<0x141>(w);  // row_half_mirror
  dpp_row_step<0x140>(w);  // row_mirror
}

}  // namespace abcmb
