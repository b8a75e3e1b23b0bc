// Source functions, line-of-sight transfer, C_l k-quadrature, l-spline
// operator and P(k), all on the GPU.
//
//   abcmb_sources  : not-a-knot cubic spline (in log10 k) of the 10 LOS
//                    fields from the PT k-grid to the transfer grid
//                    (abcmb/spectrum.py:664-678), fused with the source
//                    functions S_T0, S_T1, S_T2, S_E (:681-696) and the
//                    trapezoid-in-ln-a weight / aH (:755-772).
//   abcmb_los_cl   : Delta_l(k) = sum_i S(k, i) phi_l(k (tau0 - tau_i)) on the
//                    tabulated Bessel functions (:697-783, lookups :48-109)
//                    and the C_l trapezoid in k (:785-797).  FP64 or FP32.
//   abcmb_los_tangent : the same C_l in FP64 with their derivatives along the
//                    directions of the sources and of the conformal distance
//                    (tangent-linear LOS, exact derivatives of the Bessel
//                    lookups), and along ln A_s and n_s.
//   abcmb_apply_operator : the l-spline (:598-601) as a precomputed linear map.
//   abcmb_pk_lin   : P(k, z=0) (:250-302).
#include <cmath>
#include <cstdint>

#include "common.h"
#include "dual.h"
#include "wave.h"

namespace abcmb {
namespace {

using namespace cst;

// fields of the perturbation table used by the sources, in spline order
__constant__ int LOS_FIELDS[10] = {ABCMB_PT_G_DELTA, ABCMB_PT_B_THETA, ABCMB_PT_THETA_B_PRIME, ABCMB_PT_G_SIGMA,
                                   ABCMB_PT_G_G0,    ABCMB_PT_G_G2,   ABCMB_PT_ETA,           ABCMB_PT_ETA_PRIME,
                                   ABCMB_PT_ALPHA,   ABCMB_PT_ALPHA_PRIME};

// ---- not-a-knot spline: data-independent Thomas factors on log10 k ----
// ws layout: x[nk], h[nk], w[nk] (multipliers), bp[nk] (pivots), cc[nk] (super)
__global__ void spline_factor_kernel(const AbcmbCosmo* __restrict__ cosmos, int n_cosmo, size_t ws_stride,
                                     double* __restrict__ ws_all) {
  int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= n_cosmo) return;
  const double* k = cosmos[ci].k;
  int nk = cosmos[ci].nk;
  double* ws = ws_all + (size_t)ci * ws_stride;
  double* x = ws;
  double* h = x + nk;
  double* w = h + nk;
  double* bp = w + nk;
  double* cc = bp + nk;
  for (int i = 0; i < nk; ++i) x[i] = log10(k[i]);
  for (int i = 0; i < nk - 1; ++i) h[i] = x[i + 1] - x[i];
  int m = nk - 2;
  // rows r = 0..m-1 correspond to M_{r+1}
  for (int r = 0; r < m; ++r) {
    int i = r + 1;
    double a = h[i - 1], b = 2.0 * (h[i - 1] + h[i]), c = h[i];
    if (r == 0) { b = h[0] * (1.0 + h[0] / h[1]) + 2.0 * (h[0] + h[1]); c = h[1] - h[0] * h[0] / h[1]; }
    if (r == m - 1) {
      double h2 = h[nk - 2], h3 = h[nk - 3];
      a = h3 - h2 * h2 / h3;
      b = 2.0 * (h3 + h2) + h2 * (1.0 + h2 / h3);
      if (m == 1) c = 0.0;
    }
    if (r == 0) { w[r] = 0.0; bp[r] = b; }
    else { w[r] = a / bp[r - 1]; bp[r] = b - w[r] * cc[r - 1]; }
    cc[r] = c;
  }
}

// M for one (field, save-point) row.  out M: [10][nk][n_tau]
__global__ void spline_m_kernel(const AbcmbCosmo* __restrict__ cosmos, int n_save, size_t ws_stride,
                                const double* __restrict__ ws_all, size_t m_stride, double* __restrict__ M_all) {
  int n_tau = n_save - 1;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  int f = blockIdx.y;
  int ci = blockIdx.z;
  if (s >= n_tau) return;
  const double* pt = cosmos[ci].pt;
  int nk = cosmos[ci].nk;
  const double* ws = ws_all + (size_t)ci * ws_stride;
  double* Mout = M_all + (size_t)ci * m_stride;
  const double* h = ws + nk;
  const double* w = h + nk;
  const double* bp = w + nk;
  const double* cc = bp + nk;
  const double* y = pt + ((size_t)LOS_FIELDS[f] * n_save + s) * nk;
  double* M = Mout + (size_t)f * nk * n_tau + s;  // M[f][i][s] at M[i * n_tau]
  int m = nk - 2;
  // forward sweep: d'_r stored in M_{r+1}
  double dprev = 0.0;
  for (int r = 0; r < m; ++r) {
    int i = r + 1;
    double d = 6.0 * ((y[i + 1] - y[i]) / h[i] - (y[i] - y[i - 1]) / h[i - 1]);
    if (r > 0) d -= w[r] * dprev;
    M[(size_t)i * n_tau] = d;
    dprev = d;
  }
  double nxt = M[(size_t)m * n_tau] / bp[m - 1];
  M[(size_t)m * n_tau] = nxt;
  for (int r = m - 2; r >= 0; --r) {
    double v = (M[(size_t)(r + 1) * n_tau] - cc[r] * nxt) / bp[r];
    M[(size_t)(r + 1) * n_tau] = v;
    nxt = v;
  }
  double M1 = M[(size_t)1 * n_tau], M2 = M[(size_t)2 * n_tau];
  M[0] = M1 * (1.0 + h[0] / h[1]) - h[0] / h[1] * M2;
  double Ma = M[(size_t)(nk - 2) * n_tau], Mb = M[(size_t)(nk - 3) * n_tau];
  M[(size_t)(nk - 1) * n_tau] = Ma * (1.0 + h[nk - 2] / h[nk - 3]) - h[nk - 2] / h[nk - 3] * Mb;
}

// the ten LOS fields of the table at (kk, saved point s) by the not-a-knot
// cubic spline in log10 k (abcmb/spectrum.py:664-678): the interval lo, the
// weights A, B (values) and cA, cB (second derivatives M)
struct SplineAt {
  int lo;
  double A, B, cA, cB;
};

__device__ __forceinline__ SplineAt spline_at(const double* __restrict__ x, int nk, double kk) {
  const double xq = log10(kk);
  int lo;
  if (xq <= x[0]) lo = 0;
  else if (xq >= x[nk - 1]) lo = nk - 2;
  else {
    int a = 0, b = nk - 1;
    while (b - a > 1) { int mid = (a + b) >> 1; if (x[mid] <= xq) a = mid; else b = mid; }
    lo = a;
  }
  SplineAt w;
  w.lo = lo;
  const double hh = x[lo + 1] - x[lo];
  w.A = (x[lo + 1] - xq) / hh;
  w.B = (xq - x[lo]) / hh;
  w.cA = (w.A * w.A * w.A - w.A) * hh * hh / 6.0;
  w.cB = (w.B * w.B * w.B - w.B) * hh * hh / 6.0;
  return w;
}

__device__ __forceinline__ void los_spline(const AbcmbCosmo& c, const double* __restrict__ ws,
                                           const double* __restrict__ Mb, int n_tau, int s, double kk,
                                           double* v) {
  const int nk = c.nk, n_save = c.n_save;
  const SplineAt w = spline_at(ws, nk, kk);
#pragma unroll
  for (int f = 0; f < 10; ++f) {
    const double* y = c.pt + ((size_t)LOS_FIELDS[f] * n_save + s) * nk;
    const double* M = Mb + (size_t)f * nk * n_tau + s;
    v[f] = w.A * y[w.lo] + w.B * y[w.lo + 1] + w.cA * M[(size_t)w.lo * n_tau] + w.cB * M[(size_t)(w.lo + 1) * n_tau];
  }
}

// the four LOS sources at one (k, saved point) (abcmb/spectrum.py:681-696)
// times the trapezoid-in-ln-a weight over aH (:755-772): v the spline values
// of the LOS fields, (g, g', aH, e^-kappa, aH' aH) the background there, wdl
// the ln a step (half at the first point).  V = double for the forward
// kernel, a dual number for the adjoint's derivatives.
template <typename V>
__device__ __forceinline__ void src_terms(const V* v, const V& g, const V& gp, const V& aH, const V& em,
                                          const V& aHd, double kk, double wdl, double sw, double isw, double dop,
                                          double pol, V* out) {
  const V wq = wdl / aH;
  const V &dg = v[0], &tb = v[1], &tbp = v[2], &sg = v[3], &g0 = v[4], &g2 = v[5], &eta = v[6], &etap = v[7],
          &al = v[8], &alp = v[9];
  const V sT0 = sw * g * (dg / 4. + aH * alp) +
                isw * (g * (eta - aH * alp - 2. * aH * al) + 2. * em * (aH * etap - aHd * al - aH * aH * alp)) +
                dop * (aH * (g * ((tbp / (kk * kk)) + alp) + gp * ((tb / (kk * kk)) + al)));
  const V sT1 = isw * em * ((aH * alp + 2. * aH * al - eta) * kk);
  const V sT2 = pol * g * (2. * sg + g0 + g2) / 8.;
  const V sE = sqrt(6.0) * g * (2. * sg + g0 + g2) / 8.;
  out[0] = wq * sT0;
  out[1] = wq * sT1;
  out[2] = wq * sT2;
  out[3] = wq * sE;
}

// sources at (s, kt).  bgv: [6][n_tau] (tau, g, g', aH, exp(-kappa), aH' aH)
__global__ void sources_kernel(const AbcmbCosmo* __restrict__ cosmos, int n_save, size_t ws_stride,
                               const double* __restrict__ ws_all, size_t m_stride, const double* __restrict__ M_all,
                               const double* __restrict__ bgv_all, const double* __restrict__ kt, int nkt, double sw,
                               double isw, double dop, double pol, double* __restrict__ src_all) {
  int n_tau = n_save - 1;
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int s = blockIdx.y;
  int ci = blockIdx.z;
  if (j >= nkt) return;
  const AbcmbCosmo& c = cosmos[ci];
  const double* ws = ws_all + (size_t)ci * ws_stride;
  const double* Mb = M_all + (size_t)ci * m_stride;
  const double* bgv = bgv_all + (size_t)ci * 6 * n_tau;
  double* src = src_all + (size_t)ci * 4 * n_tau * nkt;
  double dl = c.lna_save[n_save - 1] - c.lna_save[n_save - 2];
  const double kk = kt[j];
  double v[10];
  los_spline(c, ws, Mb, n_tau, s, kk, v);
  const double g = bgv[1 * n_tau + s], gp = bgv[2 * n_tau + s], aH = bgv[3 * n_tau + s], em = bgv[4 * n_tau + s],
               aHd = bgv[5 * n_tau + s];
  double out[4];
  src_terms<double>(v, g, gp, aH, em, aHd, kk, s == 0 ? 0.5 * dl : dl, sw, isw, dop, pol, out);
  size_t o = (size_t)s * nkt + j, plane = (size_t)n_tau * nkt;
  src[0 * plane + o] = out[0];
  src[1 * plane + o] = out[1];
  src[2 * plane + o] = out[2];
  src[3 * plane + o] = out[3];
}

// ---- Bessel lookups (abcmb/spectrum.py:48-109) ----
template <typename T>
__device__ __forceinline__ T fast_interp_t(T x, T lo, T hi, const double* __restrict__ f, int n) {
  T i = (x - lo) / (hi - lo) * (T)(n - 1);
  const T eps = (T)1e-6;
  i = fmin(fmax(i, eps), (T)n - (T)1 - eps);
  int il = (int)floor(i);
  int iu = min(il + 1, n - 1);
  T wu = i - (T)il;
  return ((T)1 - wu) * (T)f[il] + wu * (T)f[iu];
}

// j_l(x) ~ sqrt(pi / 2x) sqrt(2 / (pi r)) cos(Q - pi/4) = cos(Q - pi/4) / sqrt(x r)
template <typename T>
__device__ __forceinline__ T j_asym(T l, T x) {
  const T pi = (T)PI;
  T nu = l + (T)0.5;
  T r = sqrt(x * x - nu * nu);
  T Q = r - nu * pi / (T)2 + nu * asin(nu / x);
  return cos(Q - pi / (T)4) / sqrt(x * r);
}

// phi_0 = j_l, phi_1 = j_l', phi_2 = (3 j_l'' + j_l) / 2 from the asymptotic j_l, j_{l+1}
template <typename T>
__device__ __forceinline__ T phi_asym(int which, T l, T x, T jl, T jl1) {
  if (which == 0) return jl;
  if (which == 1) return l / x * jl - jl1;
  return (((T)3 * l * (l - (T)1) - (T)2 * x * x) * jl + (T)6 * x * jl1) / (T)2 / (x * x);
}

// operand rounding of the contraction: none, or to bfloat16 (round to nearest
// even) for the BF16 variant, whose products are then accumulated in FP32 --
// the numerics of a BF16-input, FP32-accumulate matrix contraction
template <bool BF16>
__device__ __forceinline__ float operand(float x) {
  if constexpr (BF16) {
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
  } else {
    return x;
  }
}
template <bool BF16>
__device__ __forceinline__ double operand(double x) { return x; }

// one block per (k chunk, l): partial C_l sums into part[3][n_ell][n_chunk]
template <typename T, int NTH, bool BF16 = false>
__global__ void __launch_bounds__(NTH) los_cl_kernel(const AbcmbCosmo* __restrict__ cosmos,
                                                     const double* __restrict__ src_all, int n_tau,
                                                     const double* __restrict__ bgv_all,
                                                     const double* __restrict__ kt, int nkt,
                                                     const double* __restrict__ bes, const double* __restrict__ rng,
                                                     const int32_t* __restrict__ bes_l, int n_lt, int nx,
                                                     const int32_t* __restrict__ ell_idx, int n_ell,
                                                     double* __restrict__ part_all, double* __restrict__ delta_all) {
  __shared__ double red[3][NTH / 64];
  int e = blockIdx.y;
  int ci = blockIdx.z;
  int j = blockIdx.x * NTH + threadIdx.x;
  const AbcmbCosmo& cc = cosmos[ci];
  const double tau0 = cc.tau0, A_s = cc.A_s, n_s = cc.n_s, k_pivot = cc.k_pivot;
  const double* src = src_all + (size_t)ci * 4 * n_tau * nkt;
  const double* tau_los = bgv_all + (size_t)ci * 6 * n_tau;  // row 0 of the LOS background vectors
  double* part = part_all + (size_t)ci * 3 * n_ell * gridDim.x;
  double* delta = delta_all ? delta_all + (size_t)ci * 4 * n_ell * nkt : nullptr;
  int li = ell_idx[e];
  T l = (T)bes_l[li];
  const double* t0 = bes + ((size_t)0 * n_lt + li) * nx;
  const double* t1 = bes + ((size_t)1 * n_lt + li) * nx;
  const double* t2 = bes + ((size_t)2 * n_lt + li) * nx;
  T x0lo = (T)rng[(0 * n_lt + li) * 2], x0hi = (T)rng[(0 * n_lt + li) * 2 + 1];
  T x1lo = (T)rng[(1 * n_lt + li) * 2], x1hi = (T)rng[(1 * n_lt + li) * 2 + 1];
  T x2lo = (T)rng[(2 * n_lt + li) * 2], x2hi = (T)rng[(2 * n_lt + li) * 2 + 1];
  T epsf = sqrt((T)3 / (T)8 * (l + (T)2) * (l + (T)1) * l * (l - (T)1));
  T D0 = 0, D1 = 0, D2 = 0, D3 = 0;
  double ctt = 0, cte = 0, cee = 0;
  if (j < nkt) {
    T kk = (T)kt[j];
    size_t plane = (size_t)n_tau * nkt;
    for (int s = 0; s < n_tau; ++s) {
      T x = (T)(tau0 - tau_los[s]) * kk;
      size_t o = (size_t)s * nkt + j;
      // above a table's range the three functions share j_l(x) and j_{l+1}(x)
      const bool a0 = x >= x0hi, a1 = x >= x1hi, a2 = x >= x2hi;
      T jl = (T)0, jl1 = (T)0;
      if (a0 || a1 || a2) jl = j_asym<T>(l, x);
      if (a1 || a2) jl1 = j_asym<T>(l + (T)1, x);
      T p0 = x < x0lo ? (T)0 : a0 ? phi_asym<T>(0, l, x, jl, jl1) : fast_interp_t<T>(x, x0lo, x0hi, t0, nx);
      T p1 = x < x1lo ? (T)0 : a1 ? phi_asym<T>(1, l, x, jl, jl1) : fast_interp_t<T>(x, x1lo, x1hi, t1, nx);
      T p2 = x < x2lo ? (T)0 : a2 ? phi_asym<T>(2, l, x, jl, jl1) : fast_interp_t<T>(x, x2lo, x2hi, t2, nx);
      D0 += operand<BF16>((T)src[0 * plane + o]) * operand<BF16>(p0);
      D1 += operand<BF16>((T)src[1 * plane + o]) * operand<BF16>(p1);
      D2 += operand<BF16>((T)src[2 * plane + o]) * operand<BF16>(p2);
      D3 += operand<BF16>((T)src[3 * plane + o]) * operand<BF16>(p0 / (x * x) * epsf);
    }
    if (delta) {
      delta[((size_t)0 * n_ell + e) * nkt + j] = D0;
      delta[((size_t)1 * n_ell + e) * nkt + j] = D1;
      delta[((size_t)2 * n_ell + e) * nkt + j] = D2;
      delta[((size_t)3 * n_ell + e) * nkt + j] = D3;
    }
    // trapezoid weight of node j
    double kd = kt[j];
    double wj = 0.5 * ((j + 1 < nkt ? kt[j + 1] : kd) - (j > 0 ? kt[j - 1] : kd));
    double P = 4. * PI * A_s * pow(kd / k_pivot, n_s - 1.) / kd;
    double Tt = (double)D0 + (double)D1 + (double)D2, E = (double)D3;
    ctt = wj * P * Tt * Tt;
    cte = wj * P * Tt * E;
    cee = wj * P * E * E;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    ctt += __shfl_xor(ctt, o, 64);
    cte += __shfl_xor(cte, o, 64);
    cee += __shfl_xor(cee, o, 64);
  }
  int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { red[0][wv] = ctt; red[1][wv] = cte; red[2][wv] = cee; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0, c = 0;
    for (int i = 0; i < NTH / 64; ++i) { a += red[0][i]; b += red[1][i]; c += red[2][i]; }
    int nch = gridDim.x;
    part[((size_t)0 * n_ell + e) * nch + blockIdx.x] = a;
    part[((size_t)1 * n_ell + e) * nch + blockIdx.x] = b;
    part[((size_t)2 * n_ell + e) * nch + blockIdx.x] = c;
  }
}

// ---- tangent-linear LOS ----
// The Bessel lookups are differentiated as the reference evaluates them: the
// slope of the interpolated table segment inside a table's range (the
// derivative of the linear interpolant, which finite differences approach as
// the step shrinks) and the derivative of the asymptotic form above it.
//
// j_l(x) of the asymptotic form and its x-derivative: with nu = l + 1/2,
// r = sqrt(x^2 - nu^2), Q = r - nu pi/2 + nu asin(nu/x): dQ/dx = r/x, dr/dx = x/r
__device__ __forceinline__ void sincos_t(double a, double* sn, double* c) { sincos(a, sn, c); }
__device__ __forceinline__ void sincos_t(float a, float* sn, float* c) { sincosf(a, sn, c); }

template <typename T>
__device__ __forceinline__ void j_asym_d(T l, T x, T& j, T& dj) {
  const T pi = (T)PI;
  constexpr T SQRT_2_PI = (T)0.79788456080286535588;   // sqrt(2 / pi)
  const T nu = l + (T)0.5;
  const T ix = (T)1 / x;
  const T r = sqrt(x * x - nu * nu);
  const T ir = (T)1 / r;
  const T isr = (T)1 / sqrt(r);
  const T Q = r - nu * pi / (T)2 + nu * asin(nu * ix);
  T c, sn;
  sincos_t(Q - pi / (T)4, &sn, &c);   // one argument reduction for both
  const T J = SQRT_2_PI * isr * c;
  const T pre = sqrt(pi / (T)2 * ix);
  j = pre * J;
  // dJ/dx = sqrt(2/pi) (-c r^(-3/2) / 2 dr/dx - sin(.) r^(-1/2) dQ/dx)
  const T dJ = SQRT_2_PI * isr * ((T)-0.5 * ir * (x * ir) * c - sn * (r * ix));
  dj = (T)-0.5 * ix * j + pre * dJ;
}

// table value and slope of the interpolated segment (zero where the index is clipped)
template <typename T>
__device__ __forceinline__ void table_d(T x, T lo, T hi, const double* __restrict__ f, int n, T& v, T& dv) {
  T i0 = (x - lo) / (hi - lo) * (T)(n - 1);
  const T eps = (T)1e-6;
  T i = fmin(fmax(i0, eps), (T)n - (T)1 - eps);
  int il = (int)floor(i);
  int iu = min(il + 1, n - 1);
  T wu = i - (T)il;
  const T fl = (T)f[il], fu = (T)f[iu];
  v = ((T)1 - wu) * fl + wu * fu;
  dv = (i == i0) ? (fu - fl) * (T)(n - 1) / (hi - lo) : (T)0;
}

// one block per (k chunk, block of LB sampled l, cosmology); a thread owns one
// k and LB multipoles.  part: [B][NS][n_ell][n_chunk] with NS = 3 (1 + ND + 2).
// T = double, or float with the Bessel arguments, lookups and asymptotic forms in
// FP32 and the contraction over tau accumulated in FP32 (BF16: its operands, the
// sources, their derivatives and the Bessel values and slopes, rounded to
// bfloat16 first), as the value-only kernel's reduced-precision variants
template <typename T, bool BF16, int NTH, int LB, int ND>
__global__ void __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(2))) los_tangent_kernel(
    const AbcmbCosmo* __restrict__ cosmos, const double* __restrict__ src_all,
    const double* __restrict__ dsrc_all, const double* __restrict__ dchi_all, int n_tau,
    const double* __restrict__ bgv_all, const double* __restrict__ kt, int nkt, const double* __restrict__ bes,
    const double* __restrict__ rng, const int32_t* __restrict__ bes_l, int n_lt, int nx,
    const int32_t* __restrict__ ell_idx, int n_ell, double* __restrict__ part_all) {
  constexpr int NS = 3 * (1 + ND + 2);
  constexpr int NW = NTH / 64;
  __shared__ double red[LB * NS][NW];
  const int ci = blockIdx.z;
  const int e0 = blockIdx.y * LB;
  const int j = blockIdx.x * NTH + threadIdx.x;
  const AbcmbCosmo& cc = cosmos[ci];
  const double tau0 = cc.tau0;
  const size_t plane = (size_t)n_tau * nkt;
  const double* src = src_all + (size_t)ci * 4 * plane;
  const double* dsrc = dsrc_all + (size_t)ci * ND * 4 * plane;
  const double* dchi = dchi_all + (size_t)ci * ND * n_tau;
  const double* tau_los = bgv_all + (size_t)ci * 6 * n_tau;
  const int nch = gridDim.x;
  double* part = part_all + (size_t)ci * NS * n_ell * nch;

  T lv[LB], eps_l[LB];
  const double* tb[LB][3];
  T lo[LB][3], hi[LB][3];
#pragma unroll
  for (int b = 0; b < LB; ++b) {
    const int li = ell_idx[min(e0 + b, n_ell - 1)];
    const T l = (T)bes_l[li];
    lv[b] = l;
    eps_l[b] = sqrt((T)3 / (T)8 * (l + (T)2) * (l + (T)1) * l * (l - (T)1));
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      tb[b][w] = bes + ((size_t)w * n_lt + li) * nx;
      lo[b][w] = (T)rng[((size_t)w * n_lt + li) * 2];
      hi[b][w] = (T)rng[((size_t)w * n_lt + li) * 2 + 1];
    }
  }
  T Tv[LB], E[LB], Td[LB][ND], Ed[LB][ND];
#pragma unroll
  for (int b = 0; b < LB; ++b) {
    Tv[b] = E[b] = (T)0;
#pragma unroll
    for (int d = 0; d < ND; ++d) Td[b][d] = Ed[b][d] = (T)0;
  }
  const double kk = j < nkt ? kt[j] : kt[nkt - 1];
  const T kkt = (T)kk;
  const int jj = j < nkt ? j : nkt - 1;
  for (int s = 0; s < n_tau; ++s) {
    const size_t o = (size_t)s * nkt + jj;
    const T x = (T)(tau0 - tau_los[s]) * kkt;
    T S[4], dS[ND][4], dx[ND];
#pragma unroll
    for (int q = 0; q < 4; ++q) S[q] = operand<BF16>((T)src[q * plane + o]);
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      dx[d] = operand<BF16>((T)(kk * dchi[d * n_tau + s]));
#pragma unroll
      for (int q = 0; q < 4; ++q) dS[d][q] = operand<BF16>((T)dsrc[((size_t)d * 4 + q) * plane + o]);
    }
#pragma unroll
    for (int b = 0; b < LB; ++b) {
      const T l = lv[b];
      const bool a0 = x >= hi[b][0], a1 = x >= hi[b][1], a2 = x >= hi[b][2];
      T jl = 0, djl = 0, jl1 = 0, djl1 = 0;
      if (a0 || a1 || a2) j_asym_d<T>(l, x, jl, djl);
      if (a1 || a2) j_asym_d<T>(l + (T)1, x, jl1, djl1);
      T p0 = 0, q0 = 0, p1 = 0, q1 = 0, p2 = 0, q2 = 0;  // phi_w and d phi_w / dx
      if (x >= lo[b][0]) {
        if (a0) { p0 = jl; q0 = djl; }
        else table_d<T>(x, lo[b][0], hi[b][0], tb[b][0], nx, p0, q0);
      }
      if (x >= lo[b][1]) {
        if (a1) { p1 = l / x * jl - jl1; q1 = -l / (x * x) * jl + l / x * djl - djl1; }
        else table_d<T>(x, lo[b][1], hi[b][1], tb[b][1], nx, p1, q1);
      }
      if (x >= lo[b][2]) {
        if (a2) {
          const T c3 = (T)3 * l * (l - (T)1) - (T)2 * x * x;
          const T Nn = c3 * jl + (T)6 * x * jl1;
          const T dN = (T)-4 * x * jl + c3 * djl + (T)6 * jl1 + (T)6 * x * djl1;
          p2 = Nn / (T)2 / (x * x);
          q2 = dN / ((T)2 * x * x) - Nn / (x * x * x);
        } else {
          table_d<T>(x, lo[b][2], hi[b][2], tb[b][2], nx, p2, q2);
        }
      }
      T psi = p0 / (x * x) * eps_l[b];
      T dpsi = eps_l[b] * (q0 / (x * x) - (T)2 * p0 / (x * x * x));
      p0 = operand<BF16>(p0); p1 = operand<BF16>(p1); p2 = operand<BF16>(p2);
      q0 = operand<BF16>(q0); q1 = operand<BF16>(q1); q2 = operand<BF16>(q2);
      psi = operand<BF16>(psi); dpsi = operand<BF16>(dpsi);
      Tv[b] += S[0] * p0 + S[1] * p1 + S[2] * p2;
      E[b] += S[3] * psi;
      const T G = S[0] * q0 + S[1] * q1 + S[2] * q2;
      const T H = S[3] * dpsi;
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        Td[b][d] += dS[d][0] * p0 + dS[d][1] * p1 + dS[d][2] * p2 + G * dx[d];
        Ed[b][d] += dS[d][3] * psi + H * dx[d];
      }
    }
  }
  // trapezoid in k with the primordial spectrum; derivatives by the product rule,
  // d/d ln A_s = the value, d/d n_s = the value times ln(k / k_pivot)
  const double wj = 0.5 * ((jj + 1 < nkt ? kt[jj + 1] : kk) - (jj > 0 ? kt[jj - 1] : kk));
  const double P = j < nkt ? 4. * PI * cc.A_s * pow(kk / cc.k_pivot, cc.n_s - 1.) / kk * wj : 0.0;
  const double lnk = log(kk / cc.k_pivot);
  const int wv = threadIdx.x >> 6;
#pragma unroll
  for (int b = 0; b < LB; ++b) {
    double cb[NS];
    const double tv = Tv[b], ev = E[b];
    cb[0] = P * tv * tv;
    cb[1] = P * tv * ev;
    cb[2] = P * ev * ev;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const double td = Td[b][d], ed = Ed[b][d];
      cb[3 + 3 * d + 0] = P * 2.0 * tv * td;
      cb[3 + 3 * d + 1] = P * (td * ev + tv * ed);
      cb[3 + 3 * d + 2] = P * 2.0 * ev * ed;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      cb[3 * (1 + ND) + q] = cb[q];
      cb[3 * (2 + ND) + q] = cb[q] * lnk;
    }
    wave_sum(cb);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int q = 0; q < NS; ++q) red[b * NS + q][wv] = cb[q];
    }
  }
  __syncthreads();
  for (int v = threadIdx.x; v < LB * NS; v += NTH) {
    const int b = v / NS, q = v % NS;
    if (e0 + b >= n_ell) continue;
    double t = 0.0;
    for (int w = 0; w < NW; ++w) t += red[v][w];
    part[((size_t)q * n_ell + e0 + b) * nch + blockIdx.x] = t;
  }
}

__global__ void sum_parts_kernel(const double* __restrict__ part, int n_rows, int nch, double* __restrict__ out) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  double s = 0.0;
  for (int c = 0; c < nch; ++c) s += part[(size_t)r * nch + c];
  out[r] = s;
}

__global__ void apply_operator_kernel(const double* __restrict__ op, int rows, int cols,
                                      const double* __restrict__ in, int n_vec, double* __restrict__ out) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  int m = blockIdx.y;
  if (r >= rows) return;
  const double* a = op + (size_t)r * cols;
  const double* v = in + (size_t)m * cols;
  double s = 0.0;
  for (int c = 0; c < cols; ++c) s += a[c] * v[c];
  out[(size_t)m * rows + r] = s;
}

__device__ double interp_clamped(double x, const double* __restrict__ xp, const double* __restrict__ fp, int n,
                                 int stride) {
  if (x <= xp[0]) return fp[0];
  if (x >= xp[n - 1]) return fp[(size_t)(n - 1) * stride];
  int a = 0, b = n - 1;
  while (b - a > 1) { int mid = (a + b) >> 1; if (xp[mid] <= x) a = mid; else b = mid; }
  double t = (x - xp[a]) / (xp[a + 1] - xp[a]);
  return fp[(size_t)a * stride] + t * (fp[(size_t)(a + 1) * stride] - fp[(size_t)a * stride]);
}

// Linear power spectrum at ln a = lna (abcmb/spectrum.py:268-347): the density
// contrast is interpolated linearly in ln a over the saved points, then
// linearly in k (both clamped at the ends, as jnp.interp does), squared and
// multiplied by the primordial spectrum.  kind 0: delta_m (Pk_lin); kind 1:
// (omega_b delta_b + omega_cdm delta_cdm) / omega_m (Pk_cb, with the
// reference's omega_m, massive neutrinos included, in the denominator).
__device__ double column_at(const double* __restrict__ f, const double* __restrict__ lna_save, int n_save, int nk,
                            int a, int b, double t, double lna) {
  const double va = interp_clamped(lna, lna_save, f + a, n_save, nk);
  const double vb = interp_clamped(lna, lna_save, f + b, n_save, nk);
  return va + t * (vb - va);
}

__global__ void pk_kernel(const AbcmbCosmo* __restrict__ cosmos, const double* __restrict__ kq, int nq, double lna,
                          int kind, double* __restrict__ out_all) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int ci = blockIdx.y;
  if (i >= nq) return;
  const AbcmbCosmo& cc = cosmos[ci];
  const double* pt = cc.pt;
  const double* k = cc.k;
  const double* lna_save = cc.lna_save;
  int nk = cc.nk, n_save = cc.n_save;
  double A_s = cc.A_s, n_s = cc.n_s, kp = cc.k_pivot;
  double* out = out_all + (size_t)ci * nq;
  const size_t fs = (size_t)n_save * nk;
  double x = kq[i];
  // the two bracketing k columns and the weight between them
  int a, b;
  double t;
  if (x <= k[0]) { a = b = 0; t = 0.0; }
  else if (x >= k[nk - 1]) { a = b = nk - 1; t = 0.0; }
  else {
    a = 0; b = nk - 1;
    while (b - a > 1) { int mid = (a + b) >> 1; if (k[mid] <= x) a = mid; else b = mid; }
    t = (x - k[a]) / (k[b] - k[a]);
  }
  double d;
  if (kind == 0) {
    d = column_at(pt + ABCMB_PT_DELTA_M * fs, lna_save, n_save, nk, a, b, t, lna);
  } else {
    const double dc = column_at(pt + ABCMB_PT_CDM_DELTA * fs, lna_save, n_save, nk, a, b, t, lna);
    const double db = column_at(pt + ABCMB_PT_B_DELTA * fs, lna_save, n_save, nk, a, b, t, lna);
    d = (cc.omega_b * db + cc.omega_cdm * dc) / cc.omega_m;
  }
  out[i] = d * d * (A_s * pow(x / kp, n_s - 1.) * (2 * PI * PI / (x * x * x)));
}


// ---- reverse mode of the sources, the LOS and the C_l quadrature (abcmb_los_adjoint) ----
// From the cotangents cb[3][n_ell] of the sampled C_l (TT, TE, EE):
//   lamd_kernel: per (l, k) the cotangents of Delta_T = D0 + D1 + D2 and
//     Delta_E = D3 from the k trapezoid with the primordial spectrum
//     (C_TT = sum_k w P T^2, C_TE = w P T E, C_EE = w P E^2), and the partial
//     sums of the A_s and n_s cotangents;
//   los_adj_kernel: per (k, saved point) the contraction over l: the sources'
//     cotangents sum_l lamd phi_l(k chi) and the distance's sum_l lamd S
//     phi_l'(k chi) k (the Bessel lookups differentiated as the tangent-linear
//     kernel does), then the sources' cotangents pulled back through the
//     source functions (one-direction duals of src_terms) to the ten spline
//     values (lamv) and to the LOS background at that point (partial sums per
//     block over k, with the distance's);
//   spline_adj_kernel: per (field, saved point) the spline in log10 k
//     transposed: the interpolation weights, the not-a-knot end rows and the
//     tridiagonal solve (its LU factors transposed), into the table cotangent.
template <int NTH>
__global__ void __launch_bounds__(NTH) lamd_kernel(const AbcmbCosmo* __restrict__ cosmos,
                                                   const double* __restrict__ delta_all,
                                                   const double* __restrict__ kt, int nkt,
                                                   const double* __restrict__ cb_all, int n_ell,
                                                   double* __restrict__ lamd_all, double* __restrict__ part_all) {
  __shared__ double red[2][NTH / 64];
  const int e = blockIdx.y, ci = blockIdx.z;
  const int j = blockIdx.x * NTH + threadIdx.x;
  const AbcmbCosmo& cc = cosmos[ci];
  const double* delta = delta_all + (size_t)ci * 4 * n_ell * nkt;
  const double* cb = cb_all + (size_t)ci * 3 * n_ell;
  double* lamd = lamd_all + (size_t)ci * 2 * n_ell * nkt;
  double qa = 0.0, qn = 0.0;
  if (j < nkt) {
    const double Tv = delta[((size_t)0 * n_ell + e) * nkt + j] + delta[((size_t)1 * n_ell + e) * nkt + j] +
                      delta[((size_t)2 * n_ell + e) * nkt + j];
    const double Ev = delta[((size_t)3 * n_ell + e) * nkt + j];
    const double kd = kt[j];
    const double wj = 0.5 * ((j + 1 < nkt ? kt[j + 1] : kd) - (j > 0 ? kt[j - 1] : kd));
    const double P = 4. * PI * cc.A_s * pow(kd / cc.k_pivot, cc.n_s - 1.) / kd;
    const double lt = cb[e], lte = cb[n_ell + e], le = cb[2 * n_ell + e];
    lamd[(size_t)e * nkt + j] = wj * P * (2.0 * Tv * lt + Ev * lte);
    lamd[((size_t)n_ell + e) * nkt + j] = wj * P * (Tv * lte + 2.0 * Ev * le);
    const double q = wj * P * (lt * Tv * Tv + lte * Tv * Ev + le * Ev * Ev);
    qa = q / cc.A_s;
    qn = q * log(kd / cc.k_pivot);
  }
  double v[2] = {qa, qn};
  wave_sum(v);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { red[0][wv] = v[0]; red[1][wv] = v[1]; }
  __syncthreads();
  if (threadIdx.x < 2) {
    double t = 0.0;
    for (int w = 0; w < NTH / 64; ++w) t += red[threadIdx.x][w];
    part_all[(((size_t)ci * 2 + threadIdx.x) * n_ell + e) * gridDim.x + blockIdx.x] = t;
  }
}

template <int NTH>
__global__ void __launch_bounds__(NTH) los_adj_kernel(
    const AbcmbCosmo* __restrict__ cosmos, int n_save, size_t ws_stride, const double* __restrict__ ws_all,
    size_t m_stride, const double* __restrict__ M_all, const double* __restrict__ src_all,
    const double* __restrict__ bgv_all, const double* __restrict__ kt, int nkt, double sw, double isw, double dop,
    double pol, const double* __restrict__ bes, const double* __restrict__ rng, const int32_t* __restrict__ bes_l,
    int n_lt, int nx, const int32_t* __restrict__ ell_idx, int n_ell, const double* __restrict__ lamd_all,
    double* __restrict__ lamv_all, double* __restrict__ part_all) {
  constexpr int NW = NTH / 64;
  __shared__ double red[6][NW];
  const int n_tau = n_save - 1;
  const int s = blockIdx.y, ci = blockIdx.z;
  const int j0 = blockIdx.x * NTH + threadIdx.x;
  const bool live = j0 < nkt;
  const int j = live ? j0 : nkt - 1;
  const AbcmbCosmo& c = cosmos[ci];
  const size_t plane = (size_t)n_tau * nkt;
  const double* src = src_all + (size_t)ci * 4 * plane;
  const double* bgv = bgv_all + (size_t)ci * 6 * n_tau;
  const double* lamd = lamd_all + (size_t)ci * 2 * n_ell * nkt;
  const double kk = kt[j];
  const double x = (c.tau0 - bgv[s]) * kk;
  const size_t o = (size_t)s * nkt + j;
  const double S0 = src[o], S1 = src[plane + o], S2 = src[2 * plane + o], S3 = src[3 * plane + o];
  double ls[4] = {0.0, 0.0, 0.0, 0.0};
  double lx = 0.0;
  for (int e = 0; e < n_ell; ++e) {
    const int li = ell_idx[e];
    const double l = bes_l[li];
    const double eps = sqrt(3. / 8. * (l + 2.) * (l + 1.) * l * (l - 1.));
    double lo[3], hi[3];
    const double* tb[3];
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      tb[w] = bes + ((size_t)w * n_lt + li) * nx;
      lo[w] = rng[((size_t)w * n_lt + li) * 2];
      hi[w] = rng[((size_t)w * n_lt + li) * 2 + 1];
    }
    const bool a0 = x >= hi[0], a1 = x >= hi[1], a2 = x >= hi[2];
    double jl = 0, djl = 0, jl1 = 0, djl1 = 0;
    if (a0 || a1 || a2) j_asym_d<double>(l, x, jl, djl);
    if (a1 || a2) j_asym_d<double>(l + 1.0, x, jl1, djl1);
    double p0 = 0, q0 = 0, p1 = 0, q1 = 0, p2 = 0, q2 = 0;
    if (x >= lo[0]) {
      if (a0) { p0 = jl; q0 = djl; }
      else table_d<double>(x, lo[0], hi[0], tb[0], nx, p0, q0);
    }
    if (x >= lo[1]) {
      if (a1) { p1 = l / x * jl - jl1; q1 = -l / (x * x) * jl + l / x * djl - djl1; }
      else table_d<double>(x, lo[1], hi[1], tb[1], nx, p1, q1);
    }
    if (x >= lo[2]) {
      if (a2) {
        const double c3 = 3. * l * (l - 1.) - 2. * x * x;
        const double Nn = c3 * jl + 6. * x * jl1;
        const double dN = -4. * x * jl + c3 * djl + 6. * jl1 + 6. * x * djl1;
        p2 = Nn / 2. / (x * x);
        q2 = dN / (2. * x * x) - Nn / (x * x * x);
      } else {
        table_d<double>(x, lo[2], hi[2], tb[2], nx, p2, q2);
      }
    }
    const double psi = p0 / (x * x) * eps;
    const double dpsi = eps * (q0 / (x * x) - 2. * p0 / (x * x * x));
    const double lT = lamd[(size_t)e * nkt + j], lE = lamd[((size_t)n_ell + e) * nkt + j];
    ls[0] += lT * p0;
    ls[1] += lT * p1;
    ls[2] += lT * p2;
    ls[3] += lE * psi;
    lx += lT * (S0 * q0 + S1 * q1 + S2 * q2) + lE * S3 * dpsi;
  }
  // the source functions backwards: ten spline values, then the background
  const double* ws = ws_all + (size_t)ci * ws_stride;
  const double* Mb = M_all + (size_t)ci * m_stride;
  double v[10];
  los_spline(c, ws, Mb, n_tau, s, kk, v);
  const double dl = c.lna_save[n_save - 1] - c.lna_save[n_save - 2];
  const double wdl = s == 0 ? 0.5 * dl : dl;
  double bg[5] = {bgv[1 * n_tau + s], bgv[2 * n_tau + s], bgv[3 * n_tau + s], bgv[4 * n_tau + s],
                  bgv[5 * n_tau + s]};
  double* lamv = lamv_all + (size_t)ci * 10 * plane;
  double red_v[6];
#pragma unroll 1
  for (int d = 0; d < 15; ++d) {
    D<1> vd[10], bd[5], out[4];
#pragma unroll
    for (int f = 0; f < 10; ++f) { vd[f] = D<1>(v[f]); vd[f].d[0] = f == d ? 1.0 : 0.0; }
#pragma unroll
    for (int q = 0; q < 5; ++q) { bd[q] = D<1>(bg[q]); bd[q].d[0] = 10 + q == d ? 1.0 : 0.0; }
    src_terms<D<1>>(vd, bd[0], bd[1], bd[2], bd[3], bd[4], kk, wdl, sw, isw, dop, pol, out);
    const double dv = live ? ls[0] * out[0].d[0] + ls[1] * out[1].d[0] + ls[2] * out[2].d[0] + ls[3] * out[3].d[0]
                           : 0.0;
    if (d < 10) {
      if (live) lamv[(size_t)d * plane + o] = dv;
    } else {
      red_v[d - 10] = dv;
    }
  }
  red_v[5] = live ? lx * kk : 0.0;   // the distance chi = tau0 - tau(s)
  wave_sum(red_v);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int q = 0; q < 6; ++q) red[q][wv] = red_v[q];
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double t = 0.0;
    for (int w = 0; w < NW; ++w) t += red[threadIdx.x][w];
    part_all[(((size_t)ci * 6 + threadIdx.x) * n_tau + s) * gridDim.x + blockIdx.x] = t;
  }
}

// one thread per (saved point, field, cosmology)
__global__ void spline_adj_kernel(const AbcmbCosmo* __restrict__ cosmos, int n_save, size_t ws_stride,
                                  const double* __restrict__ ws_all, const double* __restrict__ kt, int nkt,
                                  const double* __restrict__ lamv_all, size_t lm_stride, double* __restrict__ lm_all,
                                  double* const* __restrict__ ptbar) {
  const int n_tau = n_save - 1;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y, ci = blockIdx.z;
  if (s >= n_tau) return;
  const AbcmbCosmo& c = cosmos[ci];
  const int nk = c.nk;
  if (nk < 4) return;   // the not-a-knot rows need two intervals at each end
  const double* ws = ws_all + (size_t)ci * ws_stride;
  const double* x = ws;
  const double* h = x + nk;
  const double* w = h + nk;
  const double* bp = w + nk;
  const double* cc = bp + nk;
  const double* lv = lamv_all + ((size_t)ci * 10 + f) * n_tau * nkt + (size_t)s * nkt;
  double* yb = ptbar[ci] + ((size_t)LOS_FIELDS[f] * n_save + s) * nk;
  double* Lm = lm_all + (size_t)ci * lm_stride + ((size_t)f * n_tau + s) * nk;
  for (int i = 0; i < nk; ++i) Lm[i] = 0.0;
  // the interpolation weights (the interval walks forward with k)
  int lo = 0;
  for (int jj = 0; jj < nkt; ++jj) {
    const double xq = log10(kt[jj]);
    while (lo < nk - 2 && x[lo + 1] <= xq) ++lo;
    const double hh = x[lo + 1] - x[lo];
    const double A = (x[lo + 1] - xq) / hh, B = (xq - x[lo]) / hh;
    const double cA = (A * A * A - A) * hh * hh / 6.0, cB = (B * B * B - B) * hh * hh / 6.0;
    const double l = lv[jj];
    yb[lo] += A * l;
    yb[lo + 1] += B * l;
    Lm[lo] += cA * l;
    Lm[lo + 1] += cB * l;
  }
  // M_0 and M_{nk-1} from the not-a-knot rows onto the interior unknowns z_r = M_{r+1}
  const int m = nk - 2;
  const double e0 = Lm[0], e1 = Lm[nk - 1];
  for (int r = 0; r < m; ++r) Lm[r] = Lm[r + 1];
  Lm[0] += (1.0 + h[0] / h[1]) * e0;
  Lm[1] += -h[0] / h[1] * e0;
  Lm[m - 1] += (1.0 + h[nk - 2] / h[nk - 3]) * e1;
  Lm[m - 2] += -h[nk - 2] / h[nk - 3] * e1;
  // the tridiagonal solve transposed: U^T q = zbar (forward), L^T dbar = q (backward)
  Lm[0] /= bp[0];
  for (int r = 1; r < m; ++r) Lm[r] = (Lm[r] - cc[r - 1] * Lm[r - 1]) / bp[r];
  for (int r = m - 2; r >= 0; --r) Lm[r] -= w[r + 1] * Lm[r + 1];
  // d_r = 6 ((y_{i+1} - y_i) / h_i - (y_i - y_{i-1}) / h_{i-1}), i = r + 1
  for (int r = 0; r < m; ++r) {
    const int i = r + 1;
    const double db = 6.0 * Lm[r];
    yb[i + 1] += db / h[i];
    yb[i] -= db * (1.0 / h[i] + 1.0 / h[i - 1]);
    yb[i - 1] += db / h[i - 1];
  }
}

// bgv_bar[B][6][n_tau] (row 0 the cotangent of tau(s) = -that of chi(s)) and
// scal_bar[B][3] = (A_s, n_s, tau0) from the partial sums
__global__ void los_adj_final_kernel(const double* __restrict__ part6, int nch6, const double* __restrict__ part2,
                                     int nch2, int n_tau, int n_ell, double* __restrict__ bgv_bar,
                                     double* __restrict__ scal_bar) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int ci = blockIdx.y;
  if (idx < 6 * n_tau) {
    const int q = idx / n_tau, s = idx % n_tau;
    const double* p = part6 + (((size_t)ci * 6 + (q == 0 ? 5 : q - 1)) * n_tau + s) * nch6;
    double t = 0.0;
    for (int b = 0; b < nch6; ++b) t += p[b];
    bgv_bar[(size_t)ci * 6 * n_tau + idx] = q == 0 ? -t : t;
  } else if (idx < 6 * n_tau + 3) {
    const int q = idx - 6 * n_tau;
    double t = 0.0;
    if (q < 2) {
      const double* p = part2 + ((size_t)ci * 2 + q) * n_ell * nch2;
      for (int b = 0; b < n_ell * nch2; ++b) t += p[b];
    } else {
      const double* p = part6 + ((size_t)ci * 6 + 5) * n_tau * nch6;
      for (int b = 0; b < n_tau * nch6; ++b) t += p[b];
    }
    scal_bar[(size_t)ci * 3 + q] = t;
  }
}
}  // namespace
}  // namespace abcmb

using namespace abcmb;

static size_t src_ws_stride(int nk_max) { return (size_t)5 * nk_max + 8; }
static size_t src_m_stride(int nk_max, int n_save) { return (size_t)10 * nk_max * (n_save - 1); }

extern "C" size_t abcmb_sources_workspace_size(int n_cosmo, int nk_max, int n_save) {
  return sizeof(double) * (size_t)n_cosmo * (src_ws_stride(nk_max) + src_m_stride(nk_max, n_save));
}

extern "C" int abcmb_sources(const AbcmbCosmo* cosmos_dev, int n_cosmo, int nk_max, int n_save,
                             const double* bgv_dev, const double* kt_dev, int nkt, const double* scales4,
                             double* src_dev, void* workspace, size_t ws_bytes, void* stream) {
  ABCMB_REQUIRE(cosmos_dev && bgv_dev && kt_dev && src_dev && workspace, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && nk_max >= 5 && n_save >= 3 && nkt >= 2, "bad sizes nk_max=%d n_save=%d nkt=%d",
                nk_max, n_save, nkt);
  ABCMB_REQUIRE(ws_bytes >= abcmb_sources_workspace_size(n_cosmo, nk_max, n_save), "workspace too small");
  double sc[4] = {1, 1, 1, 1};
  if (scales4) for (int i = 0; i < 4; ++i) sc[i] = scales4[i];
  hipStream_t s = (hipStream_t)stream;
  size_t wss = src_ws_stride(nk_max), ms = src_m_stride(nk_max, n_save);
  double* ws = (double*)workspace;
  double* Mb = ws + (size_t)n_cosmo * wss;
  int n_tau = n_save - 1;
  hipLaunchKernelGGL(spline_factor_kernel, dim3((n_cosmo + 63) / 64), dim3(64), 0, s, cosmos_dev, n_cosmo, wss, ws);
  hipLaunchKernelGGL(spline_m_kernel, dim3((n_tau + 63) / 64, 10, n_cosmo), dim3(64), 0, s, cosmos_dev, n_save, wss,
                     ws, ms, Mb);
  hipLaunchKernelGGL(sources_kernel, dim3((nkt + 255) / 256, n_tau, n_cosmo), dim3(256), 0, s, cosmos_dev, n_save,
                     wss, ws, ms, Mb, bgv_dev, kt_dev, nkt, sc[0], sc[1], sc[2], sc[3], src_dev);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}

extern "C" size_t abcmb_los_workspace_size(int n_cosmo, int n_ell, int nkt) {
  return sizeof(double) * (size_t)n_cosmo * 3 * n_ell * ((nkt + 255) / 256);
}

extern "C" int abcmb_los_cl(const AbcmbCosmo* cosmos_dev, int n_cosmo, const double* src_dev, int n_tau,
                            const double* bgv_dev, const double* kt_dev, int nkt, const double* bes_dev,
                            const double* bes_rng_dev, const int32_t* bes_l_dev, int n_lt, int nx,
                            const int32_t* ell_idx_dev, int n_ell, int precision, double* cl_dev,
                            double* delta_dev, void* workspace, size_t ws_bytes, void* stream) {
  ABCMB_REQUIRE(cosmos_dev && src_dev && bgv_dev && kt_dev && bes_dev && bes_rng_dev && bes_l_dev && ell_idx_dev &&
                    cl_dev, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && n_tau > 0 && nkt > 1 && n_lt > 0 && nx > 1 && n_ell > 0, "bad sizes");
  ABCMB_REQUIRE(precision >= 0 && precision <= 2, "precision must be 0 (FP64), 1 (FP32) or 2 (BF16 operands)");
  ABCMB_REQUIRE(workspace && ws_bytes >= abcmb_los_workspace_size(n_cosmo, n_ell, nkt), "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  constexpr int NTH = 256;
  int nch = (nkt + NTH - 1) / NTH;
  double* part = (double*)workspace;
  dim3 grid(nch, n_ell, n_cosmo);
  if (precision == 0)
    hipLaunchKernelGGL((los_cl_kernel<double, NTH>), grid, dim3(NTH), 0, s, cosmos_dev, src_dev, n_tau, bgv_dev,
                       kt_dev, nkt, bes_dev, bes_rng_dev, bes_l_dev, n_lt, nx, ell_idx_dev, n_ell, part, delta_dev);
  else if (precision == 1)
    hipLaunchKernelGGL((los_cl_kernel<float, NTH>), grid, dim3(NTH), 0, s, cosmos_dev, src_dev, n_tau, bgv_dev,
                       kt_dev, nkt, bes_dev, bes_rng_dev, bes_l_dev, n_lt, nx, ell_idx_dev, n_ell, part, delta_dev);
  else
    hipLaunchKernelGGL((los_cl_kernel<float, NTH, true>), grid, dim3(NTH), 0, s, cosmos_dev, src_dev, n_tau,
                       bgv_dev, kt_dev, nkt, bes_dev, bes_rng_dev, bes_l_dev, n_lt, nx, ell_idx_dev, n_ell, part,
                       delta_dev);
  hipLaunchKernelGGL(sum_parts_kernel, dim3((3 * n_ell * n_cosmo + 255) / 256), dim3(256), 0, s, part,
                     3 * n_ell * n_cosmo, nch, cl_dev);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}

namespace {
constexpr int LOS_T_NTH = 128;
constexpr int LOS_T_LB = 2;
}

extern "C" size_t abcmb_los_tangent_workspace_size(int n_cosmo, int n_dir, int n_ell, int nkt) {
  return sizeof(double) * (size_t)n_cosmo * 3 * (1 + n_dir + 2) * n_ell * ((nkt + LOS_T_NTH - 1) / LOS_T_NTH);
}

extern "C" int abcmb_los_tangent(const AbcmbCosmo* cosmos_dev, int n_cosmo, const double* src_dev,
                                 const double* dsrc_dev, const double* dchi_dev, int n_dir, int n_tau,
                                 const double* bgv_dev, const double* kt_dev, int nkt, const double* bes_dev,
                                 const double* bes_rng_dev, const int32_t* bes_l_dev, int n_lt, int nx,
                                 const int32_t* ell_idx_dev, int n_ell, int precision, double* cl_dev,
                                 void* workspace, size_t ws_bytes, void* stream) {
  ABCMB_REQUIRE(cosmos_dev && src_dev && dsrc_dev && dchi_dev && bgv_dev && kt_dev && bes_dev && bes_rng_dev &&
                    bes_l_dev && ell_idx_dev && cl_dev, "null argument");
  ABCMB_REQUIRE(precision >= 0 && precision <= 2, "precision must be 0 (FP64), 1 (FP32) or 2 (BF16 operands)");
  ABCMB_REQUIRE(n_dir == ABCMB_TANGENT_DIRS, "n_dir must be %d", ABCMB_TANGENT_DIRS);
  ABCMB_REQUIRE(n_cosmo > 0 && n_tau > 0 && nkt > 1 && n_lt > 0 && nx > 1 && n_ell > 0, "bad sizes");
  ABCMB_REQUIRE(workspace && ws_bytes >= abcmb_los_tangent_workspace_size(n_cosmo, n_dir, n_ell, nkt),
                "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  constexpr int ND = ABCMB_TANGENT_DIRS;
  constexpr int NS = 3 * (1 + ND + 2);
  const int nch = (nkt + LOS_T_NTH - 1) / LOS_T_NTH;
  double* part = (double*)workspace;
  dim3 grid(nch, (n_ell + LOS_T_LB - 1) / LOS_T_LB, n_cosmo);
#define ABCMB_LOS_T(T_, BF_)                                                                                  \
  hipLaunchKernelGGL((los_tangent_kernel<T_, BF_, LOS_T_NTH, LOS_T_LB, ND>), grid, dim3(LOS_T_NTH), 0, s,            \
                     cosmos_dev, src_dev, dsrc_dev, dchi_dev, n_tau, bgv_dev, kt_dev, nkt, bes_dev, bes_rng_dev,      \
                     bes_l_dev, n_lt, nx, ell_idx_dev, n_ell, part)
  if (precision == 0) ABCMB_LOS_T(double, false);
  else if (precision == 1) ABCMB_LOS_T(float, false);
  else ABCMB_LOS_T(float, true);
#undef ABCMB_LOS_T
  hipLaunchKernelGGL(sum_parts_kernel, dim3((NS * n_ell * n_cosmo + 255) / 256), dim3(256), 0, s, part,
                     NS * n_ell * n_cosmo, nch, cl_dev);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}

extern "C" int abcmb_apply_operator(const double* op_dev, int rows, int cols, const double* in_dev, int n_vec,
                                    double* out_dev, void* stream) {
  ABCMB_REQUIRE(op_dev && in_dev && out_dev && rows > 0 && cols > 0 && n_vec > 0, "bad operator arguments");
  hipLaunchKernelGGL(apply_operator_kernel, dim3((rows + 255) / 256, n_vec), dim3(256), 0, (hipStream_t)stream,
                     op_dev, rows, cols, in_dev, n_vec, out_dev);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}

extern "C" int abcmb_pk(const AbcmbCosmo* cosmos_dev, int n_cosmo, const double* kq_dev, int nq, double z, int kind,
                        double* pk_dev, void* stream) {
  ABCMB_REQUIRE(cosmos_dev && kq_dev && pk_dev && n_cosmo > 0 && nq > 0, "bad P(k) arguments");
  ABCMB_REQUIRE(kind == ABCMB_PK_M || kind == ABCMB_PK_CB, "bad P(k) kind %d", kind);
  ABCMB_REQUIRE(z > -1.0, "redshift %g must exceed -1", z);
  hipLaunchKernelGGL(pk_kernel, dim3((nq + 255) / 256, n_cosmo), dim3(256), 0, (hipStream_t)stream, cosmos_dev,
                     kq_dev, nq, -std::log1p(z), kind, pk_dev);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}

extern "C" int abcmb_pk_lin(const AbcmbCosmo* cosmos_dev, int n_cosmo, const double* kq_dev, int nq, double* pk_dev,
                            void* stream) {
  return abcmb_pk(cosmos_dev, n_cosmo, kq_dev, nq, 0.0, ABCMB_PK_M, pk_dev, stream);
}

namespace {
constexpr int LOS_A_NTH = 256;
struct LosAdjParts {
  size_t src_ws, src, delta, lpart, lcl, lamd, part2, lamv, part6, lm, total;
};
LosAdjParts los_adj_parts(int n_cosmo, int nk_max, int n_save, int n_ell, int nkt) {
  LosAdjParts P;
  const size_t B = n_cosmo, n_tau = n_save - 1, nch = (nkt + LOS_A_NTH - 1) / LOS_A_NTH;
  size_t o = 0;
  P.src_ws = o; o += abcmb_sources_workspace_size(n_cosmo, nk_max, n_save) / sizeof(double);
  P.src = o; o += B * 4 * n_tau * nkt;
  P.delta = o; o += B * 4 * n_ell * nkt;
  P.lpart = o; o += abcmb_los_workspace_size(n_cosmo, n_ell, nkt) / sizeof(double);
  P.lcl = o; o += B * 3 * n_ell;
  P.lamd = o; o += B * 2 * n_ell * nkt;
  P.part2 = o; o += B * 2 * n_ell * nch;
  P.lamv = o; o += B * 10 * n_tau * nkt;
  P.part6 = o; o += B * 6 * n_tau * nch;
  P.lm = o; o += B * 10 * n_tau * nk_max;
  P.total = o * sizeof(double);
  return P;
}
}  // namespace

extern "C" size_t abcmb_los_adjoint_workspace_size(int n_cosmo, int nk_max, int n_save, int n_ell, int nkt) {
  if (n_cosmo <= 0 || nk_max <= 0 || n_save < 3 || n_ell <= 0 || nkt < 2) return 0;
  return los_adj_parts(n_cosmo, nk_max, n_save, n_ell, nkt).total;
}

extern "C" int abcmb_los_adjoint(const AbcmbCosmo* cosmos_dev, int n_cosmo, int nk_max, int n_save,
                                 const double* bgv_dev, const double* kt_dev, int nkt, const double* scales4,
                                 const double* bes_dev, const double* bes_rng_dev, const int32_t* bes_l_dev, int n_lt,
                                 int nx, const int32_t* ell_idx_dev, int n_ell, const double* cl_bar_dev,
                                 double* const* pt_bar_dev, double* bgv_bar_dev, double* scal_bar_dev,
                                 void* workspace, size_t ws_bytes, void* stream) {
  ABCMB_REQUIRE(cosmos_dev && bgv_dev && kt_dev && bes_dev && bes_rng_dev && bes_l_dev && ell_idx_dev && cl_bar_dev &&
                    pt_bar_dev && bgv_bar_dev && scal_bar_dev && workspace, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && nk_max >= 5 && n_save >= 3 && nkt >= 2 && n_lt > 0 && nx > 1 && n_ell > 0,
                "bad sizes");
  const LosAdjParts P = los_adj_parts(n_cosmo, nk_max, n_save, n_ell, nkt);
  ABCMB_REQUIRE(ws_bytes >= P.total, "workspace too small (%zu < %zu)", ws_bytes, P.total);
  double sc[4] = {1, 1, 1, 1};
  if (scales4) for (int i = 0; i < 4; ++i) sc[i] = scales4[i];
  hipStream_t s = (hipStream_t)stream;
  double* ws = (double*)workspace;
  const int n_tau = n_save - 1;
  const size_t wss = src_ws_stride(nk_max), ms = src_m_stride(nk_max, n_save);
  double* spl = ws + P.src_ws;
  double* Mb = spl + (size_t)n_cosmo * wss;
  // the forward sources and LOS with the transfer functions kept
  int rc = abcmb_sources(cosmos_dev, n_cosmo, nk_max, n_save, bgv_dev, kt_dev, nkt, sc, ws + P.src, spl,
                         abcmb_sources_workspace_size(n_cosmo, nk_max, n_save), stream);
  if (rc != ABCMB_OK) return rc;
  rc = abcmb_los_cl(cosmos_dev, n_cosmo, ws + P.src, n_tau, bgv_dev, kt_dev, nkt, bes_dev, bes_rng_dev, bes_l_dev,
                    n_lt, nx, ell_idx_dev, n_ell, 0, ws + P.lcl, ws + P.delta, ws + P.lpart,
                    abcmb_los_workspace_size(n_cosmo, n_ell, nkt), stream);
  if (rc != ABCMB_OK) return rc;
  const int nch = (nkt + LOS_A_NTH - 1) / LOS_A_NTH;
  hipLaunchKernelGGL((lamd_kernel<LOS_A_NTH>), dim3(nch, n_ell, n_cosmo), dim3(LOS_A_NTH), 0, s, cosmos_dev,
                     ws + P.delta, kt_dev, nkt, cl_bar_dev, n_ell, ws + P.lamd, ws + P.part2);
  hipLaunchKernelGGL((los_adj_kernel<LOS_A_NTH>), dim3(nch, n_tau, n_cosmo), dim3(LOS_A_NTH), 0, s, cosmos_dev,
                     n_save, wss, spl, ms, Mb, ws + P.src, bgv_dev, kt_dev, nkt, sc[0], sc[1], sc[2], sc[3], bes_dev,
                     bes_rng_dev, bes_l_dev, n_lt, nx, ell_idx_dev, n_ell, ws + P.lamd, ws + P.lamv, ws + P.part6);
  hipLaunchKernelGGL(spline_adj_kernel, dim3((n_tau + 63) / 64, 10, n_cosmo), dim3(64), 0, s, cosmos_dev, n_save, wss,
                     spl, kt_dev, nkt, ws + P.lamv, (size_t)10 * n_tau * nk_max, ws + P.lm, pt_bar_dev);
  hipLaunchKernelGGL(los_adj_final_kernel, dim3((6 * n_tau + 3 + 255) / 256, n_cosmo), dim3(256), 0, s,
                     ws + P.part6, nch, ws + P.part2, nch, n_tau, n_ell, bgv_bar_dev, scal_bar_dev);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}
