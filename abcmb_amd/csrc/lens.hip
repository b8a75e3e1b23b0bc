// CMB lensing on the GPU.
//
//   abcmb_lensing_clpp : Limber C_l^phiphi (abcmb/spectrum.py:349-435), one
//                        thread per multipole, trapezoid over 500 ln a.
//   abcmb_lens         : correlation-function lensing of TT/TE/EE plus BB
//                        (abcmb/spectrum.py:437-572).  The reduced Wigner
//                        d^l_mn(mu) are generated by the three-term recursion in
//                        l (abcmb/ABCMBTools.py:16-185) in registers, one
//                        Gauss-Legendre node per thread, never stored:
//                          pass 1: Cgl(mu), Cgl2(mu)  (d11, d1-1)
//                          pass 2: xi, xi+, xi-, xi_x (12 d's, X-functions)
//                          pass 3: back-transform sum_mu w xi d (d00, d20,
//                                  d22, d2-2), reduced per block in l order
//                                  (deterministic), then across blocks.
#include <cmath>
#include <cstdint>

#include "common.h"
#include "dual.h"
#include "wave.h"

namespace abcmb {
namespace {

using namespace cst;

// ---- Limber lensing potential ----
// lens_all: [B][4][n_lna] = ln a axis, aH, tau, and chi(lna_rec) in [3][0].
//
// One thread per (l, cosmology); the ln a nodes go through LDS in tiles.  The
// Limber integrand at node i and multipole l is
//   v = chi / aH * W^2 * 9 / (8 pi^2) * Om(z)^2 aH^4 * P(kk) / kk,
//   P(kk) = d^2 A_s (kk / k_p)^(n_s - 1) 2 pi^2 / kk^3,  kk = (l + 1/2) / chi,
// so v = C_i * d_i(kk)^2 * (l + 1/2)^(n_s - 5), with everything but the
// transfer function d in a per-node factor C_i (chi^(5 - n_s) included) and
// the power of (l + 1/2) out of the sum.  Per node the tile also holds 1/chi,
// the ln a step of the trapezoid and the bracket of ln a(z) on the saved-time
// grid; the k bracket of kk walks forward with i (chi falls, kk grows).
constexpr int CLPP_TILE = 256;
constexpr int CLPP_THREADS = 128;
__global__ void __launch_bounds__(CLPP_THREADS) clpp_kernel(const AbcmbCosmo* __restrict__ cosmos,
                                                            const double* __restrict__ lens_all, int n_lna,
                                                            const double* __restrict__ ells, int n_l,
                                                            double* __restrict__ out_all) {
  __shared__ double s_ichi[CLPP_TILE], s_c[CLPP_TILE], s_dl[CLPP_TILE], s_u[CLPP_TILE];
  __shared__ int s_ia[CLPP_TILE], s_ib[CLPP_TILE];
  const int il = blockIdx.x * blockDim.x + threadIdx.x;
  const int ci = blockIdx.y;
  const AbcmbCosmo& c = cosmos[ci];
  const double* lna_ax = lens_all + (size_t)ci * 4 * n_lna;
  const double* aH_ax = lna_ax + n_lna;
  const double* tau_ax = lna_ax + 2 * n_lna;
  const double chi_rec = lna_ax[3 * n_lna];
  const double Om = c.omega_m / (c.h * c.h), OL = c.omega_Lambda / (c.h * c.h);
  const double ns = c.n_s;
  // node factor constants: 9 / (8 pi^2) * A_s * 2 pi^2 * k_p^(1 - n_s)
  const double cnode = 9. / 8. / (PI * PI) * c.A_s * (2 * PI * PI) * pow(c.k_pivot, 1. - ns);
  const int n_save = c.n_save, nk = c.nk;
  const double* lna_save = c.lna_save;
  const double* kgrid = c.k;
  const double* dm = c.pt + (size_t)ABCMB_PT_DELTA_M * n_save * nk;
  const double e5 = (il < n_l ? ells[il] : 0.0) + 0.5;
  double s = 0.0, prev = 0.0;
  int ka = 0;
  for (int i0 = 0; i0 < n_lna; i0 += CLPP_TILE) {
    __syncthreads();
    for (int j = threadIdx.x; j < CLPP_TILE && i0 + j < n_lna; j += blockDim.x) {
      const int i = i0 + j;
      const double lna = lna_ax[i];
      double C = 0.0, ichi = 0.0, u = 0.0;
      int ia = 0, ib = 0;
      if (lna < 0.0) {
        const double chi = c.tau0 - tau_ax[i];
        const double win = (chi_rec - chi) / chi_rec / chi;
        const double a = exp(lna);
        const double z = 1. / a - 1.;
        const double aH = aH_ax[i];
        const double z3 = (1. + z) * (1. + z) * (1. + z);
        const double omz = (Om * z3) / ((Om * z3) + OL);
        const double aH2 = aH * aH;
        C = chi / aH * win * win * omz * omz * (aH2 * aH2) * cnode * pow(chi, 5. - ns);
        ichi = 1.0 / chi;
        const double lq = -log(1. + z);
        if (lq <= lna_save[0]) { ia = ib = 0; }
        else if (lq >= lna_save[n_save - 1]) { ia = ib = n_save - 1; }
        else {
          int lo = 0, hi = n_save - 1;
          while (hi - lo > 1) { int mid = (lo + hi) >> 1; if (lna_save[mid] <= lq) lo = mid; else hi = mid; }
          ia = lo; ib = lo + 1;
          u = (lq - lna_save[lo]) / (lna_save[lo + 1] - lna_save[lo]);
        }
      }
      s_ichi[j] = ichi; s_c[j] = C; s_u[j] = u; s_ia[j] = ia; s_ib[j] = ib;
      s_dl[j] = i > 0 ? lna - lna_ax[i - 1] : 0.0;
    }
    __syncthreads();
    if (il < n_l) {
      const int m = min(CLPP_TILE, n_lna - i0);
      for (int j = 0; j < m; ++j) {
        double v = 0.0;
        const double C = s_c[j];
        if (C != 0.0) {
          const double kk = e5 * s_ichi[j];
          while (ka < nk - 2 && kgrid[ka + 1] <= kk) ++ka;
          int k0, k1;
          double t;
          if (kk <= kgrid[0]) { k0 = k1 = 0; t = 0.0; }
          else if (kk >= kgrid[nk - 1]) { k0 = k1 = nk - 1; t = 0.0; }
          else { k0 = ka; k1 = ka + 1; t = (kk - kgrid[k0]) / (kgrid[k1] - kgrid[k0]); }
          const double u = s_u[j];
          const double* ra = dm + (size_t)s_ia[j] * nk;
          const double* rb = dm + (size_t)s_ib[j] * nk;
          const double da = ra[k0] + u * (rb[k0] - ra[k0]);
          const double db = ra[k1] + u * (rb[k1] - ra[k1]);
          const double d = da + t * (db - da);
          v = C * (d * d);
        }
        if (i0 + j > 0) s += 0.5 * (prev + v) * s_dl[j];
        prev = v;
      }
    }
  }
  if (il < n_l) out_all[(size_t)ci * n_l + il] = 8. * PI * PI / (e5 * e5 * e5) * pow(e5, ns - 5.) * s;
}

// ---- Wigner d recursion tables ----
// pairs: 0 d00, 1 d11, 2 d1-1, 3 d2-2, 4 d22, 5 d31, 6 d40, 7 d3-3, 8 d4-4, 9 d20, 10 d3-1, 11 d4-2
__constant__ int PAIR_M[12] = {0, 1, 1, 2, 2, 3, 4, 3, 4, 2, 3, 4};
__constant__ int PAIR_N[12] = {0, 1, -1, -2, 2, 1, 0, -3, -4, 0, -1, -2};
constexpr int NPAIR = 12;

// coef[q][l][3] = (A, B, C) of d^{l+1} = (A mu + B) d^l + C d^{l-1} in the
// normalised recursion; lfac[l][LFAC] = sqrt(2/(2l+1)), the four X-function
// square roots, 8 / (l(l+1)) and 2 / sqrt(l(l+1)) (0 at l = 0).  l = 0 .. LT-1
// with LT = n_l + 3.
constexpr int LFAC = 7;
__global__ void wigner_coef_kernel(int LT, double* __restrict__ coef, double* __restrict__ lfac) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < NPAIR * LT) {
    int q = idx / LT, li = idx % LT;
    double l = li, m = PAIR_M[q], n = PAIR_N[q];
    double nA = sqrt((2 * l + 3) / (2 * l + 1));
    double nC = sqrt((2 * l + 3) / (2 * l - 1));
    double den = sqrt((l + 1) * (l + 1) - m * m) * sqrt((l + 1) * (l + 1) - n * n);
    double A = nA * (l + 1) * (2 * l + 1) / den;
    double B = -A * m * n / l / (l + 1);
    double C = -nC * sqrt(l * l - m * m) * sqrt(l * l - n * n) / den * (l + 1) / l;
    // the reference maps non-finite coefficients to 0 (nan_to_num)
    coef[(size_t)idx * 3 + 0] = isfinite(A) ? A : 0.0;
    coef[(size_t)idx * 3 + 1] = isfinite(B) ? B : 0.0;
    coef[(size_t)idx * 3 + 2] = isfinite(C) ? C : 0.0;
  }
  if (idx < LT) {
    double l = idx;
    double* lf = lfac + idx * LFAC;
    lf[0] = sqrt(2. / (2. * l + 1.));
    lf[1] = sqrt(fmax((l + 2) * (l - 1) * l * (l + 1), 0.0));
    lf[2] = sqrt(fmax((l + 2) * (l - 1), 0.0));
    lf[3] = sqrt(fmax((l + 3) * (l - 2), 0.0));
    lf[4] = sqrt(fmax((l + 4) * (l + 3) * (l - 2) * (l - 3), 0.0));
    lf[5] = idx > 0 ? 8. / l / (l + 1) : 0.0;
    lf[6] = idx > 0 ? 2. / sqrt(l * (l + 1)) : 0.0;
  }
}

__device__ double fact(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

// d^m_{mn}(mu) in the normalised recursion (times sqrt((2m+1)/2))
__device__ double wigner_start(int m, int n, double mu) {
  double beta = acos(mu);
  double norm = sqrt((2 * m + 1) / 2.0) * sqrt(fact(2 * m) / (fact(m + n) * fact(m - n)));
  double c = cos(beta / 2.), s = -sin(beta / 2.);
  double v = norm;
  for (int i = 0; i < m + n; ++i) v *= c;
  for (int i = 0; i < m - n; ++i) v *= s;
  return v;
}

struct Rec {
  double dl, dlm1;
};

// advance pair q from l to l+1 with its coefficients (A, B, C) at l
__device__ __forceinline__ void rec_step_c(Rec& r, double A, double B, double C, double mu) {
  double dn = A * mu * r.dl + B * r.dl + C * r.dlm1;
  r.dlm1 = r.dl;
  r.dl = dn;
}
__device__ __forceinline__ void rec_step(Rec& r, const double* __restrict__ cq, int l, double mu) {
  rec_step_c(r, cq[l * 3 + 0], cq[l * 3 + 1], cq[l * 3 + 2], mu);
}

// The l loops read the recursion coefficients and l factors of a tile of KT
// multipoles from LDS, staged by the block (one wavefront) before the tile:
// reads at the same address in every lane, issued together, instead of a
// chain of dependent scalar loads per multipole.
constexpr int KT = 16;
struct CoefTile {
  double c[KT][NPAIR][4];   // (A, B, C, pad) of every pair at l = l0 + i
  double f[KT][8];          // lfac at l = l0 + i (LFAC entries, padded)
};

// stage the tile of multipoles l0 .. l0 + KT - 1 (those below LT)
__device__ __forceinline__ void load_tile(CoefTile& T, const double* __restrict__ coef,
                                          const double* __restrict__ lfac, int LT, int l0) {
  __syncthreads();
  for (int idx = threadIdx.x; idx < KT * NPAIR * 3; idx += blockDim.x) {
    const int i = idx / (NPAIR * 3), q = (idx / 3) % NPAIR, e = idx % 3;
    const int l = l0 + i;
    T.c[i][q][e] = l < LT ? coef[((size_t)q * LT + l) * 3 + e] : 0.0;
  }
  for (int idx = threadIdx.x; idx < KT * LFAC; idx += blockDim.x) {
    const int i = idx / LFAC, e = idx % LFAC;
    const int l = l0 + i;
    T.f[i][e] = l < LT ? lfac[(size_t)l * LFAC + e] : 0.0;
  }
  __syncthreads();
}

// pass 1: Cgl(mu), Cgl2(mu)
__global__ void __launch_bounds__(64) cgl_kernel(const double* __restrict__ clpp_all, int n_l, const double* __restrict__ mu_arr, int n_mu,
                           const double* __restrict__ coef, const double* __restrict__ lfac, int LT,
                           double* __restrict__ cgl_all) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int ci = blockIdx.y;
  if (i >= n_mu) return;
  const double* clpp = clpp_all + (size_t)ci * n_l;
  double* cgl = cgl_all + (size_t)ci * 2 * n_mu;
  double mu = mu_arr[i];
  const double* c11 = coef + (size_t)1 * LT * 3;
  const double* c1m = coef + (size_t)2 * LT * 3;
  Rec a{wigner_start(1, 1, mu), 0.0}, b{wigner_start(1, -1, mu), 0.0};
  rec_step(a, c11, 1, mu);
  rec_step(b, c1m, 1, mu);
  double s1 = 0.0, s2 = 0.0;
  for (int j = 0; j < n_l; ++j) {
    int l = j + 2;
    double nf = lfac[l * LFAC];
    double f = (2. * l + 1) * l * (l + 1) * clpp[j];
    s1 += f * (a.dl * nf);
    s2 += f * (b.dl * nf);
    rec_step(a, c11, l, mu);
    rec_step(b, c1m, l, mu);
  }
  cgl[i] = s1 / 4. / PI;
  cgl[n_mu + i] = s2 / 4. / PI;
}

// pass 2: correlation functions
__global__ void __launch_bounds__(64) ksi_kernel(const double* __restrict__ cl_all, const double* __restrict__ cgl_all, int n_l,
                           const double* __restrict__ mu_arr, int n_mu, const double* __restrict__ coef,
                           const double* __restrict__ lfac, int LT, double* __restrict__ ksi_all) {
  __shared__ CoefTile T;
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i0 < n_mu;
  const int i = live ? i0 : n_mu - 1;   // spare lanes follow the last node and store nothing
  int ci = blockIdx.y;
  const double* cl_in = cl_all + (size_t)ci * 3 * n_l;
  const double* cgl = cgl_all + (size_t)ci * 2 * n_mu;
  double* ksi = ksi_all + (size_t)ci * 4 * n_mu;
  const double* TT = cl_in;
  const double* TE = cl_in + n_l;
  const double* EE = cl_in + 2 * n_l;
  double mu = mu_arr[i];
  double s2 = cgl[n_mu - 1] - cgl[i];
  double C2 = cgl[n_mu + i];
  Rec R[NPAIR];
#pragma unroll
  for (int q = 0; q < NPAIR; ++q) R[q] = Rec{0.0, 0.0};
  // d00 from l=0, d1n from l=1, d2n from l=2 are live at l=2; d3n, d4n start later
  R[0] = Rec{wigner_start(0, 0, mu), 0.0};
  rec_step(R[0], coef + 0, 0, mu);
  rec_step(R[0], coef + 0, 1, mu);
  R[1] = Rec{wigner_start(1, 1, mu), 0.0};
  rec_step(R[1], coef + (size_t)1 * LT * 3, 1, mu);
  R[2] = Rec{wigner_start(1, -1, mu), 0.0};
  rec_step(R[2], coef + (size_t)2 * LT * 3, 1, mu);
  R[3] = Rec{wigner_start(2, -2, mu), 0.0};
  R[4] = Rec{wigner_start(2, 2, mu), 0.0};
  R[9] = Rec{wigner_start(2, 0, mu), 0.0};
  double e2 = exp(2. * s2 / 4.), e4 = exp(4. * s2 / 4.), e83 = exp(8. / 3. * s2 / 4.);
  double e203 = exp(20. / 3. * s2 / 4.), e10 = exp(10. * s2 / 4.);
  // E(l) = exp(-l(l+1) s2 / 4) by products: E(l+1) = E(l) r(l) with
  // r(l) = g^(l+1), g = exp(-s2 / 2)
  const double g = exp(-0.5 * s2);
  double E = exp(-1.5 * s2), r = g * g * g;
  double a0 = 0, ap = 0, am = 0, ax = 0;
  for (int j = 0; j < n_l; ++j) {
    int l = j + 2;
    const int it = j % KT;
    if (it == 0) load_tile(T, coef, lfac, LT, l);
    if (l == 3) {
      R[5] = Rec{wigner_start(3, 1, mu), 0.0};
      R[7] = Rec{wigner_start(3, -3, mu), 0.0};
      R[10] = Rec{wigner_start(3, -1, mu), 0.0};
    }
    if (l == 4) {
      R[6] = Rec{wigner_start(4, 0, mu), 0.0};
      R[8] = Rec{wigner_start(4, -4, mu), 0.0};
      R[11] = Rec{wigner_start(4, -2, mu), 0.0};
    }
    const double* lf = T.f[it];
    double nf = lf[0];
    double d[NPAIR];
#pragma unroll
    for (int q = 0; q < NPAIR; ++q) d[q] = R[q].dl * nf;
    double dl = l;
    double llp1 = dl * (dl + 1);
    double X000 = E;
    double X000p = -llp1 / 4. * X000;
    double X220 = 1. / 4. * lf[1] * E * e2;
    double X022 = E * e4;
    double X022p = -(llp1 - 4) / 4 * X022;
    double X121 = -1. / 2. * lf[2] * E * e83;
    double X132 = -1. / 2. * lf[3] * E * e203;
    double X242 = 1. / 4. * lf[4] * E * e10;
    double d00 = d[0], d11 = d[1], d1m1 = d[2], d2m2 = d[3], d22 = d[4], d31 = d[5];
    double d40 = d[6], d3m3 = d[7], d4m4 = d[8], d20 = d[9], d3m1 = d[10], d4m2 = d[11];
    double pre = 2. * dl + 1;
    a0 += pre * TT[j] * (X000 * X000 * d00 + lf[5] * C2 * X000p * X000p * d1m1 +
                         C2 * C2 * (X000p * X000p * d00 + X220 * X220 * d2m2));
    ap += pre * EE[j] * (X022 * X022 * d22 + 2 * C2 * X132 * X121 * d31 +
                         C2 * C2 * (X022p * X022p * d22 + X242 * X220 * d40));
    am += pre * EE[j] * (X022 * X022 * d2m2 + C2 * (X121 * X121 * d1m1 + X132 * X132 * d3m3) +
                         1. / 2. * C2 * C2 * (2 * X022p * X022p * d2m2 + X220 * X220 * d00 + X242 * X242 * d4m4));
    ax += pre * TE[j] * (X022 * X000 * d20 + C2 * lf[6] * X000p * (X121 * d11 + X132 * d3m1) +
                         1. / 2. * C2 * C2 * ((2 * X022p * X000p + X220 * X220) * d20 + X220 * X242 * d4m2));
    E *= r;
    r *= g;
    if (l >= 4) {
#pragma unroll
      for (int q = 0; q < NPAIR; ++q) rec_step_c(R[q], T.c[it][q][0], T.c[it][q][1], T.c[it][q][2], mu);
    } else {
#pragma unroll
      for (int q = 0; q < NPAIR; ++q)
        if (l >= PAIR_M[q]) rec_step_c(R[q], T.c[it][q][0], T.c[it][q][1], T.c[it][q][2], mu);
    }
  }
  if (live) {
    ksi[0 * n_mu + i] = a0 / 4. / PI;
    ksi[1 * n_mu + i] = ap / 4. / PI;
    ksi[2 * n_mu + i] = am / 4. / PI;
    ksi[3 * n_mu + i] = ax / 4. / PI;
  }
}

// pass 3: per-block partial back-transform sums, part[blk][4][n_l].  A
// thread carries BACK_MU nodes mu (i, i + 64, ...), so each wavefront sum
// over the block covers 64 BACK_MU nodes and the recursions of the nodes are
// independent work between the sums.
constexpr int BACK_MU = 4;
__global__ void __launch_bounds__(64) back_kernel(const double* __restrict__ ksi_all, int n_l,
                                                  const double* __restrict__ mu_arr, const double* __restrict__ w_arr,
                                                  int n_mu, const double* __restrict__ coef,
                                                  const double* __restrict__ lfac, int LT, double* __restrict__ part_all) {
  int ci = blockIdx.y;
  const double* ksi = ksi_all + (size_t)ci * 4 * n_mu;
  double* part = part_all + (size_t)ci * gridDim.x * 4 * n_l;
  double mu[BACK_MU], k0[BACK_MU], kp[BACK_MU], km[BACK_MU], kx[BACK_MU];
  Rec r00[BACK_MU], r20[BACK_MU], r22[BACK_MU], r2m2[BACK_MU];
  const double* c00 = coef;
  const double* c2m2 = coef + (size_t)3 * LT * 3;
  const double* c22 = coef + (size_t)4 * LT * 3;
  const double* c20 = coef + (size_t)9 * LT * 3;
#pragma unroll
  for (int m = 0; m < BACK_MU; ++m) {
    const int i = blockIdx.x * 64 * BACK_MU + m * 64 + threadIdx.x;
    const bool live = i < n_mu;
    mu[m] = live ? mu_arr[i] : 0.0;
    const double w = live ? w_arr[i] : 0.0;
    k0[m] = live ? ksi[0 * n_mu + i] * w : 0.0;
    kp[m] = live ? ksi[1 * n_mu + i] * w : 0.0;
    km[m] = live ? ksi[2 * n_mu + i] * w : 0.0;
    kx[m] = live ? ksi[3 * n_mu + i] * w : 0.0;
    r00[m] = Rec{wigner_start(0, 0, mu[m]), 0.0};
    r20[m] = Rec{wigner_start(2, 0, mu[m]), 0.0};
    r22[m] = Rec{wigner_start(2, 2, mu[m]), 0.0};
    r2m2[m] = Rec{wigner_start(2, -2, mu[m]), 0.0};
    rec_step(r00[m], c00, 0, mu[m]);
    rec_step(r00[m], c00, 1, mu[m]);
  }
  double* P = part + (size_t)blockIdx.x * 4 * n_l;
  for (int j = 0; j < n_l; ++j) {
    int l = j + 2;
    double nf = lfac[l * LFAC];
    double vt = 0.0, vx = 0.0, ve = 0.0, vb = 0.0;
#pragma unroll
    for (int m = 0; m < BACK_MU; ++m) {
      double d00 = r00[m].dl * nf, d20 = r20[m].dl * nf, d22 = r22[m].dl * nf, d2m2 = r2m2[m].dl * nf;
      vt += k0[m] * d00;
      vx += kx[m] * d20;
      ve += kp[m] * d22 + km[m] * d2m2;
      vb += kp[m] * d22 - km[m] * d2m2;
    }
    // transposed wavefront sums: row rho of w holds the total of value rho
    const double v4[4] = {vt, vx, ve, vb};
    double w[RowSums<4>::H2];
    wave_sum_rows<4>(v4, w);
    if ((threadIdx.x & 15) == 0) P[(threadIdx.x >> 4) * n_l + j] = w[0];
#pragma unroll
    for (int m = 0; m < BACK_MU; ++m) {
      rec_step(r00[m], c00, l, mu[m]);
      rec_step(r20[m], c20, l, mu[m]);
      rec_step(r22[m], c22, l, mu[m]);
      rec_step(r2m2[m], c2m2, l, mu[m]);
    }
  }
}

__global__ void back_final_kernel(const double* __restrict__ part_all, int nblk, int n_l, double* __restrict__ out_all) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  int ci = blockIdx.y;
  if (idx >= 4 * n_l) return;
  const double* part = part_all + (size_t)ci * nblk * 4 * n_l;
  double* out = out_all + (size_t)ci * 4 * n_l;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[(size_t)b * 4 * n_l + idx];
  int X = idx / n_l;
  double f = X < 2 ? 2 * PI : 1. / 2. * 2 * PI;
  out[idx] = s * f;
}


// ---- lensing contractions on the matrix cores ----
// The two passes of the correlation-function lensing that are plain
// contractions over a cosmology-independent operand become batched GEMMs
// across the cosmologies of a call, with the reduced Wigner d-matrices
// generated once and kept in HBM:
//   Cgl_c(mu)  = sum_l d11(mu, l)  F_c(l),  Cgl2_c(mu) = sum_l d1-1(mu, l) F_c(l),
//                F_c(l) = (2l+1) l(l+1) C^phiphi_c(l) / 4pi;
//   G_qc(l)    = sum_mu d_q(mu, l) w(mu) xi_qc(mu),  q = 00, 22, 2-2, 20,
// then TT = 2pi G_00, TE = 2pi G_20, EE/BB = pi (G_22 +- G_2-2).  The
// correlation pass between them (exp-weighted X functions of sigma^2(mu) per
// cosmology) is not a product of shared operands and stays on the vector
// units (ksi_kernel).
//
// d-matrix layout: [6][n_mu][n_l] of T (l = 2 .. n_l + 1 along rows), pairs in
// DM_PAIR order; the FP64 recursion tables (coef, lfac) follow the matrices.
constexpr int N_DM = 6;
constexpr int GEMM_MAX_SPLIT = 16;   // most slices of K in a lensing GEMM
__constant__ int DM_PAIR[N_DM] = {0, 1, 2, 3, 4, 9};   // d00 d11 d1-1 d2-2 d22 d20
enum { DM_00 = 0, DM_11, DM_1M1, DM_2M2, DM_22, DM_20 };

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <typename T> __device__ __forceinline__ T to_t(double v);
template <> __device__ __forceinline__ double to_t<double>(double v) { return v; }
template <> __device__ __forceinline__ float to_t<float>(double v) { return (float)v; }
template <> __device__ __forceinline__ __bf16 to_t<__bf16>(double v) { return (__bf16)(float)v; }

// one thread per node mu: the six d's along l, normalised as the vector path
// normalises them (the recursion value times sqrt(2 / (2l + 1)))
template <typename T>
__global__ void dmat_kernel(const double* __restrict__ mu_arr, int n_mu, int n_l, const double* __restrict__ coef,
                            const double* __restrict__ lfac, int LT, T* __restrict__ dm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= n_mu) return;
  const int q = DM_PAIR[p];
  const int m = PAIR_M[q], n = PAIR_N[q];
  const double mu = mu_arr[i];
  const double* cq = coef + (size_t)q * LT * 3;
  Rec r{wigner_start(m, n, mu), 0.0};
  for (int l = m; l < 2; ++l) rec_step(r, cq, l, mu);
  T* row = dm + ((size_t)p * n_mu + i) * n_l;
  for (int j = 0; j < n_l; ++j) {
    const int l = j + 2;
    row[j] = to_t<T>(r.dl * lfac[l * LFAC]);
    rec_step(r, cq, l, mu);
  }
}

// C(m, n) = sum_k A(m, k) B(k, n), A(m, k) = A[m sam + k sak], B(k, n) =
// B[k sbk + n sbn], C(m, n) -> C[m scm + n scn] (FP64).  One wavefront per 32 x
// 32 tile of C as 2 x 2 MFMA blocks of 16 x 16; operands straight from global
// memory (the d-matrix streams once per 32 or 64 columns of C, the batch
// operand stays in L2).  MFMA operand layout (16 x 16 x K blocks): lane l
// holds A(l % 16, k-slice l / 16), B(k-slice l / 16, l % 16); the four
// accumulator values of lane l hold column l % 16 and rows 4 (l / 16) + v
// (FP32 accumulation) or l / 16 + 4 v (FP64), as measured on gfx950.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  static constexpr int KS = 4, KV = 1;
  static constexpr bool ROW_STRIDED = true;   // accumulator v of lane l: row l / 16 + 4 v
  typedef double vec;
  typedef f64x4 acc;
  __device__ static acc run(vec a, vec b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template <>
struct Mfma<float> {
  static constexpr int KS = 4, KV = 1;
  static constexpr bool ROW_STRIDED = false;  // accumulator v of lane l: row 4 (l / 16) + v
  typedef float vec;
  typedef f32x4 acc;
  __device__ static acc run(vec a, vec b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};
template <>
struct Mfma<__bf16> {
  static constexpr int KS = 32, KV = 8;
  static constexpr bool ROW_STRIDED = false;
  typedef bf16x8 vec;
  typedef f32x4 acc;
  __device__ static acc run(vec a, vec b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};

template <typename T>
__device__ __forceinline__ typename Mfma<T>::vec load_op(const T* __restrict__ p, size_t s_row, size_t s_k, int row,
                                                         int rows, int k0, int K) {
  typename Mfma<T>::vec v;
  if constexpr (Mfma<T>::KV == 1) {
    v = (row < rows && k0 < K) ? p[(size_t)row * s_row + (size_t)k0 * s_k] : T(0);
  } else {
#pragma unroll
    for (int e = 0; e < Mfma<T>::KV; ++e) {
      const int k = k0 + e;
      v[e] = (row < rows && k < K) ? p[(size_t)row * s_row + (size_t)k * s_k] : (__bf16)0.0f;
    }
  }
  return v;
}

// K is cut into gridDim.z slices of kslice (a multiple of the MFMA depth); slice
// z writes its partial product to Cm + z * szc, and a reduction pass adds the
// slices in order (deterministic).
template <typename T, int NC>
__global__ void __launch_bounds__(64) gemm_mfma_kernel(int M, int N, int K, int kslice, const T* __restrict__ A,
                                                       size_t sam, size_t sak, const T* __restrict__ B, size_t sbk,
                                                       size_t sbn, double* __restrict__ Cm, size_t scm, size_t scn,
                                                       size_t szc) {
  using F = Mfma<T>;
  const int lane = threadIdx.x;
  const int li = lane & 15, lk = (lane >> 4) * F::KV;
  const int m0 = blockIdx.x * 32, n0 = blockIdx.y * 16 * NC;
  const int kb = blockIdx.z * kslice;
  const int ke = min(K, kb + kslice);
  Cm += (size_t)blockIdx.z * szc;
  typename F::acc acc[2][NC];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[r][c] = typename F::acc{0, 0, 0, 0};
  for (int k0 = kb; k0 < ke; k0 += F::KS) {
    const int k = k0 + lk;
    typename F::vec a[2], b[NC];
#pragma unroll
    for (int r = 0; r < 2; ++r) a[r] = load_op(A, sam, sak, m0 + 16 * r + li, M, k, ke);
#pragma unroll
    for (int c = 0; c < NC; ++c) b[c] = load_op(B, sbn, sbk, n0 + 16 * c + li, N, k, ke);
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[r][c] = F::run(a[r], b[c], acc[r][c]);
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + 16 * r + (F::ROW_STRIDED ? (lane >> 4) + 4 * v : 4 * (lane >> 4) + v);
        const int n = n0 + 16 * c + li;
        if (m < M && n < N) Cm[(size_t)m * scm + (size_t)n * scn] = (double)acc[r][c][v];
      }
}

// F[j][c] = (2l+1) l(l+1) C^phiphi_c(l) / 4pi, l = j + 2
template <typename T>
__global__ void cgl_prep_kernel(const double* __restrict__ clpp, int n_l, int n_cosmo, T* __restrict__ F) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (j >= n_l) return;
  const double l = j + 2;
  F[(size_t)j * n_cosmo + c] = to_t<T>((2. * l + 1) * l * (l + 1) * clpp[(size_t)c * n_l + j] / 4. / PI);
}

// X[q][mu][c] = w(mu) xi_qc(mu) with the correlation functions in GEMM order
// (d00 <- xi, d22 <- xi+, d2-2 <- xi-, d20 <- xi_x)
template <typename T>
__global__ void back_prep_kernel(const double* __restrict__ ksi_all, const double* __restrict__ w, int n_mu,
                                 int n_cosmo, T* __restrict__ X) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n_mu) return;
  const double* ksi = ksi_all + (size_t)c * 4 * n_mu;
  const double wi = w[i];
#pragma unroll
  for (int q = 0; q < 4; ++q) X[((size_t)q * n_mu + i) * n_cosmo + c] = to_t<T>(wi * ksi[q * n_mu + i]);
}

// out[c][4][n_l]: TT = 2pi G0, TE = 2pi G3, EE = pi (G1 + G2), BB = pi (G1 - G2);
// G[q][j][c] in the back_prep order (q: xi, xi+, xi-, xi_x)
__global__ void back_combine_kernel(const double* __restrict__ G, int n_l, int n_cosmo, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (j >= n_l) return;
  const size_t L = (size_t)n_l * n_cosmo, o = (size_t)j * n_cosmo + c;
  const double g0 = G[o], gp = G[L + o], gm = G[2 * L + o], gx = G[3 * L + o];
  double* oc = out + (size_t)c * 4 * n_l;
  oc[j] = 2 * PI * g0;
  oc[n_l + j] = 2 * PI * gx;
  oc[2 * n_l + j] = PI * (gp + gm);
  oc[3 * n_l + j] = PI * (gp - gm);
}

size_t elem_bytes(int precision) { return precision == 0 ? 8 : precision == 1 ? 4 : 2; }

// d-matrix buffer: matrices, then coef and lfac (FP64), 256-byte aligned
size_t dmat_parts(int n_mu, int n_l, int precision, size_t* off) {
  const int LT = n_l + 3;
  size_t o = 0;
  off[0] = 0;
  o += (size_t)N_DM * n_mu * n_l * elem_bytes(precision);
  o = (o + 255) / 256 * 256;
  off[1] = o;
  o += sizeof(double) * NPAIR * LT * 3;
  off[2] = o;
  o += sizeof(double) * LT * LFAC;
  return o;
}

// workspace of abcmb_lens_mfma: cgl, ksi, F, X, G
size_t lens_mfma_parts(int n_cosmo, int n_mu, int n_l, int precision, size_t* off) {  // off[6]
  const size_t eb = elem_bytes(precision);
  auto al = [](size_t v) { return (v + 255) / 256 * 256; };
  size_t o = 0;
  off[0] = o; o = al(o + sizeof(double) * n_cosmo * 2 * n_mu);         // cgl
  off[1] = o; o = al(o + sizeof(double) * n_cosmo * 4 * n_mu);         // ksi
  off[2] = o; o = al(o + eb * n_l * n_cosmo);                          // F
  off[3] = o; o = al(o + eb * 4 * n_mu * n_cosmo);                     // X
  off[4] = o; o = al(o + sizeof(double) * 4 * n_l * n_cosmo);          // G
  const int mn = n_mu > n_l ? n_mu : n_l;
  off[5] = o; o = al(o + sizeof(double) * GEMM_MAX_SPLIT * mn * n_cosmo);  // split-K partials
  return o;
}

// C(m, n) = sum over the S slices of P[z][m][n] (row-major partials)
__global__ void gemm_reduce_kernel(const double* __restrict__ P, int S, int M, int N, double* __restrict__ Cm,
                                   size_t scm, size_t scn) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)M * N) return;
  const int m = (int)(idx / N), n = (int)(idx % N);
  double v = 0.0;
  for (int z = 0; z < S; ++z) v += P[(size_t)z * M * N + idx];
  Cm[(size_t)m * scm + (size_t)n * scn] = v;
}

// slices of K for a skinny product (few cosmologies): enough workgroups to
// cover the CUs, each slice a multiple of the MFMA depth
inline int gemm_split(int tiles, int K, int ks) {
  int S = (512 + tiles - 1) / tiles;
  if (S > GEMM_MAX_SPLIT) S = GEMM_MAX_SPLIT;
  const int kmax = (K + ks - 1) / ks;   // no empty slice
  if (S > kmax) S = kmax;
  return S < 1 ? 1 : S;
}

// the column tile (32 or 64 cosmologies) is a tuned launch shape; part holds
// GEMM_MAX_SPLIT * M * N doubles for the split-K partials
template <typename T>
int launch_gemm(int M, int N, int K, const T* A, size_t sam, size_t sak, const T* B, size_t sbk, size_t sbn, double* C,
                size_t scm, size_t scn, double* part, hipStream_t s) {
  const bool wide = tune_get(ABCMB_TUNE_LENS_GEMM_COLS) == 64;
  const int tn = wide ? 64 : 32;
  const int tiles = ((M + 31) / 32) * ((N + tn - 1) / tn);
  const int S = gemm_split(tiles, K, Mfma<T>::KS);
  const int kslice = ((K + S - 1) / S + Mfma<T>::KS - 1) / Mfma<T>::KS * Mfma<T>::KS;
  const int Sk = (K + kslice - 1) / kslice;
  double* out = Sk > 1 ? part : C;
  const size_t om = Sk > 1 ? (size_t)N : scm, on = Sk > 1 ? 1 : scn, oz = (size_t)M * N;
  const dim3 grid((M + 31) / 32, (N + tn - 1) / tn, Sk);
  if (wide)
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 4>), grid, dim3(64), 0, s, M, N, K, kslice, A, sam, sak, B, sbk, sbn, out,
                       om, on, oz);
  else
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 2>), grid, dim3(64), 0, s, M, N, K, kslice, A, sam, sak, B, sbk, sbn, out,
                       om, on, oz);
  ABCMB_HIP_CHECK(hipGetLastError());
  if (Sk > 1) {
    hipLaunchKernelGGL(gemm_reduce_kernel, dim3((unsigned)(((size_t)M * N + 255) / 256)), dim3(256), 0, s, part, Sk, M,
                       N, C, scm, scn);
    ABCMB_HIP_CHECK(hipGetLastError());
  }
  return ABCMB_OK;
}

template <typename T>
int lens_mfma(int n_cosmo, const double* cl_in, const double* clpp, int n_l, const double* mu, const double* w,
              int n_mu, const void* dmat, int precision, double* out, void* workspace, hipStream_t s) {
  size_t doff[3], off[6];
  dmat_parts(n_mu, n_l, precision, doff);
  lens_mfma_parts(n_cosmo, n_mu, n_l, precision, off);
  const T* dm = (const T*)dmat;
  const double* coef = (const double*)((const char*)dmat + doff[1]);
  const double* lfac = (const double*)((const char*)dmat + doff[2]);
  char* ws = (char*)workspace;
  double* cgl = (double*)(ws + off[0]);
  double* ksi = (double*)(ws + off[1]);
  T* F = (T*)(ws + off[2]);
  T* X = (T*)(ws + off[3]);
  double* G = (double*)(ws + off[4]);
  double* part = (double*)(ws + off[5]);
  const int LT = n_l + 3;
  const size_t plane = (size_t)n_mu * n_l;
  hipLaunchKernelGGL((cgl_prep_kernel<T>), dim3((n_l + 255) / 256, n_cosmo), dim3(256), 0, s, clpp, n_l, n_cosmo, F);
  ABCMB_HIP_CHECK(hipGetLastError());
  // Cgl, Cgl2: (mu, c) = sum_l d(mu, l) F(l, c), written into cgl[c][0 or 1][mu]
  for (int h = 0; h < 2; ++h) {
    int rc = launch_gemm<T>(n_mu, n_cosmo, n_l, dm + (h == 0 ? DM_11 : DM_1M1) * plane, n_l, 1, F, n_cosmo, 1,
                            cgl + (size_t)h * n_mu, 1, (size_t)2 * n_mu, part, s);
    if (rc != ABCMB_OK) return rc;
  }
  const int nblk = (n_mu + 63) / 64;
  hipLaunchKernelGGL(ksi_kernel, dim3(nblk, n_cosmo), dim3(64), 0, s, cl_in, cgl, n_l, mu, n_mu, coef, lfac, LT, ksi);
  hipLaunchKernelGGL((back_prep_kernel<T>), dim3((n_mu + 255) / 256, n_cosmo), dim3(256), 0, s, ksi, w, n_mu,
                     n_cosmo, X);
  ABCMB_HIP_CHECK(hipGetLastError());
  // G[q](l, c) = sum_mu d_q(mu, l) X[q](mu, c)
  const int dq[4] = {DM_00, DM_22, DM_2M2, DM_20};
  for (int q = 0; q < 4; ++q) {
    int rc = launch_gemm<T>(n_l, n_cosmo, n_mu, dm + dq[q] * plane, 1, n_l, X + (size_t)q * n_mu * n_cosmo, n_cosmo,
                            1, G + (size_t)q * n_l * n_cosmo, n_cosmo, 1, part, s);
    if (rc != ABCMB_OK) return rc;
  }
  hipLaunchKernelGGL(back_combine_kernel, dim3((n_l + 255) / 256, n_cosmo), dim3(256), 0, s, G, n_l, n_cosmo, out);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}


// ---- reverse mode of the lensing (abcmb_lens_adjoint, abcmb_lensing_clpp_adjoint) ----
// From the cotangents lb[4][n_l] of the lensed TT, TE, EE, BB, the passes of
// abcmb_lens backwards:
//   back-transform: xib_q(mu) = f w(mu) sum_l d_q(mu, l) lb(l) (TT: d00, TE: d20
//     with f = 2 pi; xi+ with lb_EE + lb_BB against d22 and xi- with lb_EE - lb_BB
//     against d2-2, f = pi), one thread per node;
//   correlation pass: per node the xi's are linear in the unlensed spectra, so
//     clb(l) gets pre(l) K(l, mu) xib(mu) / 4pi summed over the nodes (per-block
//     sums, then across blocks), and the dependence on sigma^2(mu) and
//     Cgl2(mu) is carried by two-direction dual numbers through the same X
//     functions, giving their cotangents per node;
//   Cgl pass: cgl_bar(mu) = -s2_bar(mu) (plus the sum of s2_bar at the last
//     node, which every sigma^2 references) and cgl2_bar = C2_bar, then
//     clppb(l) = (2l+1) l (l+1) / 4pi sum_mu (cgl_bar d11 + cgl2_bar d1-1).
__global__ void __launch_bounds__(64) back_adj_kernel(const double* __restrict__ lb_all, int n_l,
                                                      const double* __restrict__ mu_arr,
                                                      const double* __restrict__ w_arr, int n_mu,
                                                      const double* __restrict__ coef,
                                                      const double* __restrict__ lfac, int LT,
                                                      double* __restrict__ xib_all) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int ci = blockIdx.y;
  if (i >= n_mu) return;
  const double* lb = lb_all + (size_t)ci * 4 * n_l;
  double* xib = xib_all + (size_t)ci * 4 * n_mu;
  const double mu = mu_arr[i], w = w_arr[i];
  const double* c00 = coef;
  const double* c2m2 = coef + (size_t)3 * LT * 3;
  const double* c22 = coef + (size_t)4 * LT * 3;
  const double* c20 = coef + (size_t)9 * LT * 3;
  Rec r00{wigner_start(0, 0, mu), 0.0}, r20{wigner_start(2, 0, mu), 0.0};
  Rec r22{wigner_start(2, 2, mu), 0.0}, r2m2{wigner_start(2, -2, mu), 0.0};
  rec_step(r00, c00, 0, mu);
  rec_step(r00, c00, 1, mu);
  double s0 = 0.0, sx = 0.0, sp = 0.0, sm = 0.0;
  for (int j = 0; j < n_l; ++j) {
    const int l = j + 2;
    const double nf = lfac[l * LFAC];
    s0 += lb[j] * (r00.dl * nf);
    sx += lb[n_l + j] * (r20.dl * nf);
    sp += (lb[2 * n_l + j] + lb[3 * n_l + j]) * (r22.dl * nf);
    sm += (lb[2 * n_l + j] - lb[3 * n_l + j]) * (r2m2.dl * nf);
    rec_step(r00, c00, l, mu);
    rec_step(r20, c20, l, mu);
    rec_step(r22, c22, l, mu);
    rec_step(r2m2, c2m2, l, mu);
  }
  xib[0 * n_mu + i] = 2. * PI * w * s0;
  xib[1 * n_mu + i] = PI * w * sp;
  xib[2 * n_mu + i] = PI * w * sm;
  xib[3 * n_mu + i] = 2. * PI * w * sx;
}

// partials part[blk][3][n_l] of the unlensed spectra's cotangents (TT, TE, EE),
// and s2c2b[2][n_mu]: the cotangents of sigma^2 and Cgl2 at each node
__global__ void __launch_bounds__(64) ksi_adj_kernel(const double* __restrict__ cl_all,
                                                     const double* __restrict__ cgl_all, int n_l,
                                                     const double* __restrict__ mu_arr, int n_mu,
                                                     const double* __restrict__ coef,
                                                     const double* __restrict__ lfac, int LT,
                                                     const double* __restrict__ xib_all, double* __restrict__ part_all,
                                                     double* __restrict__ s2c2b_all) {
  __shared__ CoefTile T;
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i0 < n_mu;
  const int i = live ? i0 : n_mu - 1;   // spare lanes follow the last node with zero cotangents
  const int ci = blockIdx.y;
  const double* cl_in = cl_all + (size_t)ci * 3 * n_l;
  const double* cgl = cgl_all + (size_t)ci * 2 * n_mu;
  const double* xib = xib_all + (size_t)ci * 4 * n_mu;
  double* part = part_all + ((size_t)ci * gridDim.x + blockIdx.x) * 3 * n_l;
  const double* TT = cl_in;
  const double* TE = cl_in + n_l;
  const double* EE = cl_in + 2 * n_l;
  const double b0 = live ? xib[0 * n_mu + i] / (4. * PI) : 0.0;
  const double bp = live ? xib[1 * n_mu + i] / (4. * PI) : 0.0;
  const double bm = live ? xib[2 * n_mu + i] / (4. * PI) : 0.0;
  const double bx = live ? xib[3 * n_mu + i] / (4. * PI) : 0.0;
  const double mu = mu_arr[i];
  D<2> s2(cgl[n_mu - 1] - cgl[i]), C2(cgl[n_mu + i]);
  s2.d[0] = 1.0;
  C2.d[1] = 1.0;
  Rec R[NPAIR];
#pragma unroll
  for (int q = 0; q < NPAIR; ++q) R[q] = Rec{0.0, 0.0};
  R[0] = Rec{wigner_start(0, 0, mu), 0.0};
  rec_step(R[0], coef + 0, 0, mu);
  rec_step(R[0], coef + 0, 1, mu);
  R[1] = Rec{wigner_start(1, 1, mu), 0.0};
  rec_step(R[1], coef + (size_t)1 * LT * 3, 1, mu);
  R[2] = Rec{wigner_start(1, -1, mu), 0.0};
  rec_step(R[2], coef + (size_t)2 * LT * 3, 1, mu);
  R[3] = Rec{wigner_start(2, -2, mu), 0.0};
  R[4] = Rec{wigner_start(2, 2, mu), 0.0};
  R[9] = Rec{wigner_start(2, 0, mu), 0.0};
  const D<2> e2 = dexp(2. * s2 / 4.), e4 = dexp(4. * s2 / 4.), e83 = dexp(8. / 3. * s2 / 4.);
  const D<2> e203 = dexp(20. / 3. * s2 / 4.), e10 = dexp(10. * s2 / 4.);
  const D<2> g = dexp(-0.5 * s2);
  D<2> E = dexp(-1.5 * s2), r = g * g * g;
  const D<2> C22 = C2 * C2;
  double s2b = 0.0, c2b = 0.0;
  for (int j = 0; j < n_l; ++j) {
    const int l = j + 2;
    const int it = j % KT;
    if (it == 0) load_tile(T, coef, lfac, LT, l);
    if (l == 3) {
      R[5] = Rec{wigner_start(3, 1, mu), 0.0};
      R[7] = Rec{wigner_start(3, -3, mu), 0.0};
      R[10] = Rec{wigner_start(3, -1, mu), 0.0};
    }
    if (l == 4) {
      R[6] = Rec{wigner_start(4, 0, mu), 0.0};
      R[8] = Rec{wigner_start(4, -4, mu), 0.0};
      R[11] = Rec{wigner_start(4, -2, mu), 0.0};
    }
    const double* lf = T.f[it];
    const double nf = lf[0];
    double d[NPAIR];
#pragma unroll
    for (int q = 0; q < NPAIR; ++q) d[q] = R[q].dl * nf;
    const double dl = l;
    const double llp1 = dl * (dl + 1);
    const D<2> X000 = E;
    const D<2> X000p = -llp1 / 4. * X000;
    const D<2> X220 = 1. / 4. * lf[1] * E * e2;
    const D<2> X022 = E * e4;
    const D<2> X022p = -(llp1 - 4) / 4 * X022;
    const D<2> X121 = -1. / 2. * lf[2] * E * e83;
    const D<2> X132 = -1. / 2. * lf[3] * E * e203;
    const D<2> X242 = 1. / 4. * lf[4] * E * e10;
    const double d00 = d[0], d11 = d[1], d1m1 = d[2], d2m2 = d[3], d22 = d[4], d31 = d[5];
    const double d40 = d[6], d3m3 = d[7], d4m4 = d[8], d20 = d[9], d3m1 = d[10], d4m2 = d[11];
    const double pre = 2. * dl + 1;
    const D<2> K0 = X000 * X000 * d00 + lf[5] * C2 * X000p * X000p * d1m1 +
                    C22 * (X000p * X000p * d00 + X220 * X220 * d2m2);
    const D<2> Kp = X022 * X022 * d22 + 2 * C2 * X132 * X121 * d31 + C22 * (X022p * X022p * d22 + X242 * X220 * d40);
    const D<2> Km = X022 * X022 * d2m2 + C2 * (X121 * X121 * d1m1 + X132 * X132 * d3m3) +
                    1. / 2. * C22 * (2 * X022p * X022p * d2m2 + X220 * X220 * d00 + X242 * X242 * d4m4);
    const D<2> Kx = X022 * X000 * d20 + C2 * lf[6] * X000p * (X121 * d11 + X132 * d3m1) +
                    1. / 2. * C22 * ((2 * X022p * X000p + X220 * X220) * d20 + X220 * X242 * d4m2);
    const double v3[3] = {pre * K0.v * b0, pre * Kx.v * bx, pre * (Kp.v * bp + Km.v * bm)};
    double w[RowSums<3>::H2];
    wave_sum_rows<3>(v3, w);
    if ((threadIdx.x & 15) == 0 && (threadIdx.x >> 4) < 3) part[(threadIdx.x >> 4) * n_l + j] = w[0];
    const double tt = TT[j], te = TE[j], ee = EE[j];
    s2b += pre * (tt * K0.d[0] * b0 + te * Kx.d[0] * bx + ee * (Kp.d[0] * bp + Km.d[0] * bm));
    c2b += pre * (tt * K0.d[1] * b0 + te * Kx.d[1] * bx + ee * (Kp.d[1] * bp + Km.d[1] * bm));
    E = E * r;
    r = r * g;
    if (l >= 4) {
#pragma unroll
      for (int q = 0; q < NPAIR; ++q) rec_step_c(R[q], T.c[it][q][0], T.c[it][q][1], T.c[it][q][2], mu);
    } else {
#pragma unroll
      for (int q = 0; q < NPAIR; ++q)
        if (l >= PAIR_M[q]) rec_step_c(R[q], T.c[it][q][0], T.c[it][q][1], T.c[it][q][2], mu);
    }
  }
  if (live) {
    s2c2b_all[(size_t)ci * 2 * n_mu + i] = s2b;
    s2c2b_all[(size_t)ci * 2 * n_mu + n_mu + i] = c2b;
  }
}

// partials part[blk][n_l] of the C_l^phiphi cotangents
__global__ void __launch_bounds__(64) cgl_adj_kernel(const double* __restrict__ s2c2b_all, int n_l,
                                                     const double* __restrict__ mu_arr, int n_mu,
                                                     const double* __restrict__ coef,
                                                     const double* __restrict__ lfac, int LT,
                                                     double* __restrict__ part_all) {
  const int i0 = blockIdx.x * 64 + threadIdx.x;
  const bool live = i0 < n_mu;
  const int i = live ? i0 : n_mu - 1;
  const int ci = blockIdx.y;
  const double* s2b = s2c2b_all + (size_t)ci * 2 * n_mu;
  const double* c2b = s2b + n_mu;
  double* part = part_all + ((size_t)ci * gridDim.x + blockIdx.x) * n_l;
  // every sigma^2(mu) = Cgl(mu_last) - Cgl(mu): the last node collects them all
  double tot = 0.0;
  for (int q = threadIdx.x; q < n_mu; q += 64) tot += s2b[q];
  double tv[1] = {tot};
  wave_sum(tv);
  const double lcg = live ? -s2b[i] + (i == n_mu - 1 ? tv[0] : 0.0) : 0.0;
  const double lcg2 = live ? c2b[i] : 0.0;
  const double mu = mu_arr[i];
  const double* c11 = coef + (size_t)1 * LT * 3;
  const double* c1m = coef + (size_t)2 * LT * 3;
  Rec a{wigner_start(1, 1, mu), 0.0}, b{wigner_start(1, -1, mu), 0.0};
  rec_step(a, c11, 1, mu);
  rec_step(b, c1m, 1, mu);
  for (int j = 0; j < n_l; ++j) {
    const int l = j + 2;
    const double nf = lfac[l * LFAC];
    double v[1] = {(2. * l + 1) * l * (l + 1) / (4. * PI) * (lcg * a.dl + lcg2 * b.dl) * nf};
    wave_sum(v);
    if (threadIdx.x == 0) part[j] = v[0];
    rec_step(a, c11, l, mu);
    rec_step(b, c1m, l, mu);
  }
}

// out[c][r] = sum over blocks of part[c][blk][r]
__global__ void sum_blocks_kernel(const double* __restrict__ part_all, int nblk, int n_row,
                                  double* __restrict__ out_all) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int ci = blockIdx.y;
  if (r >= n_row) return;
  const double* part = part_all + (size_t)ci * nblk * n_row;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[(size_t)b * n_row + r];
  out_all[(size_t)ci * n_row + r] = s;
}

// Limber C_l^phiphi backwards to the delta_m field of the perturbation table:
// one thread per (ln a node, cosmology) walks the multipoles (the k bracket of
// (l + 1/2) / chi moves forward with l) and adds 2 C d (trapezoid weight)
// (l factor) clppb(l) times the bilinear weights into the four table entries,
// kept in registers while the k bracket stays and added to the table (atomic
// adds) when it moves.
constexpr int CLPP_ADJ_THREADS = 64;
__global__ void __launch_bounds__(CLPP_ADJ_THREADS) clpp_adj_kernel(const AbcmbCosmo* __restrict__ cosmos,
                                                                    const double* __restrict__ lens_all, int n_lna,
                                                                    const double* __restrict__ ells, int n_l,
                                                                    const double* __restrict__ clppb_all,
                                                                    double* const* __restrict__ ptbar) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int ci = blockIdx.y;
  if (i >= n_lna) return;
  const AbcmbCosmo& c = cosmos[ci];
  const double* lna_ax = lens_all + (size_t)ci * 4 * n_lna;
  const double* aH_ax = lna_ax + n_lna;
  const double* tau_ax = lna_ax + 2 * n_lna;
  const double chi_rec = lna_ax[3 * n_lna];
  const double lna = lna_ax[i];
  if (!(lna < 0.0)) return;
  const double Om = c.omega_m / (c.h * c.h), OL = c.omega_Lambda / (c.h * c.h);
  const double ns = c.n_s;
  const double cnode = 9. / 8. / (PI * PI) * c.A_s * (2 * PI * PI) * pow(c.k_pivot, 1. - ns);
  const int n_save = c.n_save, nk = c.nk;
  const double* lna_save = c.lna_save;
  const double* kgrid = c.k;
  const double* dm = c.pt + (size_t)ABCMB_PT_DELTA_M * n_save * nk;
  double* dmb = ptbar[ci] + (size_t)ABCMB_PT_DELTA_M * n_save * nk;
  const double* clppb = clppb_all + (size_t)ci * n_l;
  const double chi = c.tau0 - tau_ax[i];
  const double win = (chi_rec - chi) / chi_rec / chi;
  const double a = exp(lna);
  const double z = 1. / a - 1.;
  const double aH = aH_ax[i];
  const double z3 = (1. + z) * (1. + z) * (1. + z);
  const double omz = (Om * z3) / ((Om * z3) + OL);
  const double aH2 = aH * aH;
  const double C = chi / aH * win * win * omz * omz * (aH2 * aH2) * cnode * pow(chi, 5. - ns);
  const double ichi = 1.0 / chi;
  const double lq = -log(1. + z);
  int ia, ib;
  double u = 0.0;
  if (lq <= lna_save[0]) { ia = ib = 0; }
  else if (lq >= lna_save[n_save - 1]) { ia = ib = n_save - 1; }
  else {
    int lo = 0, hi = n_save - 1;
    while (hi - lo > 1) { int mid = (lo + hi) >> 1; if (lna_save[mid] <= lq) lo = mid; else hi = mid; }
    ia = lo; ib = lo + 1;
    u = (lq - lna_save[lo]) / (lna_save[lo + 1] - lna_save[lo]);
  }
  // the trapezoid weight of node i (the sum runs over nodes 1 .. n_lna - 1)
  const double tw = 0.5 * ((i > 0 ? lna - lna_ax[i - 1] : 0.0) + (i + 1 < n_lna ? lna_ax[i + 1] - lna : 0.0));
  const double* ra = dm + (size_t)ia * nk;
  const double* rb = dm + (size_t)ib * nk;
  double* wa = dmb + (size_t)ia * nk;
  double* wb = dmb + (size_t)ib * nk;
  int ka = 0, cur0 = -1, cur1 = -1;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};   // (ia, k0), (ib, k0), (ia, k1), (ib, k1)
  auto flush = [&]() {
    if (cur0 >= 0) {
      atomicAdd(wa + cur0, acc[0]);
      atomicAdd(wb + cur0, acc[1]);
      atomicAdd(wa + cur1, acc[2]);
      atomicAdd(wb + cur1, acc[3]);
    }
    acc[0] = acc[1] = acc[2] = acc[3] = 0.0;
  };
  for (int il = 0; il < n_l; ++il) {
    const double e5 = ells[il] + 0.5;
    const double lfac = 8. * PI * PI / (e5 * e5 * e5) * pow(e5, ns - 5.);
    const double kk = e5 * ichi;
    while (ka < nk - 2 && kgrid[ka + 1] <= kk) ++ka;
    int k0, k1;
    double t;
    if (kk <= kgrid[0]) { k0 = k1 = 0; t = 0.0; }
    else if (kk >= kgrid[nk - 1]) { k0 = k1 = nk - 1; t = 0.0; }
    else { k0 = ka; k1 = ka + 1; t = (kk - kgrid[k0]) / (kgrid[k1] - kgrid[k0]); }
    const double da = ra[k0] + u * (rb[k0] - ra[k0]);
    const double db = ra[k1] + u * (rb[k1] - ra[k1]);
    const double d = da + t * (db - da);
    const double ld = 2. * C * d * tw * lfac * clppb[il];
    if (k0 != cur0 || k1 != cur1) {
      flush();
      cur0 = k0;
      cur1 = k1;
    }
    acc[0] += (1. - t) * (1. - u) * ld;
    acc[1] += (1. - t) * u * ld;
    acc[2] += t * (1. - u) * ld;
    acc[3] += t * u * ld;
  }
  flush();
}
}  // namespace
}  // namespace abcmb

using namespace abcmb;

extern "C" int abcmb_lensing_clpp(const AbcmbCosmo* cosmos_dev, int n_cosmo, const double* lens_dev, int n_lna,
                                  const double* ells_dev, int n_l, double* clpp_dev, void* stream) {
  ABCMB_REQUIRE(cosmos_dev && lens_dev && ells_dev && clpp_dev, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && n_lna >= 2 && n_l > 0, "bad sizes");
  hipLaunchKernelGGL(clpp_kernel, dim3((n_l + CLPP_THREADS - 1) / CLPP_THREADS, n_cosmo), dim3(CLPP_THREADS), 0,
                     (hipStream_t)stream, cosmos_dev, lens_dev, n_lna, ells_dev, n_l, clpp_dev);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}

static size_t lens_ws_parts(int n_cosmo, int n_mu, int n_l, size_t* off) {
  int LT = n_l + 3;
  size_t o = 0;
  off[0] = o; o += (size_t)NPAIR * LT * 3;                               // coef
  off[1] = o; o += (size_t)LT * LFAC;                                     // lfac
  off[2] = o; o += (size_t)n_cosmo * 2 * n_mu;                           // cgl
  off[3] = o; o += (size_t)n_cosmo * 4 * n_mu;                           // ksi
  off[4] = o; o += (size_t)n_cosmo * ((n_mu + 63) / 64) * 4 * n_l;       // partials
  return o;
}

extern "C" size_t abcmb_lens_workspace_size(int n_cosmo, int n_mu, int n_l) {
  size_t off[5];
  return sizeof(double) * lens_ws_parts(n_cosmo, n_mu, n_l, off);
}

extern "C" int abcmb_lens(int n_cosmo, const double* cl_in_dev, const double* clpp_dev, int n_l, const double* mu_dev,
                          const double* w_dev, int n_mu, double* out_dev, void* workspace, size_t ws_bytes,
                          void* stream) {
  ABCMB_REQUIRE(cl_in_dev && clpp_dev && mu_dev && w_dev && out_dev && workspace, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && n_l >= 3 && n_mu >= 2, "bad sizes");
  ABCMB_REQUIRE(ws_bytes >= abcmb_lens_workspace_size(n_cosmo, n_mu, n_l), "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  size_t off[5];
  lens_ws_parts(n_cosmo, n_mu, n_l, off);
  double* ws = (double*)workspace;
  double* coef = ws + off[0];
  double* lfac = ws + off[1];
  double* cgl = ws + off[2];
  double* ksi = ws + off[3];
  double* part = ws + off[4];
  int LT = n_l + 3;
  int nblk = (n_mu + 63) / 64;
  hipLaunchKernelGGL(wigner_coef_kernel, dim3((NPAIR * LT + 255) / 256), dim3(256), 0, s, LT, coef, lfac);
  hipLaunchKernelGGL(cgl_kernel, dim3(nblk, n_cosmo), dim3(64), 0, s, clpp_dev, n_l, mu_dev, n_mu, coef, lfac, LT,
                     cgl);
  hipLaunchKernelGGL(ksi_kernel, dim3(nblk, n_cosmo), dim3(64), 0, s, cl_in_dev, cgl, n_l, mu_dev, n_mu, coef, lfac,
                     LT, ksi);
  const int nblk_back = (n_mu + 64 * BACK_MU - 1) / (64 * BACK_MU);  // <= nblk: fits the part buffer
  hipLaunchKernelGGL(back_kernel, dim3(nblk_back, n_cosmo), dim3(64), 0, s, ksi, n_l, mu_dev, w_dev, n_mu, coef, lfac, LT,
                     part);
  hipLaunchKernelGGL(back_final_kernel, dim3((4 * n_l + 255) / 256, n_cosmo), dim3(256), 0, s, part, nblk_back, n_l,
                     out_dev);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}

extern "C" size_t abcmb_lens_dmat_size(int n_mu, int n_l, int precision) {
  size_t off[3];
  if (precision < 0 || precision > 2 || n_mu <= 0 || n_l <= 0) return 0;
  return dmat_parts(n_mu, n_l, precision, off);
}

extern "C" int abcmb_lens_dmat(const double* mu_dev, int n_mu, int n_l, int precision, void* dmat_dev,
                               size_t dmat_bytes, void* stream) {
  ABCMB_REQUIRE(mu_dev && dmat_dev, "null argument");
  ABCMB_REQUIRE(n_mu >= 2 && n_l >= 3, "bad sizes");
  ABCMB_REQUIRE(precision >= 0 && precision <= 2, "precision must be 0 (FP64), 1 (FP32) or 2 (BF16)");
  ABCMB_REQUIRE(dmat_bytes >= abcmb_lens_dmat_size(n_mu, n_l, precision), "d-matrix buffer too small");
  hipStream_t s = (hipStream_t)stream;
  size_t off[3];
  dmat_parts(n_mu, n_l, precision, off);
  double* coef = (double*)((char*)dmat_dev + off[1]);
  double* lfac = (double*)((char*)dmat_dev + off[2]);
  const int LT = n_l + 3;
  hipLaunchKernelGGL(wigner_coef_kernel, dim3((NPAIR * LT + 255) / 256), dim3(256), 0, s, LT, coef, lfac);
  const dim3 grid((n_mu + 63) / 64, N_DM);
  if (precision == 0)
    hipLaunchKernelGGL(dmat_kernel<double>, grid, dim3(64), 0, s, mu_dev, n_mu, n_l, coef, lfac, LT, (double*)dmat_dev);
  else if (precision == 1)
    hipLaunchKernelGGL(dmat_kernel<float>, grid, dim3(64), 0, s, mu_dev, n_mu, n_l, coef, lfac, LT, (float*)dmat_dev);
  else
    hipLaunchKernelGGL(dmat_kernel<__bf16>, grid, dim3(64), 0, s, mu_dev, n_mu, n_l, coef, lfac, LT, (__bf16*)dmat_dev);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}

extern "C" size_t abcmb_lens_mfma_workspace_size(int n_cosmo, int n_mu, int n_l, int precision) {
  size_t off[6];
  if (precision < 0 || precision > 2 || n_cosmo <= 0 || n_mu <= 0 || n_l <= 0) return 0;
  return lens_mfma_parts(n_cosmo, n_mu, n_l, precision, off);
}

extern "C" int abcmb_lens_mfma(int n_cosmo, const double* cl_in_dev, const double* clpp_dev, int n_l,
                               const double* mu_dev, const double* w_dev, int n_mu, const void* dmat_dev,
                               int precision, double* out_dev, void* workspace, size_t ws_bytes, void* stream) {
  ABCMB_REQUIRE(cl_in_dev && clpp_dev && mu_dev && w_dev && dmat_dev && out_dev && workspace, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && n_l >= 3 && n_mu >= 2, "bad sizes");
  ABCMB_REQUIRE(precision >= 0 && precision <= 2, "precision must be 0 (FP64), 1 (FP32) or 2 (BF16)");
  ABCMB_REQUIRE(ws_bytes >= abcmb_lens_mfma_workspace_size(n_cosmo, n_mu, n_l, precision), "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  if (precision == 0)
    return lens_mfma<double>(n_cosmo, cl_in_dev, clpp_dev, n_l, mu_dev, w_dev, n_mu, dmat_dev, 0, out_dev, workspace, s);
  if (precision == 1)
    return lens_mfma<float>(n_cosmo, cl_in_dev, clpp_dev, n_l, mu_dev, w_dev, n_mu, dmat_dev, 1, out_dev, workspace, s);
  return lens_mfma<__bf16>(n_cosmo, cl_in_dev, clpp_dev, n_l, mu_dev, w_dev, n_mu, dmat_dev, 2, out_dev, workspace, s);
}

static size_t lens_adj_parts(int n_cosmo, int n_mu, int n_l, size_t* off) {
  const int LT = n_l + 3;
  const size_t nblk = (n_mu + 63) / 64;
  size_t o = 0;
  off[0] = o; o += (size_t)NPAIR * LT * 3;                 // coef
  off[1] = o; o += (size_t)LT * LFAC;                       // lfac
  off[2] = o; o += (size_t)n_cosmo * 2 * n_mu;             // cgl
  off[3] = o; o += (size_t)n_cosmo * 4 * n_mu;             // xi cotangents
  off[4] = o; o += (size_t)n_cosmo * 2 * n_mu;             // sigma^2, Cgl2 cotangents
  off[5] = o; o += (size_t)n_cosmo * nblk * 3 * n_l;       // partials
  return o;
}

extern "C" size_t abcmb_lens_adjoint_workspace_size(int n_cosmo, int n_mu, int n_l) {
  size_t off[6];
  return sizeof(double) * lens_adj_parts(n_cosmo, n_mu, n_l, off);
}

extern "C" int abcmb_lens_adjoint(int n_cosmo, const double* cl_in_dev, const double* clpp_dev, int n_l,
                                  const double* mu_dev, const double* w_dev, int n_mu, const double* lensed_bar_dev,
                                  double* cl_bar_dev, double* clpp_bar_dev, void* workspace, size_t ws_bytes,
                                  void* stream) {
  ABCMB_REQUIRE(cl_in_dev && clpp_dev && mu_dev && w_dev && lensed_bar_dev && cl_bar_dev && clpp_bar_dev &&
                    workspace, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && n_l >= 3 && n_mu >= 2, "bad sizes");
  ABCMB_REQUIRE(ws_bytes >= abcmb_lens_adjoint_workspace_size(n_cosmo, n_mu, n_l), "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  size_t off[6];
  lens_adj_parts(n_cosmo, n_mu, n_l, off);
  double* ws = (double*)workspace;
  double* coef = ws + off[0];
  double* lfac = ws + off[1];
  double* cgl = ws + off[2];
  double* xib = ws + off[3];
  double* s2c2b = ws + off[4];
  double* part = ws + off[5];
  const int LT = n_l + 3;
  const int nblk = (n_mu + 63) / 64;
  hipLaunchKernelGGL(wigner_coef_kernel, dim3((NPAIR * LT + 255) / 256), dim3(256), 0, s, LT, coef, lfac);
  hipLaunchKernelGGL(cgl_kernel, dim3(nblk, n_cosmo), dim3(64), 0, s, clpp_dev, n_l, mu_dev, n_mu, coef, lfac, LT,
                     cgl);
  hipLaunchKernelGGL(back_adj_kernel, dim3(nblk, n_cosmo), dim3(64), 0, s, lensed_bar_dev, n_l, mu_dev, w_dev, n_mu,
                     coef, lfac, LT, xib);
  hipLaunchKernelGGL(ksi_adj_kernel, dim3(nblk, n_cosmo), dim3(64), 0, s, cl_in_dev, cgl, n_l, mu_dev, n_mu, coef,
                     lfac, LT, xib, part, s2c2b);
  hipLaunchKernelGGL(sum_blocks_kernel, dim3((3 * n_l + 255) / 256, n_cosmo), dim3(256), 0, s, part, nblk, 3 * n_l,
                     cl_bar_dev);
  hipLaunchKernelGGL(cgl_adj_kernel, dim3(nblk, n_cosmo), dim3(64), 0, s, s2c2b, n_l, mu_dev, n_mu, coef, lfac, LT,
                     part);
  hipLaunchKernelGGL(sum_blocks_kernel, dim3((n_l + 255) / 256, n_cosmo), dim3(256), 0, s, part, nblk, n_l,
                     clpp_bar_dev);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}

extern "C" int abcmb_lensing_clpp_adjoint(const AbcmbCosmo* cosmos_dev, int n_cosmo, const double* lens_dev,
                                          int n_lna, const double* ells_dev, int n_l, const double* clpp_bar_dev,
                                          double* const* pt_bar_dev, void* stream) {
  ABCMB_REQUIRE(cosmos_dev && lens_dev && ells_dev && clpp_bar_dev && pt_bar_dev, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && n_lna >= 2 && n_l > 0, "bad sizes");
  hipLaunchKernelGGL(clpp_adj_kernel, dim3((n_lna + CLPP_ADJ_THREADS - 1) / CLPP_ADJ_THREADS, n_cosmo),
                     dim3(CLPP_ADJ_THREADS), 0, (hipStream_t)stream, cosmos_dev, lens_dev, n_lna, ells_dev, n_l,
                     clpp_bar_dev, pt_bar_dev);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}
