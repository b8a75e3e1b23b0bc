// Perturbation ODE kernel: one (cosmology, k-mode) per workgroup, FP64.
//
// Integrates the linear Einstein-Boltzmann hierarchy of ABCMB
// (abcmb/perturbations.py:201-248 and the species y_prime functions in
// abcmb/species.py) from the per-k start time to ln a = 0 with the
// Kvaerno5 ESDIRK method, PID step control and chord Newton
// (abcmb/perturbations.py:250-316), writes cubic-Hermite dense output at the
// cosmology's saved ln a points and assembles the perturbation-table fields
// in the same pass (abcmb/perturbations.py:318-404).
//
// Layout on the workgroup (NT threads, one state row per thread, n <= NT):
//   - the right-hand side is linear, y' = A(ln a) y.  Each row keeps a
//     compact description of its row of A: up to 5 (column, coefficient,
//     background-scalar) terms plus its coupling to the two metric sums
//     h' and eta' (a rank-2 part of A).  Background scalars at a given
//     ln a are computed once per evaluation time and shared through LDS;
//   - the chord matrix M = I - h*gamma*A(t_n) is never formed: A is block
//     diagonal (one block per Boltzmann hierarchy) plus the rank-2 metric
//     coupling, so each thread inverts its row of the small diagonal block
//     in registers and M^-1 is applied with the Woodbury identity;
//   - the background scalars at the step's six abscissae are evaluated in
//     parallel, one lane per time, once per step;
//   - stage derivatives, state and error live in registers of the row's
//     thread; workgroup sums (metric sums, error norms, pivot search) use
//     wavefront shuffles plus one LDS slot per wavefront.
#include <cmath>
#include <cstdint>

#include "common.h"

namespace abcmb {
namespace {

using namespace cst;

// ---------------------------------------------------------------- tableau
constexpr double KV_G = 0.26;
__constant__ double KV_C[7] = {0.0, 0.52, 1.230333209967908, 0.8957659843500759, 0.43639360985864756, 1.0, 1.0};
__constant__ double KV_A[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {0.26, 0, 0, 0, 0, 0},
    {0.13, 0.84033320996790809, 0, 0, 0, 0},
    {0.22371961478320505, 0.47675532319799699, -0.06470895363112615, 0, 0, 0},
    {0.16648564323248321, 0.10450018841591720, 0.03631482272098715, -0.13090704451073998, 0, 0},
    {0.13855640231268224, 0.0, -0.04245337201752043, 0.02446657898003141, 0.61943039072480676, 0},
    {0.13659751177640291, 0.0, -0.05496908796538376, -0.04118626728321046, 0.62993304899016403,
     0.06962479448202728}};
__constant__ double KV_E[7] = {0.13659751177640291 - 0.13855640231268224, 0.0,
                               -0.05496908796538376 + 0.04245337201752043,
                               -0.04118626728321046 - 0.02446657898003141,
                               0.62993304899016403 - 0.61943039072480676,
                               0.06962479448202728 - 0.26, 0.26};

__constant__ double Q3[3] = {0.913201, 3.37517, 7.79184};
__constant__ double W3[3] = {0.0687359, 3.31435, 2.29911};
__constant__ double Q5[5] = {0.583165, 2.0, 4.0, 7.26582, 13.0};
__constant__ double W5[5] = {0.0081201, 0.689407, 2.8063, 2.05156, 0.12681};

// background scalar slots (per evaluation time)
enum {
  SC_ONE = 0, SC_IAH, SC_ITC, SC_ITAU, SC_CS2, SC_RTC,
  SC_H_ETA, SC_H_C, SC_H_B, SC_H_G, SC_H_U,
  SC_E_B, SC_E_G, SC_E_U,
  SC_P0,            // 5 slots: q_i / (eps_i aH)
  SC_HN0 = SC_P0 + 5,  // 5 slots: h'-weight of Psi0_i
  SC_EN = SC_HN0 + 5,  // eta'-weight base of kPsi1_i
  SC_ZERO,
  SC_COUNT
};

// ---------------------------------------------------------------- background
__device__ __forceinline__ double fast_interp(double x, double lo, double hi, const double* __restrict__ f, int n) {
  double i = (x - lo) / (hi - lo) * (n - 1);
  const double eps = 1e-6;
  i = fmin(fmax(i, eps), n - 1.0 - eps);
  int il = (int)floor(i);
  int iu = min(il + 1, n - 1);
  double wu = i - il;
  return (1.0 - wu) * f[il] + wu * f[iu];
}

struct Bg {
  double a, aH, tau, tauc, cs2, R;
  double rc, rb, rg, ru, rn, pn;
  double Tn, xn;
};

__device__ __forceinline__ double rho_unit() { return 3.0 * H0_OVER_H * H0_OVER_H / 8.0 / PI / GN; }

// densities, aH (abcmb/background.py:101-187; species rho at abcmb/species.py)
__device__ void bg_rho(const AbcmbCosmo& c, double lna, Bg& b) {
  const double hc3 = (C_CM * HBAR) * (C_CM * HBAR) * (C_CM * HBAR);
  double a = exp(lna);
  double a3 = a * a * a, a4 = a3 * a;
  double t04 = c.TCMB0 * c.TCMB0 * c.TCMB0 * c.TCMB0;
  double tu4 = c.T_ur * c.T_ur * c.T_ur * c.T_ur;
  b.a = a;
  double rde = c.omega_Lambda * rho_unit();
  b.rc = c.omega_cdm * rho_unit() / a3;
  b.rb = c.omega_b * rho_unit() / a3;
  b.rg = PI * PI / 15.0 * t04 / a4 / hc3;
  b.ru = c.N_ur * 2.0 * 7.0 / 8.0 * PI * PI / 30.0 * tu4 * t04 / a4 / hc3;
  b.rn = b.pn = 0.0;
  b.Tn = c.T_ncdm * c.TCMB0 / a;
  b.xn = c.m_ncdm / b.Tn;
  if (c.has_ncdm) {
    double ir = 0.0, ip = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      double q = Q5[i];
      double e = sqrt(q * q + b.xn * b.xn);
      ir += (1.0 + exp(-q)) / (q * q) * e * W5[i];
      ip += (1.0 + exp(-q)) / e * W5[i];
    }
    double T4 = b.Tn * b.Tn * b.Tn * b.Tn;
    double pre = T4 / (PI * PI) / (HBAR * HBAR * HBAR) / (C_CM * C_CM * C_CM);
    b.rn = c.N_ncdm * 4.0 * pre * ir;
    b.pn = c.N_ncdm * 4.0 / 3.0 * pre * ip;
  }
  double rt = rde + b.rc + b.rb + b.rg + b.ru + b.rn;
  b.aH = a * sqrt(8.0 * PI * GN * rt / 3.0) / C_MPC;
}

// x_e lookup with the reference's span quirk (abcmb/background.py:568-603)
__device__ __forceinline__ double bg_xe(const AbcmbCosmo& c, double lna) {
  if (lna < c.xe_x0) return c.xe_tab[0];
  if (lna >= c.xe_lna_last) return c.xe_last;
  return fast_interp(lna, c.xe_x0, c.xe_x0 + c.xe_n * c.xe_dx, c.xe_tab, c.xe_n);
}

__device__ __noinline__ void bg_eval(const AbcmbCosmo& c, double lna, Bg& b) {
  bg_rho(c, lna, b);
  b.tau = fast_interp(lna, ABCMB_TAU_TAB_LO, 0.0, c.tau_tab, ABCMB_TAU_TAB_N);
  double xe = bg_xe(c, lna);
  double nH = (1 - c.YHe) * 3.0 * c.omega_b * H0_OVER_H * H0_OVER_H / 8 / PI / GN / MH / (b.a * b.a * b.a);
  b.tauc = 1.0 / b.a / (nH * xe) / SIGMA_T / C_CM * C_MPC;
  // T_m (abcmb/background.py:605-657)
  double Tg = c.TCMB0 / b.a;
  double Tm;
  if (lna < c.tm_x0) {
    double fhe = c.YHe * MH / MHE / (1 - c.YHe);
    double stef = PI * PI / (60.0 * HBAR * HBAR * HBAR * C_CM * C_CM);
    double gc = xe / (1.0 + fhe + xe) * 8.0 * SIGMA_T * (4.0 * stef) * Tg * Tg * Tg * Tg / (3.0 * ME);
    double H = b.aH * C_MPC / b.a;
    Tm = Tg * (1.0 - H / gc);
  } else if (lna >= c.tm_lna_last) {
    Tm = c.tm_last;
  } else {
    Tm = fast_interp(lna, c.tm_x0, c.tm_x0 + c.tm_n * c.tm_dx, c.tm_tab, c.tm_n);
  }
  // c_s^2 (abcmb/species.py:1163-1220)
  double mu = MH / ((1.0 + xe) * (1.0 - c.YHe) + MH / MHE * c.YHe);
  b.R = 4.0 * b.rg / 3.0 / b.rb;
  b.cs2 = Tm / mu * (5.0 / 3.0 - 2.0 / 3.0 * mu * b.R / ME / b.aH / b.tauc * (Tg / Tm - 1.0));
}

// ---------------------------------------------------------------- row description
struct Row {
  int col[5];
  double cs[5];
  int si[5];
  double uh, ue;      // d(row)/d(h'), d(row)/d(eta')
  int whi, wei;       // this row as a column of h' / eta' (scalar slot)
  double whs, wes;    // static factor of that weight
};

struct Layout {
  int lg, lp, lu, ln, nq, n;
  int ig, ip, iu, inc;
};

__device__ __noinline__ void row_desc(const Layout& L, double k, int r, Row& w) {
  for (int j = 0; j < 5; ++j) { w.col[j] = r < L.n ? r : 0; w.cs[j] = 0.0; w.si[j] = SC_ZERO; }
  w.uh = w.ue = 0.0;
  w.whi = w.wei = SC_ZERO;
  w.whs = w.wes = 0.0;
  int nt = 0;
  auto add = [&](int c, double v, int s) { w.col[nt] = c; w.cs[nt] = v; w.si[nt] = s; ++nt; };
  const double k2 = k * k;
  if (r == 0) { w.ue = 1.0; w.whi = SC_H_ETA; w.whs = 1.0; return; }
  if (r == 1) { w.uh = -0.5; w.whi = SC_H_C; w.whs = 1.0; return; }
  if (r == 2) { add(3, -1.0, SC_IAH); w.uh = -0.5; w.whi = SC_H_B; w.whs = 1.0; return; }
  if (r == 3) {
    add(3, -1.0, SC_ONE); add(2, k2, SC_CS2); add(L.ig + 1, 1.0, SC_RTC); add(3, -1.0, SC_RTC);
    w.wei = SC_E_B; w.wes = 1.0;
    return;
  }
  if (r >= L.ig && r < L.ip) {  // photon temperature F_l
    int l = r - L.ig, b = L.ig;
    if (l == 0) { add(b + 1, -4.0 / 3.0, SC_IAH); w.uh = -2.0 / 3.0; w.whi = SC_H_G; w.whs = 1.0; }
    else if (l == 1) {
      add(b, k2 / 4.0, SC_IAH); add(b + 2, -k2, SC_IAH); add(3, 1.0, SC_ITC); add(b + 1, -1.0, SC_ITC);
      w.wei = SC_E_G; w.wes = 1.0;
    } else if (l == 2) {
      add(b + 1, 4.0 / 15.0, SC_IAH); add(b + 3, -3.0 / 10.0 * k, SC_IAH); add(b + 2, -9.0 / 10.0, SC_ITC);
      add(L.ip, 1.0 / 20.0, SC_ITC); add(L.ip + 2, 1.0 / 20.0, SC_ITC);
      w.uh = 2.0 / 15.0; w.ue = 4.0 / 5.0;
    } else if (l == 3) {
      add(b + 2, 6.0 * k / 7.0, SC_IAH); add(b + 4, -4.0 * k / 7.0, SC_IAH); add(b + 3, -1.0, SC_ITC);
    } else if (l < L.lg) {
      add(b + l - 1, l * k / (2.0 * l + 1.0), SC_IAH); add(b + l + 1, -(l + 1) * k / (2.0 * l + 1.0), SC_IAH);
      add(b + l, -1.0, SC_ITC);
    } else {
      add(b + l - 1, k, SC_IAH); add(b + l, -(double)(L.lg + 1), SC_ITAU); add(b + l, -1.0, SC_ITC);
    }
    return;
  }
  if (r >= L.ip && r < L.iu) {  // photon polarisation G_l
    int l = r - L.ip, b = L.ip;
    if (l == 0) {
      add(b + 1, -k, SC_IAH); add(b, -0.5, SC_ITC); add(L.ig + 2, 1.0, SC_ITC); add(b + 2, 0.5, SC_ITC);
    } else if (l == 1) {
      add(b, k / 3.0, SC_IAH); add(b + 2, -2.0 * k / 3.0, SC_IAH); add(b + 1, -1.0, SC_ITC);
    } else if (l == 2) {
      add(b + 1, 2.0 * k / 5.0, SC_IAH); add(b + 3, -3.0 * k / 5.0, SC_IAH); add(b + 2, -0.9, SC_ITC);
      add(L.ig + 2, 0.2, SC_ITC); add(b, 0.1, SC_ITC);
    } else if (l < L.lp) {
      add(b + l - 1, l * k / (2.0 * l + 1.0), SC_IAH); add(b + l + 1, -(l + 1) * k / (2.0 * l + 1.0), SC_IAH);
      add(b + l, -1.0, SC_ITC);
    } else {
      add(b + l - 1, k, SC_IAH); add(b + l, -(double)(L.lp + 1), SC_ITAU); add(b + l, -1.0, SC_ITC);
    }
    return;
  }
  if (r >= L.iu && r < L.inc) {  // massless neutrinos
    int l = r - L.iu, b = L.iu;
    if (l == 0) { add(b + 1, -4.0 / 3.0, SC_IAH); w.uh = -2.0 / 3.0; w.whi = SC_H_U; w.whs = 1.0; }
    else if (l == 1) { add(b, k2 / 4.0, SC_IAH); add(b + 2, -k2, SC_IAH); w.wei = SC_E_U; w.wes = 1.0; }
    else if (l == 2) {
      add(b + 1, 4.0 / 15.0, SC_IAH); add(b + 3, -3.0 / 10.0 * k, SC_IAH); w.uh = 2.0 / 15.0; w.ue = 4.0 / 5.0;
    } else if (l == 3) { add(b + 2, 6.0 * k / 7.0, SC_IAH); add(b + 4, -4.0 * k / 7.0, SC_IAH); }
    else if (l < L.lu) {
      add(b + l - 1, l * k / (2.0 * l + 1.0), SC_IAH); add(b + l + 1, -(l + 1) * k / (2.0 * l + 1.0), SC_IAH);
    } else { add(b + l - 1, k, SC_IAH); add(b + l, -(double)(L.lu + 1), SC_ITAU); }
    return;
  }
  if (r >= L.inc && r < L.n) {  // massive neutrinos, bin i
    int nl = L.ln + 1;
    int i = (r - L.inc) / nl, l = (r - L.inc) % nl, b = L.inc + i * nl;
    double q = L.nq == 5 ? Q5[i] : Q3[i];
    double wq = L.nq == 5 ? W5[i] : W3[i];
    double dlf = -q / (1.0 + exp(-q));
    int sp = SC_P0 + i;
    if (l == 0) { add(b + 1, -1.0, sp); w.uh = dlf / 6.0; w.whi = SC_HN0 + i; w.whs = 1.0; }
    else if (l == 1) {
      add(b, k2 / 3.0, sp); add(b + 2, -2.0 * k2 / 3.0, sp);
      w.wei = SC_EN; w.wes = wq * (1.0 + exp(-q)) / q;
    } else if (l == 2) {
      add(b + 1, 2.0 / 5.0, sp); add(b + 3, -3.0 * k / 5.0, sp); w.uh = -dlf / 15.0; w.ue = -2.0 / 5.0 * dlf;
    } else if (l < L.ln) {
      add(b + l - 1, l * k / (2.0 * l + 1.0), sp); add(b + l + 1, -(l + 1) * k / (2.0 * l + 1.0), sp);
    } else { add(b + l - 1, k, sp); add(b + l, -(double)(L.ln + 1), SC_ITAU); }
    return;
  }
}

// scalars of A(ln a) shared by all rows (written by thread 0)
__device__ __noinline__ void fill_scalars(const AbcmbCosmo& c, const Layout& L, double k, const Bg& b, double* sc) {
  double ah = b.aH;
  sc[SC_ONE] = 1.0;
  sc[SC_IAH] = 1.0 / ah;
  sc[SC_ITC] = 1.0 / (ah * b.tauc);
  sc[SC_ITAU] = 1.0 / (ah * b.tau);
  sc[SC_CS2] = b.cs2 / ah;
  sc[SC_RTC] = b.R / (b.tauc * ah);
  double kh = 2.0 / (ah * ah) * 4.0 * PI * GN * b.a * b.a / (C_MPC * C_MPC);
  sc[SC_H_ETA] = 2.0 / (ah * ah) * k * k;
  sc[SC_H_C] = kh * b.rc;
  sc[SC_H_B] = kh * b.rb;
  sc[SC_H_G] = kh * b.rg;
  sc[SC_H_U] = kh * b.ru;
  double ke = 4.0 * PI * GN * b.a * b.a / ah / (k * k) / (C_MPC * C_MPC);
  sc[SC_E_B] = ke * b.rb;
  sc[SC_E_G] = ke * (b.rg + b.rg / 3.0);
  sc[SC_E_U] = ke * (b.ru + b.ru / 3.0);
  double pre = 0.0;
  if (c.has_ncdm) pre = c.N_ncdm * b.Tn * b.Tn * b.Tn * b.Tn / (PI * PI) / (HBAR * HBAR * HBAR) / (C_CM * C_CM * C_CM);
  for (int i = 0; i < 5; ++i) {
    double q = i < L.nq ? (L.nq == 5 ? Q5[i] : Q3[i]) : 1.0;
    double wq = i < L.nq ? (L.nq == 5 ? W5[i] : W3[i]) : 0.0;
    double e = sqrt(q * q + b.xn * b.xn);
    sc[SC_P0 + i] = q / e / ah;
    sc[SC_HN0 + i] = kh * 4.0 * pre * wq * (1.0 + exp(-q)) * e / (q * q);
  }
  sc[SC_EN] = ke * 4.0 * pre;
  sc[SC_ZERO] = 0.0;
}

// ---------------------------------------------------------------- the kernel
// The chord matrix is I - g A(t_n + jac_offset h) (the oracle uses the same time).
// A(ln a) = B + u_h v_h^T + u_e v_e^T.  B is block diagonal with blocks
// {eta}, {delta_c}, {delta_b, theta_b, F_0..F_lg, G_0..G_lp}, {N_0..N_lu} and
// one block per massive-neutrino momentum bin; the rank-2 part is the metric
// coupling through h' and eta' (column weights v = d h'/dy, d eta'/dy, row
// weights u = d y'_r / d h', d y'_r / d eta').  Each thread keeps its row of
// (I - g B_blk)^-1 in registers, built by Gauss-Jordan inside the block in
// natural order (the block pivots stay >= 1 for these matrices; a pivot
// below 1e-12 is reported as a failed step), and (I - g A)^-1 is applied
// with the Woodbury identity through a 2x2 capacitance matrix.

// chord updates below this (in units of the error tolerance) are rounding noise
constexpr double NEWTON_SMALL = 1e-9;
constexpr int NTIMES = 7;  // t, the distinct stage abscissae t + c_s h, and the chord-Jacobian time
constexpr int JAC_SLOT = 6;
__constant__ double TIME_C[NTIMES - 1] = {0.0, 0.52, 1.230333209967908, 0.8957659843500759, 0.43639360985864756,
                                          1.0};
__constant__ int STAGE_SLOT[7] = {0, 1, 2, 3, 4, 5, 5};
constexpr int NBLK = 10;

struct Blk {
  int id, start, size, lr;
};

__device__ Blk row_block(const Layout& L, int r) {
  Blk b;
  if (r == 0) { b.id = 0; b.start = 0; b.size = 1; }
  else if (r == 1) { b.id = 1; b.start = 1; b.size = 1; }
  else if (r < L.iu) { b.id = 2; b.start = 2; b.size = L.iu - 2; }
  else if (r < L.inc) { b.id = 3; b.start = L.iu; b.size = L.lu + 1; }
  else if (r < L.n) {
    int i = (r - L.inc) / (L.ln + 1);
    b.id = 4 + i; b.start = L.inc + i * (L.ln + 1); b.size = L.ln + 1;
  } else { b.id = NBLK - 1; b.start = r; b.size = 0; }
  b.lr = r - b.start;
  return b;
}

template <int NT, int BMAX>
struct Smem {
  double sc[NTIMES][SC_COUNT];
  double zl[NT + BMAX];
  double rl[NT + BMAX];
  double uh[NT + BMAX];
  double ue[NT + BMAX];
  double pbuf[2][NBLK][BMAX];
  double red[8 * (NT / 64)];
  double Fs[7][NT];  // stage derivatives, one column per thread
};

// workgroup sum of NV values (wave shuffles + one LDS slot per wave)
template <int NT, int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red) {
#pragma unroll
  for (int q = 0; q < NV; ++q) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
  }
  if constexpr (NT > 64) {
    constexpr int NW = NT / 64;
    int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < NV; ++q) red[q * NW + wv] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int i = 0; i < NW; ++i) s += red[q * NW + i];
      v[q] = s;
    }
  }
}

// background scalars of A at t + TIME_C[i] * hh, i < cnt (one lane per time)
template <int NT, int BMAX>
__device__ void set_times(Smem<NT, BMAX>& sm, const AbcmbCosmo& c, const Layout& L, double k, double t, double hh,
                          int cnt, double jac_offset) {
  __syncthreads();
  int r = threadIdx.x;
  if (r < cnt) {
    Bg b;
    bg_eval(c, t + (r == JAC_SLOT ? jac_offset : TIME_C[r]) * hh, b);
    fill_scalars(c, L, k, b, sm.sc[r]);
  }
  __syncthreads();
}

struct RowState {
  Row w;
  Blk B;
  int lc[5];  // term columns local to the block
  int n;
};

// f_r(t_slot, z)
template <int NT, int BMAX>
__device__ __forceinline__ double rhs_eval(Smem<NT, BMAX>& sm, const RowState& R, int slot, double z) {
  const int r = threadIdx.x;
  sm.zl[r] = z;
  __syncthreads();
  const double* sc = sm.sc[slot];
  double v[2] = {R.w.whs * sc[R.w.whi] * z, R.w.wes * sc[R.w.wei] * z};
  block_sum<NT, 2>(v, sm.red);
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < 5; ++j) acc += R.w.cs[j] * sc[R.w.si[j]] * sm.zl[R.w.col[j]];
  acc += R.w.uh * v[0] + R.w.ue * v[1];
  return r < R.n ? acc : 0.0;
}

template <int BMAX>
struct Inv {
  double row[BMAX];   // this thread's row of (I - g B_blk)^-1 (block-local columns)
  double Yh, Ye;      // rows of S^-1 g u_h, S^-1 g u_e
  double c00, c01, c10, c11;  // inverse capacitance
  double Vh, Ve;      // this row's weights in h', eta' at t_n
};

// build and invert S = I - g B(t_n) block-wise, then the Woodbury terms
template <int NT, int BMAX>
__device__ bool factor(Smem<NT, BMAX>& sm, const RowState& R, int maxb, double g, Inv<BMAX>& F) {
  const double* sc = sm.sc[JAC_SLOT];
  const Blk& B = R.B;
#pragma unroll
  for (int c = 0; c < BMAX; ++c) F.row[c] = (c == B.lr) ? 1.0 : 0.0;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    double v = g * R.w.cs[j] * sc[R.w.si[j]];
#pragma unroll
    for (int c = 0; c < BMAX; ++c) F.row[c] -= (R.lc[j] == c) ? v : 0.0;
  }
  double bad = 0.0;
  for (int p = 0; p < maxb; ++p) {
    double* pb = sm.pbuf[p & 1][B.id];
    const bool in = p < B.size;
    if (in && B.lr == p) {
      double d = 0.0;
#pragma unroll
      for (int c = 0; c < BMAX; ++c) d = (c == p) ? F.row[c] : d;
      if (!(fabs(d) > 1e-12)) bad = 1.0;
      double inv = 1.0 / d;
#pragma unroll
      for (int c = 0; c < BMAX; ++c) {
        double v = (c == p) ? inv : F.row[c] * inv;
        F.row[c] = v;
        pb[c] = v;
      }
    }
    __syncthreads();
    if (in && B.lr != p && B.lr < B.size) {
      double f = 0.0;
#pragma unroll
      for (int c = 0; c < BMAX; ++c) f = (c == p) ? F.row[c] : f;
#pragma unroll
      for (int c = 0; c < BMAX; ++c) {
        double pv = pb[c];
        F.row[c] = (c == p) ? -f * pv : F.row[c] - f * pv;
      }
    }
  }
  double yh = 0.0, ye = 0.0;
#pragma unroll
  for (int c = 0; c < BMAX; ++c) {
    yh += F.row[c] * sm.uh[B.start + c];
    ye += F.row[c] * sm.ue[B.start + c];
  }
  F.Yh = g * yh;
  F.Ye = g * ye;
  F.Vh = R.w.whs * sc[R.w.whi];
  F.Ve = R.w.wes * sc[R.w.wei];
  double v[5] = {F.Vh * F.Yh, F.Vh * F.Ye, F.Ve * F.Yh, F.Ve * F.Ye, bad};
  block_sum<NT, 5>(v, sm.red);
  double a = 1.0 - v[0], b = -v[1], cc = -v[2], d = 1.0 - v[3];
  double det = a * d - b * cc;
  F.c00 = d / det; F.c01 = -b / det; F.c10 = -cc / det; F.c11 = a / det;
  return v[4] == 0.0 && isfinite(F.c00) && isfinite(F.c11) && det != 0.0;
}

// x = (I - g A(t_n))^-1 b
template <int NT, int BMAX>
__device__ __forceinline__ double solve(Smem<NT, BMAX>& sm, const RowState& R, const Inv<BMAX>& F, double b) {
  const int r = threadIdx.x;
  sm.rl[r] = r < R.n ? b : 0.0;
  __syncthreads();
  double z = 0.0;
#pragma unroll
  for (int c = 0; c < BMAX; ++c) z += F.row[c] * sm.rl[R.B.start + c];
  double v[2] = {F.Vh * z, F.Ve * z};
  block_sum<NT, 2>(v, sm.red);
  double a = F.c00 * v[0] + F.c01 * v[1];
  double bb = F.c10 * v[0] + F.c11 * v[1];
  return r < R.n ? z + F.Yh * a + F.Ye * bb : 0.0;
}

__device__ __forceinline__ double rms_term(double e, double y0, double y1, double rtol, double atol) {
  double m = fmax(fabs(y0), fabs(y1));
  double v = e / (atol + rtol * m);
  return v * v;
}

// initial conditions (abcmb/perturbations.py:161-199 and species y_ini)
__device__ __noinline__ double ic_row(const AbcmbCosmo& c, const Layout& L, double k, double ti, int r) {
  double om = c.om, Rn = c.R_nu, Rb = c.R_b;
  double kt = k * ti;
  double dg = -kt * kt / 3. * (1. - om * ti / 5.);
  double thb = -k * k * k * k * ti * ti * ti / 36. * (1. - 3. * (1. + 5. * Rb - Rn) / 20. / (1. - Rn) * om * ti);
  double tn = -k * kt * kt * kt / 36. / (4. * Rn + 15.) *
              (4. * Rn + 11. + 12. - 3. * (8. * Rn * Rn + 50. * Rn + 275.) / 20. / (2. * Rn + 15.) * ti * om);
  double sn = kt * kt / (45. + 12. * Rn) * 2. * (1. + (4. * Rn - 5.) / 4. / (2. * Rn + 15.) * ti * om);
  if (r == 0)
    return 1. - k * k * ti * ti / 12. / (15. + 4. * Rn) *
                    (5. + 4. * Rn - (16. * Rn * Rn + 280. * Rn + 325) / 10. / (2. * Rn + 15.) * ti * om);
  if (r == 1 || r == 2) return -kt * kt / 4. * (1. - om * ti / 5.);
  if (r == 3) return thb;
  if (r == L.ig) return dg;
  if (r == L.ig + 1) return thb;
  if (r == L.iu) return dg;
  if (r == L.iu + 1) return tn;
  if (r == L.iu + 2) return sn;
  if (r >= L.inc && r < L.n) {
    int nl = L.ln + 1;
    int i = (r - L.inc) / nl, l = (r - L.inc) % nl;
    double q = L.nq == 5 ? Q5[i] : Q3[i];
    double f = q / (1. + exp(-q));
    if (l == 0) return dg / 4. * f;
    if (l == 1) return tn / 3. * f;
    if (l == 2) return sn / 2. * f;
  }
  return 0.0;
}

// output-table fields at one saved point from the state in LDS (zl)
// (abcmb/perturbations.py:318-404; output_perturbations in abcmb/species.py)
__device__ __noinline__ double output_field(const AbcmbCosmo& c, const Layout& L, double k, double lna, const double* y, int f) {
  Bg b;
  bg_eval(c, lna, b);
  double a = b.a, aH = b.aH;
  double nrd = 0.0, nrt = 0.0, nrs = 0.0;
  if (L.nq > 0) {
    int nl = L.ln + 1;
    double s0 = 0, s1 = 0, s2 = 0;
    for (int i = 0; i < L.nq; ++i) {
      double q = L.nq == 5 ? Q5[i] : Q3[i];
      double wq = L.nq == 5 ? W5[i] : W3[i];
      double e = sqrt(q * q + b.xn * b.xn);
      const double* P = y + L.inc + i * nl;
      s0 += wq * (1.0 + exp(-q)) * e / (q * q) * P[0];
      s1 += wq * (1.0 + exp(-q)) / q * P[1];
      s2 += wq * (1.0 + exp(-q)) / e * P[2];
    }
    double pre = c.N_ncdm * b.Tn * b.Tn * b.Tn * b.Tn / (PI * PI) / (HBAR * HBAR * HBAR) / (C_CM * C_CM * C_CM);
    nrd = s0 * 4.0 * pre; nrt = s1 * 4.0 * pre; nrs = s2 * 8.0 / 3.0 * pre;
  }
  switch (f) {
    case ABCMB_PT_THETA_B_PRIME:
      return -y[3] + b.cs2 / aH * (k * k * y[2]) + b.R / aH / b.tauc * (y[L.ig + 1] - y[3]);
    case ABCMB_PT_ETA: return y[0];
    case ABCMB_PT_CDM_DELTA: return y[1];
    case ABCMB_PT_B_DELTA: return y[2];
    case ABCMB_PT_B_THETA: return y[3];
    case ABCMB_PT_G_DELTA: return y[L.ig];
    case ABCMB_PT_G_THETA: return y[L.ig + 1];
    case ABCMB_PT_G_SIGMA: return y[L.ig + 2];
    case ABCMB_PT_G_G0: return y[L.ip];
    case ABCMB_PT_G_G2: return y[L.ip + 2];
    case ABCMB_PT_UR_DELTA: return y[L.iu];
    case ABCMB_PT_UR_THETA: return y[L.iu + 1];
    case ABCMB_PT_UR_SIGMA: return y[L.iu + 2];
    case ABCMB_PT_NCDM_DELTA: return L.nq > 0 ? nrd / b.rn : 0.0;
    case ABCMB_PT_NCDM_THETA: return L.nq > 0 ? nrt / (b.rn + b.pn) : 0.0;
    case ABCMB_PT_NCDM_SIGMA: return L.nq > 0 ? nrs / (b.rn + b.pn) : 0.0;
    default: break;
  }
  double srd = b.rc * y[1] + b.rb * y[2] + b.rg * y[L.ig] + b.ru * y[L.iu];
  double srt = b.rb * y[3] + (b.rg + b.rg / 3.) * y[L.ig + 1] + (b.ru + b.ru / 3.) * y[L.iu + 1];
  double srs = (b.rg + b.rg / 3.) * y[L.ig + 2] + (b.ru + b.ru / 3.) * y[L.iu + 2];
  double srdm = b.rc * y[1] + b.rb * y[2], srm = b.rc + b.rb;
  if (L.nq > 0) { srd += nrd; srt += nrt; srs += nrs; srdm += nrd; srm += b.rn; }
  if (f == ABCMB_PT_DELTA_M) return srdm / srm;
  double KG = 4. * PI * GN * a * a / (C_MPC * C_MPC);
  double hp = 2. / (aH * aH) * (k * k * y[0] + KG * srd);
  double ep = KG / aH * srt / (k * k);
  if (f == ABCMB_PT_H_PRIME) return hp;
  if (f == ABCMB_PT_ETA_PRIME) return ep;
  double al = aH * (hp + 6. * ep) / 2. / (k * k);
  if (f == ABCMB_PT_ALPHA) return al;
  return y[0] / aH - 2. * al - 3. * KG / aH * srs / (k * k);  // ABCMB_PT_ALPHA_PRIME
}

template <int NT, int BMAX>
__global__ void __launch_bounds__(NT) pert_fwd_kernel(const AbcmbCosmo* __restrict__ cosmos,
                                                      const AbcmbPertItem* __restrict__ items,
                                                      AbcmbOdeSettings st, int32_t* __restrict__ status_out,
                                                      int32_t* __restrict__ steps_out, int32_t* __restrict__ iters_out,
                                                      double* __restrict__ state_out) {
  __shared__ Smem<NT, BMAX> sm;
  const int r = threadIdx.x;
  const int item = blockIdx.x;
  const AbcmbPertItem I = items[item];
  const AbcmbCosmo& c = cosmos[I.cosmo];
  const double k = I.k;
  Layout L;
  L.lg = st.l_g; L.lp = st.l_pol; L.lu = st.l_ur; L.ln = st.l_ncdm; L.nq = st.n_q; L.n = st.n;
  L.ig = 4; L.ip = 4 + L.lg + 1; L.iu = L.ip + L.lp + 1; L.inc = L.iu + L.lu + 1;
  const int n = L.n;
  RowState R;
  row_desc(L, k, r, R.w);
  R.B = row_block(L, r);
  R.n = n;
#pragma unroll
  for (int j = 0; j < 5; ++j) R.lc[j] = R.w.col[j] - R.B.start;
  const int maxb = max(L.iu - 2, max(L.lu + 1, L.nq > 0 ? L.ln + 1 : 1));
  sm.zl[r] = 0.0; sm.rl[r] = 0.0;
  sm.uh[r] = r < n ? R.w.uh : 0.0;
  sm.ue[r] = r < n ? R.w.ue : 0.0;
  if (r < BMAX) { sm.zl[NT + r] = 0.0; sm.rl[NT + r] = 0.0; sm.uh[NT + r] = 0.0; sm.ue[NT + r] = 0.0; }

  const double rtol = I.rtol, atol = I.atol;
  const int n_save = c.n_save;
  const double* __restrict__ lna_save = c.lna_save;

  double t = I.lna_start;
  double ti = fast_interp(t, ABCMB_TAU_TAB_LO, 0.0, c.tau_tab, ABCMB_TAU_TAB_N);
  double y = r < n ? ic_row(c, L, k, ti, r) : 0.0;
  set_times(sm, c, L, k, t, 0.0, 1, st.jac_offset);
  double f0 = rhs_eval(sm, R, 0, y);
  double h = st.dt0;
  const double b1 = st.pcoeff + st.icoeff + st.dcoeff, b2 = -(st.pcoeff + 2.0 * st.dcoeff);
  double prev_inv = 1.0;
  int isave = 0, steps = 0, iters = 0, status = ABCMB_K_OK;
  Inv<BMAX> FI;

  auto save_point = [&](double ys, double ls) {
    __syncthreads();
    sm.zl[r] = ys;
    __syncthreads();
    if (state_out && r < n) state_out[((size_t)item * n_save + isave) * n + r] = ys;
    if (r < ABCMB_N_PT_FIELDS) {
      double v = output_field(c, L, k, ls, sm.zl, r);
      c.pt[((size_t)r * n_save + isave) * c.nk + I.ik] = v;
    }
  };

  while (isave < n_save && lna_save[isave] <= t) { save_point(y, lna_save[isave]); ++isave; }

  while (t < 0.0) {
    if (steps >= st.max_steps) { status = ABCMB_K_MAX_STEPS; break; }
    double tn = t + h;
    if (tn > 0.0) tn = 0.0;
    const double hh = tn - t;
    if (!(hh > 0.0)) { status = ABCMB_K_STEP_UNDERFLOW; break; }
    ++steps;
    const double g = hh * KV_G;
    set_times(sm, c, L, k, t, hh, NTIMES, st.jac_offset);
    bool ok = factor(sm, R, maxb, g, FI);
    sm.Fs[0][r] = f0;
    double z = y;
#pragma unroll 1
    for (int s = 1; s < 7; ++s) {
      if (!ok) break;
      double acc = 0.0;
      for (int j = 0; j < s; ++j) acc += KV_A[s][j] * sm.Fs[j][r];
      const double base = y + hh * acc;
      z = base + g * sm.Fs[s - 1][r];
      const int slot = STAGE_SLOT[s];
      bool conv = false;
      double prev = 0.0;
      for (int it = 0; it < st.max_newton; ++it) {
        ++iters;
        double fz = rhs_eval(sm, R, slot, z);
        double dz = solve(sm, R, FI, -(z - base - g * fz));
        z += dz;
        double v[1] = {0.0};
        if (r < n) { double q = dz / (atol + rtol * fabs(z)); v[0] = q * q; }
        block_sum<NT, 1>(v, sm.red);
        double nrm = sqrt(v[0] / n);
        if (!isfinite(nrm)) break;
        if (nrm < NEWTON_SMALL) { conv = true; break; }
        if (it > 0) {
          double th = nrm / prev;
          if (th > 2.0) break;
          double fac = nrm * th / (1.0 - th);
          if (fac > 0.0 && fac < st.newton_kappa) { conv = true; break; }
        }
        prev = nrm;
      }
      if (!conv) { ok = false; break; }
      sm.Fs[s][r] = (z - base) / g;
    }
    double enorm = INFINITY;
    const double ynew = z;
    if (ok) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 7; ++j) acc += KV_E[j] * sm.Fs[j][r];
      double v[1] = {r < n ? rms_term(hh * acc, y, ynew, rtol, atol) : 0.0};
      block_sum<NT, 1>(v, sm.red);
      enorm = sqrt(v[0] / n);
      if (!isfinite(enorm)) enorm = INFINITY;
    }
    const bool accept = enorm < 1.0;
    const double inv = enorm == 0.0 ? INFINITY : 1.0 / enorm;
    const double fmin = accept ? 1.0 : st.factormin;
    double fac = st.safety * pow(inv, b1 / 5.0) * pow(prev_inv, b2 / 5.0);
    if (!(fac >= fmin)) fac = fmin;
    if (fac > st.factormax) fac = st.factormax;
    if (accept) {
      double fnew = rhs_eval(sm, R, 5, ynew);
      while (isave < n_save && lna_save[isave] <= tn) {
        double th = (lna_save[isave] - t) / hh;
        double h00 = 2 * th * th * th - 3 * th * th + 1, h10 = th * th * th - 2 * th * th + th;
        double h01 = -2 * th * th * th + 3 * th * th, h11 = th * th * th - th * th;
        double ys = h00 * y + h10 * hh * f0 + h01 * ynew + h11 * hh * fnew;
        save_point(ys, lna_save[isave]);
        ++isave;
      }
      t = tn;
      y = ynew;
      f0 = fnew;
      prev_inv = inv;
    }
    h = hh * fac;
  }
  if (r == 0) {
    status_out[item] = status;
    steps_out[item] = steps;
    if (iters_out) iters_out[item] = iters;
  }
}

template <int NT, int BMAX>
int launch(const AbcmbCosmo* cosmos, const AbcmbPertItem* items, int n_items, const AbcmbOdeSettings& st,
           int32_t* status, int32_t* steps, int32_t* iters, double* state, hipStream_t stream) {
  hipLaunchKernelGGL((pert_fwd_kernel<NT, BMAX>), dim3(n_items), dim3(NT), 0, stream, cosmos, items, st, status,
                     steps, iters, state);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}

}  // namespace
}  // namespace abcmb

extern "C" size_t abcmb_pert_workspace_size(int n_items) {
  (void)n_items;
  return 0;
}

extern "C" int abcmb_pert_fwd(const AbcmbCosmo* cosmos_dev, int n_cosmo, const AbcmbPertItem* items_dev,
                              int n_items, const AbcmbOdeSettings* settings, int32_t* status_dev,
                              int32_t* steps_dev, int32_t* iters_dev, double* state_dev, void* workspace,
                              size_t ws_bytes, void* stream) {
  using namespace abcmb;
  (void)workspace;
  (void)ws_bytes;
  ABCMB_REQUIRE(cosmos_dev && items_dev && settings && status_dev && steps_dev, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && n_items > 0, "empty launch");
  const AbcmbOdeSettings& st = *settings;
  ABCMB_REQUIRE(st.l_g >= 5 && st.l_pol >= 4 && st.l_ur >= 5, "hierarchy cut-offs too small");
  ABCMB_REQUIRE(st.n_q == 0 || ((st.n_q == 3 || st.n_q == 5) && st.l_ncdm >= 4), "bad massive-nu layout");
  int n_expect = 4 + (st.l_g + 1) + (st.l_pol + 1) + (st.l_ur + 1) + st.n_q * (st.l_ncdm + 1);
  ABCMB_REQUIRE(st.n == n_expect, "settings.n=%d does not match the layout (%d)", st.n, n_expect);
  ABCMB_REQUIRE(st.max_newton > 0 && st.max_steps > 0, "bad iteration limits");
  int maxb = 2 + (st.l_g + 1) + (st.l_pol + 1);
  if (st.l_ur + 1 > maxb) maxb = st.l_ur + 1;
  if (st.n_q > 0 && st.l_ncdm + 1 > maxb) maxb = st.l_ncdm + 1;
  ABCMB_REQUIRE(maxb <= 64, "a hierarchy block of %d rows exceeds 64", maxb);
  hipStream_t s = (hipStream_t)stream;
  if (maxb <= 32) {
    if (st.n <= 64) return launch<64, 32>(cosmos_dev, items_dev, n_items, st, status_dev, steps_dev, iters_dev, state_dev, s);
    if (st.n <= 128) return launch<128, 32>(cosmos_dev, items_dev, n_items, st, status_dev, steps_dev, iters_dev, state_dev, s);
    if (st.n <= 192) return launch<192, 32>(cosmos_dev, items_dev, n_items, st, status_dev, steps_dev, iters_dev, state_dev, s);
  } else {
    if (st.n <= 64) return launch<64, 64>(cosmos_dev, items_dev, n_items, st, status_dev, steps_dev, iters_dev, state_dev, s);
    if (st.n <= 128) return launch<128, 64>(cosmos_dev, items_dev, n_items, st, status_dev, steps_dev, iters_dev, state_dev, s);
    if (st.n <= 192) return launch<192, 64>(cosmos_dev, items_dev, n_items, st, status_dev, steps_dev, iters_dev, state_dev, s);
  }
  set_error("state size %d exceeds 192", st.n);
  return ABCMB_EINVAL;
}
