// Shared device code of the perturbation ODE kernels (forward, tangent-linear
// and adjoint): the Kvaerno5 tableau, background and the scalars of A(ln a),
// the row descriptors, the block Gauss-Jordan chord factorisation, the
// Woodbury solve, the initial conditions, the output-table fields and one
// integration of a work item (ode_item).  Included by pert.hip and adjoint.hip;
// everything is in an anonymous namespace, so each includes its own copy.
//
// Perturbation ODE kernel: one (cosmology, k-mode) per workgroup, FP64.
//
// Integrates the linear Einstein-Boltzmann hierarchy of ABCMB
// (abcmb/perturbations.py:201-248 and the species y_prime functions in
// abcmb/species.py) from the per-k start time to ln a = 0 with the
// Kvaerno5 ESDIRK method, PID step control and chord Newton
// (abcmb/perturbations.py:250-316), writes cubic-Hermite dense output at the
// cosmology's saved ln a points and assembles the perturbation-table fields
// (abcmb/perturbations.py:318-404): from the saved state rows in a second
// pass (fields_cosmo_kernel / fields_kernel) where the rows fit the table's
// slots, otherwise in the kernel at each saved point.
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
//   - the background scalars at the step's seven times are evaluated in
//     parallel, one lane per time, once per step;
//   - state and error live in registers of the row's thread, the stage
//     derivatives in an LDS column per thread; workgroup sums (metric sums,
//     Woodbury sums, Newton and error norms) are transposed wavefront sums
//     (half-wave swaps, then DPP within rows) plus one LDS slot per
//     (value, wavefront) and one barrier.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "common.h"
#include "dual.h"
#include "wave.h"

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

// massive-neutrino momentum quadrature: nodes q, weights w, and 1 + e^-q at
// the nodes (the occupation factor's numerator), folded into the weights the
// background and the source terms use, so that no step evaluates exp(-q)
constexpr double Q3V[3] = {0.913201, 3.37517, 7.79184};
constexpr double W3V[3] = {0.0687359, 3.31435, 2.29911};
constexpr double E3V[3] = {1.4012378040061892, 1.0342123017259768, 1.0004130920948129};
constexpr double Q5V[5] = {0.583165, 2.0, 4.0, 7.26582, 13.0};
constexpr double W5V[5] = {0.0081201, 0.689407, 2.8063, 2.05156, 0.12681};
constexpr double E5V[5] = {1.5581290895929683, 1.1353352832366128, 1.0183156388887342, 1.0006990278317496,
                           1.000002260329407};
__constant__ double Q3[3] = {Q3V[0], Q3V[1], Q3V[2]};
__constant__ double W3[3] = {W3V[0], W3V[1], W3V[2]};
__constant__ double Q5[5] = {Q5V[0], Q5V[1], Q5V[2], Q5V[3], Q5V[4]};
__constant__ double W5[5] = {W5V[0], W5V[1], W5V[2], W5V[3], W5V[4]};
#define ABCMB_WR(Q, W, E, i) (E[i] * W[i] / (Q[i] * Q[i]))
// w (1 + e^-q) / q^2 (density: times e = sqrt(q^2 + x^2)) and w (1 + e^-q) (pressure: over e)
__constant__ double WR5[5] = {ABCMB_WR(Q5V, W5V, E5V, 0), ABCMB_WR(Q5V, W5V, E5V, 1), ABCMB_WR(Q5V, W5V, E5V, 2),
                              ABCMB_WR(Q5V, W5V, E5V, 3), ABCMB_WR(Q5V, W5V, E5V, 4)};
__constant__ double WP5[5] = {E5V[0] * W5V[0], E5V[1] * W5V[1], E5V[2] * W5V[2], E5V[3] * W5V[3], E5V[4] * W5V[4]};
__constant__ double WR3[3] = {ABCMB_WR(Q3V, W3V, E3V, 0), ABCMB_WR(Q3V, W3V, E3V, 1), ABCMB_WR(Q3V, W3V, E3V, 2)};
#undef ABCMB_WR


// ---------------------------------------------------------------- cosmology fields
// AbcmbCosmo's double fields in declaration order; the tangent inputs index
// them the same way.
enum CosmoField {
  F_H = 0, F_OMEGA_B, F_OMEGA_CDM, F_OMEGA_LAMBDA, F_A_S, F_N_S, F_K_PIVOT, F_TCMB0, F_YHE,
  F_N_UR, F_T_UR, F_N_NCDM, F_T_NCDM, F_M_NCDM, F_R_NU, F_R_B, F_OM, F_OMEGA_R, F_OMEGA_M, F_TAU0,
  F_XE_X0, F_XE_DX, F_XE_LNA_LAST, F_XE_LAST, F_TM_X0, F_TM_DX, F_TM_LNA_LAST, F_TM_LAST,
  F_LNA_REC, F_LNA_TRANSFER_START, F_Z_REION, F_FLUID0,
  F_COUNT = F_FLUID0 + ABCMB_FLUID_PARAMS
};
static_assert(F_COUNT == ABCMB_COSMO_DOUBLES, "field list out of step with AbcmbCosmo");
static_assert(offsetof(AbcmbCosmo, fluid_par) == F_FLUID0 * sizeof(double), "AbcmbCosmo doubles are not packed");

// ---------------------------------------------------------------- user fluids
// The model's user-defined fluids (abcmb/species.py:11-249) arrive as
// generated code (abcmb_amd/fluidgen.py) in namespace fluid: their summed
// density and pressure, their rows' terms (columns, static coefficients and
// background scalar slots), their h'/eta' couplings, initial conditions and
// contributions to the metric sums.  The default header has none, and every
// use below is behind `if constexpr (fluid::kActive)` / `fluid::kRows > 0`,
// so the standard kernels compile exactly as without this hook.  The helpers
// give the generated expressions one spelling for plain doubles and duals.
namespace fluid {
__device__ __forceinline__ double fexp(double x) { return exp(x); }
__device__ __forceinline__ double flog(double x) { return log(x); }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double fpow(double x, double p) { return pow(x, p); }
template <int N> __device__ __forceinline__ D<N> fexp(const D<N>& x) { return dexp(x); }
template <int N> __device__ __forceinline__ D<N> flog(const D<N>& x) { return dlog(x); }
template <int N> __device__ __forceinline__ D<N> fsqrt(const D<N>& x) { return dsqrt(x); }
template <int N> __device__ __forceinline__ D<N> fpow(const D<N>& x, double p) { return dpow(x, p); }
}  // namespace fluid
#ifndef ABCMB_FLUID_DEV
#define ABCMB_FLUID_DEV "fluid_none_dev.h"
#endif
#include ABCMB_FLUID_DEV

// background scalar slots (per evaluation time); the user fluids' slots follow
enum {
  SC_ONE = 0, SC_IAH, SC_ITC, SC_ITAU, SC_CS2, SC_RTC,
  SC_H_ETA, SC_H_C, SC_H_B, SC_H_G, SC_H_U,
  SC_E_B, SC_E_G, SC_E_U,
  SC_P0,            // 5 slots: q_i / (eps_i aH)
  SC_HN0 = SC_P0 + 5,  // 5 slots: h'-weight of Psi0_i
  SC_EN = SC_HN0 + 5,  // eta'-weight base of kPsi1_i
  SC_ZERO,
  SC_FL0,              // fluid::kSlots slots of the user fluids' rows
  SC_COUNT = SC_FL0 + fluid::kSlots
};

// A cosmology seen by the kernels: its fields and tables, with their
// derivatives along N parameter directions when N > 0.
template <int N>
struct Cosmo {
  const AbcmbCosmo& c;
  const AbcmbTangent* t;  // null when N == 0
  __device__ __forceinline__ D<N> operator()(int f) const {
    D<N> r(reinterpret_cast<const double*>(&c)[f]);
    if constexpr (N > 0) {
      ABCMB_DLOOP(N) r.d[i_] = t->par[i_ * ABCMB_COSMO_DOUBLES + f];
    }
    return r;
  }
  __device__ __forceinline__ const double* dtau() const { if constexpr (N > 0) return t->tau_tab; else return nullptr; }
  __device__ __forceinline__ const double* dxe() const { if constexpr (N > 0) return t->xe_tab; else return nullptr; }
  __device__ __forceinline__ const double* dtm() const { if constexpr (N > 0) return t->tm_tab; else return nullptr; }
};

// table entry i of f (and of its tangent tables fd[N][n])
template <int N>
__device__ __forceinline__ D<N> tab_at(const double* __restrict__ f, const double* __restrict__ fd, int n, int i) {
  D<N> r(f[i]);
  if constexpr (N > 0) { ABCMB_DLOOP(N) r.d[i_] = fd[i_ * n + i]; }
  return r;
}

// ---------------------------------------------------------------- background
// linear interpolation on a uniform grid with the reference's end clipping
// (abcmb/ABCMBTools.py:268-308); the abscissa does not depend on the parameters
template <int N>
__device__ __forceinline__ D<N> fast_interp(double x, double lo, double hi, const double* __restrict__ f,
                                            const double* __restrict__ fd, int n) {
  double i = (x - lo) / (hi - lo) * (n - 1);
  const double eps = 1e-6;
  i = fmin(fmax(i, eps), n - 1.0 - eps);
  int il = (int)floor(i);
  int iu = min(il + 1, n - 1);
  double wu = i - il;
  D<N> r((1.0 - wu) * f[il] + wu * f[iu]);
  if constexpr (N > 0) { ABCMB_DLOOP(N) r.d[i_] = (1.0 - wu) * fd[i_ * n + il] + wu * fd[i_ * n + iu]; }
  return r;
}

template <int N>
struct Bg {
  double a;  // exp(ln a): the time grid does not move with the parameters
  D<N> aH, tau, tauc, cs2, R;
  D<N> rc, rb, rg, ru, rn, pn;
  D<N> Tn, xn;
};

// densities, aH (abcmb/background.py:101-187; species rho at abcmb/species.py).
// Quotients by a, by dual quantities and by constants are written as products
// with reciprocals (one division per reciprocal), which moves results by at
// most a few units in the last place.
template <int N>
__device__ void bg_rho(const Cosmo<N>& P, double lna, Bg<N>& b) {
  constexpr double hc3 = (C_CM * HBAR) * (C_CM * HBAR) * (C_CM * HBAR);
  constexpr double K_RHO = 3.0 * H0_OVER_H * H0_OVER_H / 8.0 / PI / GN;  // critical density per h^2
  constexpr double K_G = PI * PI / 15.0 / hc3;
  constexpr double K_U = 2.0 * 7.0 / 8.0 * PI * PI / 30.0 / hc3;
  constexpr double K_N = 1.0 / (PI * PI) / (HBAR * HBAR * HBAR) / (C_CM * C_CM * C_CM);
  constexpr double K_H = 8.0 * PI * GN / 3.0;
  const double a = exp(lna), ia = 1.0 / a;
  const double ia3 = ia * ia * ia, ia4 = ia3 * ia;
  const D<N> T0 = P(F_TCMB0), Tu = P(F_T_UR);
  const D<N> t04 = T0 * T0 * T0 * T0;
  const D<N> tu4 = Tu * Tu * Tu * Tu;
  b.a = a;
  const int off = P.c.species_off;
  const D<N> rde = (off & ABCMB_NO_LAMBDA) ? D<N>(0.0) : P(F_OMEGA_LAMBDA) * K_RHO;
  b.rc = (off & ABCMB_NO_CDM) ? D<N>(0.0) : P(F_OMEGA_CDM) * (K_RHO * ia3);
  b.rb = P(F_OMEGA_B) * (K_RHO * ia3);
  b.rg = t04 * (K_G * ia4);
  b.ru = P(F_N_UR) * tu4 * t04 * (K_U * ia4);
  b.rn = D<N>(0.0);
  b.pn = D<N>(0.0);
  b.Tn = P(F_T_NCDM) * T0 * ia;
  b.xn = P(F_M_NCDM) * dinv(b.Tn);
  if (P.c.has_ncdm) {
    D<N> ir(0.0), ip(0.0);
    const D<N> x2 = b.xn * b.xn;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const double q = Q5[i];
      D<N> e = dsqrt(x2 + q * q);
      ir += WR5[i] * e;
      ip += WP5[i] * dinv(e);
    }
    D<N> pre = b.Tn * b.Tn * b.Tn * b.Tn * K_N;
    b.rn = P(F_N_NCDM) * 4.0 * pre * ir;
    b.pn = P(F_N_NCDM) * (4.0 / 3.0) * pre * ip;
  }
  D<N> rt = rde + b.rc + b.rb + b.rg + b.ru + b.rn;
  if constexpr (fluid::kActive) {
    D<N> rf, pf, rm;
    fluid::bg<N>(P, lna, a, rf, pf, rm);
    rt += rf;
  }
  b.aH = a * dsqrt(K_H * rt) * (1.0 / C_MPC);
}

// x_e lookup with the reference's span quirk (abcmb/background.py:568-603)
template <int N>
__device__ __forceinline__ D<N> bg_xe(const Cosmo<N>& P, double lna) {
  const AbcmbCosmo& c = P.c;
  if (lna < c.xe_x0) return tab_at<N>(c.xe_tab, P.dxe(), c.xe_n, 0);
  if (lna >= c.xe_lna_last) return P(F_XE_LAST);
  return fast_interp<N>(lna, c.xe_x0, c.xe_x0 + c.xe_n * c.xe_dx, c.xe_tab, P.dxe(), c.xe_n);
}

template <int N>
__device__ __forceinline__ void bg_eval(const Cosmo<N>& P, double lna, Bg<N>& b) {
  constexpr double K_NH = 3.0 * H0_OVER_H * H0_OVER_H / 8 / PI / GN / MH;
  constexpr double K_TC = C_MPC / SIGMA_T / C_CM;
  constexpr double STEF = PI * PI / (60.0 * HBAR * HBAR * HBAR * C_CM * C_CM);
  constexpr double K_GC = 8.0 * SIGMA_T * (4.0 * STEF) / (3.0 * ME);
  const AbcmbCosmo& c = P.c;
  bg_rho(P, lna, b);
  const double ia = 1.0 / b.a;
  b.tau = fast_interp<N>(lna, ABCMB_TAU_TAB_LO, 0.0, c.tau_tab, P.dtau(), ABCMB_TAU_TAB_N);
  D<N> xe = bg_xe(P, lna);
  const D<N> YHe = P(F_YHE);
  D<N> nH = (1 - YHe) * P(F_OMEGA_B) * (K_NH * ia * ia * ia);
  b.tauc = dinv(nH * xe) * (ia * K_TC);
  // T_m (abcmb/background.py:605-657)
  D<N> Tg = P(F_TCMB0) * ia;
  D<N> Tm;
  if (lna < c.tm_x0) {
    D<N> fhe = YHe * (MH / MHE) * dinv(1 - YHe);
    D<N> Tg2 = Tg * Tg;
    D<N> gc = xe * dinv(1.0 + fhe + xe) * Tg2 * Tg2 * K_GC;
    D<N> H = b.aH * (C_MPC * ia);
    Tm = Tg * (1.0 - H * dinv(gc));
  } else if (lna >= c.tm_lna_last) {
    Tm = P(F_TM_LAST);
  } else {
    Tm = fast_interp<N>(lna, c.tm_x0, c.tm_x0 + c.tm_n * c.tm_dx, c.tm_tab, P.dtm(), c.tm_n);
  }
  // c_s^2 (abcmb/species.py:1163-1220)
  D<N> mu = MH * dinv((1.0 + xe) * (1.0 - YHe) + (MH / MHE) * YHe);
  b.R = (4.0 / 3.0) * b.rg * dinv(b.rb);
  b.cs2 = Tm * dinv(mu) *
          (5.0 / 3.0 - (2.0 / 3.0 / ME) * mu * b.R * dinv(b.aH) * dinv(b.tauc) * (Tg * dinv(Tm) - 1.0));
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
  int ifl, nfl;  // user-fluid rows: the last nfl rows, from ifl
};

__device__ __forceinline__ Layout make_layout(const AbcmbOdeSettings& st) {
  Layout L;
  L.lg = st.l_g; L.lp = st.l_pol; L.lu = st.l_ur; L.ln = st.l_ncdm; L.nq = st.n_q; L.n = st.n;
  L.ig = 4; L.ip = 4 + L.lg + 1; L.iu = L.ip + L.lp + 1; L.inc = L.iu + L.lu + 1;
  L.nfl = st.n_fluid; L.ifl = L.n - L.nfl;
  return L;
}

__device__ __forceinline__ void row_desc(const Layout& L, double k, int r, Row& w) {
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
  if (r >= L.inc && r < L.ifl) {  // massive neutrinos, bin i
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
  if constexpr (fluid::kRows > 0) {
    if (r >= L.ifl && r < L.n) {  // user fluids: columns local to their block, slots after SC_FL0
      int hs = -1, es = -1;
      fluid::row(r - L.ifl, k, [&](int c, double v, int sl) { add(L.ifl + c, v, sl < 0 ? SC_ONE : SC_FL0 + sl); },
                 w.uh, w.ue, hs, es);
      if (hs >= 0) { w.whi = SC_FL0 + hs; w.whs = 1.0; }
      if (es >= 0) { w.wei = SC_FL0 + es; w.wes = 1.0; }
    }
  }
}

// row_desc out of line: the value-only kernel keeps the descriptor in a stack
// object (three waves per SIMD leave no room for it in registers); the tangent
// kernel inlines it, so the register allocator keeps the descriptor in
// registers or spills it to AGPRs instead of reloading it from scratch in the
// chord loop
__device__ __noinline__ void row_desc_call(const Layout& L, double k, int r, Row& w) { row_desc(L, k, r, w); }


// scalars of A(ln a) shared by all rows: sc[c][slot] holds the value (c = 0)
// and the derivative along direction c - 1
template <int N>
__device__ __forceinline__ void fill_scalars(const Cosmo<N>& P, const Layout& L, double k, const Bg<N>& b,
                                             double* val, double* der) {
  constexpr double K_G = 4.0 * PI * GN / (C_MPC * C_MPC);
  constexpr double K_N = 1.0 / (PI * PI) / (HBAR * HBAR * HBAR) / (C_CM * C_CM * C_CM);
  // the value into val[s] (when val is given), derivative i into der[i * SC_COUNT + s]
  auto put = [&](int s, const D<N>& x) {
    if (val) val[s] = x.v;
    if constexpr (N > 0) { ABCMB_DLOOP(N) der[i_ * SC_COUNT + s] = x.d[i_]; }
  };
  const D<N> iah = dinv(b.aH);
  const D<N> itc = iah * dinv(b.tauc);
  put(SC_ONE, D<N>(1.0));
  put(SC_IAH, iah);
  put(SC_ITC, itc);
  put(SC_ITAU, iah * dinv(b.tau));
  put(SC_CS2, b.cs2 * iah);
  put(SC_RTC, b.R * itc);
  const D<N> iah2 = iah * iah;
  const double a2 = b.a * b.a;
  D<N> kh = iah2 * (2.0 * K_G * a2);
  put(SC_H_ETA, iah2 * (2.0 * k * k));
  put(SC_H_C, kh * b.rc);
  put(SC_H_B, kh * b.rb);
  put(SC_H_G, kh * b.rg);
  put(SC_H_U, kh * b.ru);
  D<N> ke = iah * (K_G * a2 / (k * k));
  put(SC_E_B, ke * b.rb);
  put(SC_E_G, ke * (b.rg + b.rg / 3.0));
  put(SC_E_U, ke * (b.ru + b.ru / 3.0));
  D<N> pre(0.0);
  if (P.c.has_ncdm) pre = P(F_N_NCDM) * b.Tn * b.Tn * b.Tn * b.Tn * K_N;
  const D<N> x2 = b.xn * b.xn;
  const D<N> khp = kh * 4.0 * pre;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    if (i < L.nq) {
      const double q = L.nq == 5 ? Q5[i] : Q3[i];
      const double wr = L.nq == 5 ? WR5[i] : WR3[i];
      D<N> e = dsqrt(x2 + q * q);
      put(SC_P0 + i, q * dinv(e) * iah);
      put(SC_HN0 + i, khp * (wr * e));
    } else {  // no such bin: no row reads these slots
      put(SC_P0 + i, D<N>(0.0));
      put(SC_HN0 + i, D<N>(0.0));
    }
  }
  put(SC_EN, ke * 4.0 * pre);
  put(SC_ZERO, D<N>(0.0));
  if constexpr (fluid::kSlots > 0) {
    D<N> fs[fluid::kSlots > 0 ? fluid::kSlots : 1];
    fluid::slots<N>(P, log(b.a), b.a, k, b.aH, kh, ke, fs);
#pragma unroll
    for (int q = 0; q < fluid::kSlots; ++q) put(SC_FL0 + q, fs[q]);
  }
}

// the background and the scalars of A at one time, out of line: one call per
// time and lane keeps the kernel's register budget for the chord loop, and
// the background stays in the callee's registers instead of a stack object.
// sc: the scalar rows of this time, [1 + N][SC_COUNT].
template <int N>
__device__ __noinline__ void time_scalars(const Cosmo<N>& P, const Layout& L, double k, double lna,
                                          double (*sc)[SC_COUNT]) {
  Bg<N> b;
  bg_eval(P, lna, b);
  fill_scalars(P, L, k, b, sc[0], N > 0 ? sc[1] : nullptr);
}

// the same along one tangent direction d of a launch with N directions: the
// parameters and tables of direction d as a one-direction dual (the value
// arithmetic is the same operation sequence as in the N-direction
// evaluation), the value row written by direction 0 only
template <int N>
__device__ __noinline__ void time_scalars_dir(const AbcmbCosmo& c, const AbcmbTangent* T, int d, const Layout& L,
                                              double k, double lna, double (*sc)[SC_COUNT]) {
  AbcmbTangent td;
  td.par = T->par + (size_t)d * ABCMB_COSMO_DOUBLES;
  td.tau_tab = T->tau_tab + (size_t)d * ABCMB_TAU_TAB_N;
  td.xe_tab = T->xe_tab + (size_t)d * c.xe_n;
  td.tm_tab = T->tm_tab + (size_t)d * c.tm_n;
  td.pt = nullptr;
  const Cosmo<1> P1{c, &td};
  Bg<1> b;
  bg_eval(P1, lna, b);
  fill_scalars(P1, L, k, b, d == 0 ? sc[0] : nullptr, sc[1 + d]);
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
//
// With N tangent directions every state quantity is a dual number D<N>.  A
// depends on the parameters only through the background scalars, so the
// tangent of a stage solves (I - g A) zdot = basedot + g Adot z; it is
// iterated with the same chord matrix alongside the value (and beyond, until
// it has converged too).  Step sizes and acceptance are decided on the value
// alone, which stays identical to the primal kernel's.

// chord updates below this (in units of the error tolerance) are rounding noise
constexpr double NEWTON_SMALL = 1e-9;
// Tangents: once the value has converged, the tangents keep iterating (value
// fixed) until each direction's estimated remaining error -- its weighted
// update times rate / (1 - rate), with the value's contraction rate, since
// value and tangents run the same linear iteration -- is below TANGENT_RTOL
// of its weighted size (the size of the stage predictor's tangent, summed
// once per stage).  They then solve the differentiated stage equations
// rather than the derivative of a truncated chord iteration.  At most
// TANGENT_EXTRA iterations beyond max_newton.
constexpr double TANGENT_RTOL = 1e-6;
constexpr int TANGENT_EXTRA = 8;
constexpr int NTIMES = 7;  // t, the distinct stage abscissae t + c_s h, and the chord-Jacobian time
constexpr int JAC_SLOT = 6;
__constant__ double TIME_C[NTIMES - 1] = {0.0, 0.52, 1.230333209967908, 0.8957659843500759, 0.43639360985864756,
                                          1.0};
__constant__ int STAGE_SLOT[7] = {0, 1, 2, 3, 4, 5, 5};
// blocks: eta, delta_c, baryon-photon, massless nu, up to 5 massive-nu bins,
// the user fluids (when the build has any), and a last id for rows past n
constexpr int NBLK = fluid::kRows > 0 ? 11 : 10;

struct Blk {
  int id, start, size, lr;
};

__device__ Blk row_block(const Layout& L, int r) {
  Blk b;
  if (r == 0) { b.id = 0; b.start = 0; b.size = 1; }
  else if (r == 1) { b.id = 1; b.start = 1; b.size = 1; }
  else if (r < L.iu) { b.id = 2; b.start = 2; b.size = L.iu - 2; }
  else if (r < L.inc) { b.id = 3; b.start = L.iu; b.size = L.lu + 1; }
  else if (r < L.ifl) {
    int i = (r - L.inc) / (L.ln + 1);
    b.id = 4 + i; b.start = L.inc + i * (L.ln + 1); b.size = L.ln + 1;
  } else if (r < L.n) { b.id = 9; b.start = L.ifl; b.size = L.nfl; }
  else { b.id = NBLK - 1; b.start = r; b.size = 0; }
  b.lr = r - b.start;
  return b;
}

template <int NT, int BMAX, int N>
struct Smem {
  static constexpr int NC = 1 + N;                        // value + tangents
  // widest workgroup sum: the Newton test's 1 + N norms plus the 2(1 + N)
  // metric-sum increments, or a stage's 2(1 + N) metric sums plus N norms
  static constexpr int NRED = (3 + 3 * N > 8 ? 3 + 3 * N : 8);
  double sc[NTIMES][NC][SC_COUNT];
  // z by component, rows padded to a multiple of 64 so that a component's
  // column sits 512-byte strides from the next (paired reads, no address math)
  static constexpr int LDZ = (NT + BMAX + 63) / 64 * 64;
  double zl[NC][LDZ];
  // right-hand sides of the block solves by row, the components of a row
  // adjacent: one base address serves every (column, component) of a block
  alignas(16) double rl[(NT + BMAX) * NC];
  double uh[NT + BMAX];
  double ue[NT + BMAX];
  double pbuf[2][NBLK][2][BMAX];   // [buffer][block][pivot row of the pair][column]
  double red[2][NRED * (NT / 64)];  // alternating halves: one barrier per sum
  double Fs[7][NC][NT];  // stage derivatives, one column per thread
};

// workgroup sum of NV values: transposed wavefront sums (the row totals sit
// in lanes 0, 16, 32, 48), one LDS slot per (value, wave), one barrier, and
// every thread adds the waves' slots.  red holds NV * NT / 64 doubles; the
// caller alternates between two such buffers, so a buffer is written again
// only after the barrier of the sum in between, which every thread reaches
// after its reads of this one.  The barrier also orders any LDS writes made
// before the call against reads made after it.
template <int NT, int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red) {
  constexpr int NW = NT / 64;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double w[RowSums<NV>::H2];
  wave_sum_rows(v, w);
  if ((lane & 15) == 0) {
    const int rho = lane >> 4;
#pragma unroll
    for (int q = 0; q < RowSums<NV>::H2; ++q) {
      const int i = row_sum_index<NV>(q, rho);
      if (i >= 0) red[i * NW + wv] = w[q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NW; ++i) s += red[q * NW + i];
    v[q] = s;
  }
}

// block_sum on the workgroup's reduction buffers, checked against their size;
// rs selects the half and flips
template <int NT, int BMAX, int N, int NV>
__device__ __forceinline__ void wg_sum(Smem<NT, BMAX, N>& sm, int& rs, double (&v)[NV]) {
  static_assert(NV <= Smem<NT, BMAX, N>::NRED, "reduction wider than the LDS buffer");
  block_sum<NT, NV>(v, sm.red[rs]);
  rs ^= 1;
}

// background scalars of A at t + TIME_C[i] * hh, i < cnt (one lane per time)
template <int NT, int BMAX, int N>
__device__ void set_times(Smem<NT, BMAX, N>& sm, const Cosmo<N>& P, const Layout& L, double k, double t, double hh,
                          int cnt, double jac_offset) {
  __syncthreads();
  if constexpr (N == 0) {
    const int r = threadIdx.x;
    if (r < cnt) time_scalars(P, L, k, t + (r == JAC_SLOT ? jac_offset : TIME_C[r]) * hh, sm.sc[r]);
  } else {
    // one lane per (time, direction): a one-direction dual each
    static_assert(NTIMES * N <= 64, "set_times: one wavefront holds every (time, direction) lane");
    const int r = threadIdx.x / N, d = threadIdx.x % N;
    if (threadIdx.x < cnt * N)
      time_scalars_dir<N>(P.c, P.t, d, L, k, t + (r == JAC_SLOT ? jac_offset : TIME_C[r]) * hh, sm.sc[r]);
  }
  __syncthreads();
}

struct RowState {
  Row w;
  Blk B;
  int lc[5];  // term columns local to the block
  int n;
};

// x into column r of sm.zl
template <int NT, int BMAX, int N>
__device__ __forceinline__ void put_z(Smem<NT, BMAX, N>& sm, int r, const D<N>& x) {
  sm.zl[0][r] = x.v;
  if constexpr (N > 0) { ABCMB_DLOOP(N) sm.zl[1 + i_][r] = x.d[i_]; }
}

// x into row r of a row-major buffer of 1 + N components per row
template <int N>
__device__ __forceinline__ void put_row(double* buf, int r, const D<N>& x) {
  buf[r * (1 + N)] = x.v;
  if constexpr (N > 0) { ABCMB_DLOOP(N) buf[r * (1 + N) + 1 + i_] = x.d[i_]; }
}

template <int NT, int BMAX, int N>
__device__ __forceinline__ D<N> get_F(const Smem<NT, BMAX, N>& sm, int s, int r) {
  D<N> x;
  x.v = sm.Fs[s][0][r];
  if constexpr (N > 0) { ABCMB_DLOOP(N) x.d[i_] = sm.Fs[s][1 + i_][r]; }
  return x;
}

template <int NT, int BMAX, int N>
__device__ __forceinline__ void put_F(Smem<NT, BMAX, N>& sm, int s, int r, const D<N>& x) {
  sm.Fs[s][0][r] = x.v;
  if constexpr (N > 0) { ABCMB_DLOOP(N) sm.Fs[s][1 + i_][r] = x.d[i_]; }
}

// this row's terms of the metric sums h' and eta' at slot, for the value and
// each tangent: m[0] = w_h sc z, m[1] = w_e sc z, m[2 + 2i], m[3 + 2i] their
// derivatives along direction i (linear in z, so an update dz adds the terms
// of dz)
template <int NT, int BMAX, int N>
__device__ __forceinline__ void metric_terms(const Smem<NT, BMAX, N>& sm, const RowState& R, int slot,
                                             const D<N>& z, double* m) {
  const double* sc = sm.sc[slot][0];
  m[0] = R.w.whs * sc[R.w.whi] * z.v;
  m[1] = R.w.wes * sc[R.w.wei] * z.v;
  if constexpr (N > 0) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double* sd = sm.sc[slot][1 + i];
      m[2 + 2 * i] = R.w.whs * (sd[R.w.whi] * z.v + sc[R.w.whi] * z.d[i]);
      m[3 + 2 * i] = R.w.wes * (sd[R.w.wei] * z.v + sc[R.w.wei] * z.d[i]);
    }
  }
}

// f_r(t_slot, z) from z in sm.zl and the workgroup's metric sums ms: value
// sum_j cs_j sc_j z_j + u_h h' + u_e eta'; each tangent adds the derivative
// of the scalars times the value state
template <int NT, int BMAX, int N>
__device__ __forceinline__ D<N> rhs_gather(const Smem<NT, BMAX, N>& sm, const RowState& R, int slot,
                                           const double* ms) {
  const int r = threadIdx.x;
  const double* sc = sm.sc[slot][0];
  D<N> f;
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < 5; ++j) acc += R.w.cs[j] * sc[R.w.si[j]] * sm.zl[0][R.w.col[j]];
  acc += R.w.uh * ms[0] + R.w.ue * ms[1];
  f.v = r < R.n ? acc : 0.0;
  if constexpr (N > 0) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double* sd = sm.sc[slot][1 + i];
      double ad = 0.0;
#pragma unroll
      for (int j = 0; j < 5; ++j)
        ad += R.w.cs[j] * (sd[R.w.si[j]] * sm.zl[0][R.w.col[j]] + sc[R.w.si[j]] * sm.zl[1 + i][R.w.col[j]]);
      ad += R.w.uh * ms[2 + 2 * i] + R.w.ue * ms[3 + 2 * i];
      f.d[i] = r < R.n ? ad : 0.0;
    }
  }
  return f;
}

// z to sm.zl and the metric sums of z into ms (one workgroup sum); the sum's
// barrier orders the zl writes before rhs_gather's reads
template <int NT, int BMAX, int N>
__device__ __forceinline__ void stage_state(Smem<NT, BMAX, N>& sm, const RowState& R, int& rs, int slot,
                                            const D<N>& z, double (&ms)[2 * (1 + N)]) {
  put_z(sm, threadIdx.x, z);
  metric_terms(sm, R, slot, z, ms);
  wg_sum(sm, rs, ms);
}

// f_r(t_slot, z) with its own metric sums
template <int NT, int BMAX, int N>
__device__ __forceinline__ D<N> rhs_eval(Smem<NT, BMAX, N>& sm, const RowState& R, int& rs, int slot,
                                         const D<N>& z) {
  double ms[2 * (1 + N)];
  stage_state(sm, R, rs, slot, z, ms);
  return rhs_gather(sm, R, slot, ms);
}

template <int BMAX>
struct Inv {
  double row[BMAX];   // this thread's row of (I - g B_blk)^-1 (block-local columns)
  double Yh, Ye;      // rows of S^-1 g u_h, S^-1 g u_e
  double c00, c01, c10, c11;  // inverse capacitance
  double Vh, Ve;      // this row's weights in h', eta' at t_n
};

// orders LDS writes before reads among the lanes of one wavefront: a
// wavefront's LDS operations are performed in issue order, so only the
// compiler has to be kept from moving them across this point
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// build and invert S = I - g B(t_n) block-wise, then the Woodbury terms.
// wl: every hierarchy block lies inside one wavefront, so a pivot row is only
// read by lanes of its own wavefront and the pivot steps need no workgroup
// barrier
template <int NT, int BMAX, int N>
__device__ bool factor(Smem<NT, BMAX, N>& sm, const RowState& R, int& rs, int maxb, bool wl, double g,
                       Inv<BMAX>& F) {
  const double* sc = sm.sc[JAC_SLOT][0];
  const Blk& B = R.B;
  if constexpr (7 * (1 + N) >= BMAX) {
    // this row of S, scattered into a row of LDS and read back: the stage
    // derivatives are dead until the stages start, and their buffer holds a
    // row of BMAX doubles per thread (the workgroup sum at the end of this
    // function orders the read-back before any stage writes them)
    double* row = &sm.Fs[0][0][0] + threadIdx.x * BMAX;
#pragma unroll
    for (int c = 0; c < BMAX; ++c) row[c] = (c == B.lr) ? 1.0 : 0.0;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const int cj = R.lc[j];
      if (cj >= 0 && cj < BMAX) row[cj] -= g * R.w.cs[j] * sc[R.w.si[j]];
    }
#pragma unroll
    for (int c = 0; c < BMAX; ++c) F.row[c] = row[c];
  } else {
#pragma unroll
    for (int c = 0; c < BMAX; ++c) F.row[c] = (c == B.lr) ? 1.0 : 0.0;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      double v = g * R.w.cs[j] * sc[R.w.si[j]];
#pragma unroll
      for (int c = 0; c < BMAX; ++c) F.row[c] -= (R.lc[j] == c) ? v : 0.0;
    }
  }
  double bad = 0.0;
  // pivots three at a time where the stage-derivative buffer (dead during the
  // factor) holds the double-buffered pivot rows after the scattered rows
  constexpr bool TRIPLES = N > 0 && 7 * (1 + N) >= BMAX &&
                           7 * (1 + N) * NT >= NT * BMAX + 2 * NBLK * 3 * BMAX;
  if constexpr (TRIPLES) {
    // Gauss-Jordan in steps of k = min(3, rows left in the block) pivots: the
    // k pivot rows P_m are exchanged at once, their k x k pivot block D is
    // inverted in registers (padded to 3 x 3 with the identity), and every row
    // of the block becomes row + sum_m b_m P_m outside the step's columns
    // (rows outside the step: b = -t, t = (row_p .. row_p+k-1) D^-1, which
    // also fills the step's columns with -t; pivot row m: b = D^-1 row m - e_m,
    // the step's columns D^-1 row m) -- the same in-place inverse as k single
    // pivots, and the singularity test checks the pivots D's elimination meets,
    // which are the ones sequential elimination would see.
    double* const pbase = &sm.Fs[0][0][0] + NT * BMAX;
#pragma unroll
    for (int p = 0; p < BMAX; p += 3) {
      if (p < maxb) {
        double* pb = pbase + (((p / 3) & 1) * NBLK + B.id) * 3 * BMAX;
        const bool in = p < B.size;
        const int kk = in ? min(3, B.size - p) : 0;
        if (in && B.lr >= p && B.lr < p + kk) {   // the pivot rows as they stand
          double* dst = pb + (B.lr - p) * BMAX;
#pragma unroll
          for (int c = 0; c < BMAX; ++c) dst[c] = F.row[c];
        }
        if (wl) wave_sync();
        else __syncthreads();
        if (in && B.lr < B.size) {
          // rows past the step's k read row 0 again (finite; their weights are 0)
          const double* s1 = pb + (kk > 1 ? BMAX : 0);
          const double* s2 = pb + (kk > 2 ? 2 * BMAX : 0);
          double p0[BMAX], p1[BMAX], p2[BMAX];
#pragma unroll
          for (int c = 0; c < BMAX; ++c) { p0[c] = pb[c]; p1[c] = s1[c]; p2[c] = s2[c]; }
          // D padded with the identity, inverted in place without pivoting
          double Dm[3][3], Di[3][3];
#pragma unroll
          for (int m = 0; m < 3; ++m) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              const double* pm = m == 0 ? p0 : (m == 1 ? p1 : p2);
              Dm[m][q] = (m < kk && q < kk && p + q < BMAX) ? pm[p + q < BMAX ? p + q : 0] : (m == q ? 1.0 : 0.0);
              Di[m][q] = m == q ? 1.0 : 0.0;
            }
          }
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const double piv = Dm[q][q];
            if (q < kk && !(fabs(piv) > 1e-12)) bad = 1.0;
            const double iv = 1.0 / piv;
#pragma unroll
            for (int c = 0; c < 3; ++c) { Dm[q][c] *= iv; Di[q][c] *= iv; }
#pragma unroll
            for (int m = 0; m < 3; ++m) {
              if (m != q) {
                const double f = Dm[m][q];
#pragma unroll
                for (int c = 0; c < 3; ++c) { Dm[m][c] -= f * Dm[q][c]; Di[m][c] -= f * Di[q][c]; }
              }
            }
          }
          const int m0 = B.lr - p;
          const bool pivrow = m0 >= 0 && m0 < kk;
          double bt[3], rk[3];
          double mr[3];
#pragma unroll
          for (int q = 0; q < 3; ++q) mr[q] = (q < kk && p + q < BMAX) ? F.row[p + q < BMAX ? p + q : 0] : 0.0;
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const double t = mr[0] * Di[0][q] + mr[1] * Di[1][q] + mr[2] * Di[2][q];
            const double dq = m0 == 0 ? Di[0][q] : (m0 == 1 ? Di[1][q] : Di[2][q]);
            bt[q] = pivrow ? dq - (q == m0 ? 1.0 : 0.0) : -t;
            rk[q] = pivrow ? dq : -t;
          }
#pragma unroll
          for (int c = 0; c < BMAX; ++c) {
            const bool step_col = c >= p && c < p + 3;
            if (!step_col) F.row[c] = fma(bt[2], p2[c], fma(bt[1], p1[c], fma(bt[0], p0[c], F.row[c])));
          }
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            if (p + q < BMAX) {
              const int c = p + q;
              F.row[c] = q < kk ? rk[q] : fma(bt[2], p2[c], fma(bt[1], p1[c], fma(bt[0], p0[c], F.row[c])));
            }
          }
        }
      }
    }
  } else if constexpr (N > 0) {
    // Gauss-Jordan in pairs of pivots (p, p + 1): one exchange of the two pivot
    // rows per pair instead of one row per pivot.  With the pivot rows P0, P1,
    // the 2x2 pivot block D = [[a, b], [c, d]] and every row written as
    // row + b0 P0 + b1 P1 outside columns p, p + 1 (one multiply-add each):
    //   rows outside the pair:  (t0, t1) = (row_p, row_p+1) D^-1, b = -t, the
    //                           pair's columns become -t;
    //   row p:   D^-1 row 0 applied to (P0, P1) (b0 = d/det - 1, b1 = -b/det);
    //   row p+1: D^-1 row 1 (b0 = -c/det, b1 = a/det - 1).
    // The same in-place inverse as two single pivots; a block whose size is odd
    // takes its last pivot alone.  The singularity test checks the two pivots
    // sequential elimination would see: a and det / a.  Fully unrolled, so the
    // pair's columns are fixed registers (maxb is the same for every thread:
    // the guard is uniform).
  #pragma unroll
    for (int p = 0; p < BMAX; p += 2) {
      if (p < maxb) {
        double* pb0 = sm.pbuf[(p >> 1) & 1][B.id][0];
        double* pb1 = sm.pbuf[(p >> 1) & 1][B.id][1];
        const bool in = p < B.size;
        const bool pair = p + 1 < B.size;
        if (in && (B.lr == p || (pair && B.lr == p + 1))) {  // the pivot rows as they stand
          double* dst = B.lr == p ? pb0 : pb1;
  #pragma unroll
          for (int c = 0; c < BMAX; ++c) dst[c] = F.row[c];
        }
        if (wl) wave_sync();
        else __syncthreads();
        if (in && B.lr < B.size) {
          // both pivot rows first: one wait for all their loads
          const double* src1 = pair ? pb1 : pb0;
          double p0[BMAX], p1[BMAX];
  #pragma unroll
          for (int c = 0; c < BMAX; ++c) { p0[c] = pb0[c]; p1[c] = src1[c]; }
          double b0, b1, rp, rq;
          if (pair && p + 1 < BMAX) {
            const double a = p0[p], b = p0[p + 1], cq = p1[p], d = p1[p + 1];
            const double det = a * d - b * cq;
            if (!(fabs(a) > 1e-12) || !(fabs(det) > 1e-12 * fabs(a))) bad = 1.0;
            const double id = 1.0 / det;
            const double m0 = F.row[p], m1 = F.row[p + 1];
            const double t0 = (m0 * d - m1 * cq) * id, t1 = (m1 * a - m0 * b) * id;
            if (B.lr == p) { b0 = d * id - 1.0; b1 = -b * id; rp = d * id; rq = -b * id; }
            else if (B.lr == p + 1) { b0 = -cq * id; b1 = a * id - 1.0; rp = -cq * id; rq = a * id; }
            else { b0 = -t0; b1 = -t1; rp = -t0; rq = -t1; }
          } else {  // the last pivot of an odd block
            const double dd = p0[p];
            if (!(fabs(dd) > 1e-12)) bad = 1.0;
            const double inv = 1.0 / dd;
            b0 = B.lr == p ? inv - 1.0 : -F.row[p] * inv;
            b1 = 0.0;
            rp = B.lr == p ? inv : b0;
            rq = 0.0;   // column p + 1 is outside this block (zero): updated below like the others
          }
  #pragma unroll
          for (int c = 0; c < BMAX; ++c)
            if (c != p && c != p + 1) F.row[c] = fma(b1, p1[c], fma(b0, p0[c], F.row[c]));
          F.row[p] = rp;
          if (p + 1 < BMAX) {
            if (pair) F.row[p + 1] = rq;
            else F.row[p + 1] = fma(b1, p1[p + 1], fma(b0, p0[p + 1], F.row[p + 1]));
          }
        }
      }
    }
  } else {
    // value-only kernel (three wavefronts per SIMD): one pivot at a time keeps a
    // single pivot row in registers
    // fully unrolled over the pivot index, so F.row[p] is a fixed register
    // (maxb is the same for every thread: the guard is uniform)
  #pragma unroll
    for (int p = 0; p < BMAX; ++p) {
      if (p < maxb) {
        double* pb = sm.pbuf[p & 1][B.id][0];
        const bool in = p < B.size;
        const bool piv = in && B.lr == p;
        if (piv) {  // the pivot row as it stands, 1/d in the pivot column
          const double d = F.row[p];
          if (!(fabs(d) > 1e-12)) bad = 1.0;
  #pragma unroll
          for (int c = 0; c < BMAX; ++c) pb[c] = (c == p) ? 1.0 / d : F.row[c];
        }
        if (wl) wave_sync();
        else __syncthreads();
        if (in && B.lr < B.size) {
          // every row of the block in one multiply-add per column: the pivot row
          // scales itself (beta = 1/d - 1 against its own entries), the others
          // eliminate (beta = -f/d); column p becomes 1/d, resp. -f/d
          // the whole pivot row first: one wait for all its loads, instead of the
          // load-wait-multiply-add chain the scheduler otherwise emits
          double pv[BMAX];
  #pragma unroll
          for (int c = 0; c < BMAX; ++c) pv[c] = pb[c];
          const double inv = pv[p];
          const double beta = piv ? inv - 1.0 : -F.row[p] * inv;
  #pragma unroll
          for (int c = 0; c < BMAX; ++c)
            if (c != p) F.row[c] = fma(beta, pv[c], F.row[c]);
          F.row[p] = piv ? inv : beta;
        }
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
  wg_sum(sm, rs, v);
  double a = 1.0 - v[0], b = -v[1], cc = -v[2], d = 1.0 - v[3];
  double det = a * d - b * cc;
  F.c00 = d / det; F.c01 = -b / det; F.c10 = -cc / det; F.c11 = a / det;
  return v[4] == 0.0 && isfinite(F.c00) && isfinite(F.c11) && det != 0.0;
}

// x = (I - g A(t_n))^-1 b, value and tangents with the same factors
template <int NT, int BMAX, int N>
__device__ __forceinline__ D<N> solve(Smem<NT, BMAX, N>& sm, const RowState& R, int& rs, const Inv<BMAX>& F,
                                      const D<N>& b) {
  constexpr int NC = 1 + N;
  const int r = threadIdx.x;
  put_row<N>(sm.rl, r, r < R.n ? b : D<N>(0.0));
  __syncthreads();
  double z[NC];
  double v[2 * NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < BMAX; ++c) s += F.row[c] * sm.rl[(R.B.start + c) * NC + q];
    z[q] = s;
    v[2 * q] = F.Vh * s;
    v[2 * q + 1] = F.Ve * s;
  }
  wg_sum(sm, rs, v);
  double out[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    double a = F.c00 * v[2 * q] + F.c01 * v[2 * q + 1];
    double bb = F.c10 * v[2 * q] + F.c11 * v[2 * q + 1];
    out[q] = r < R.n ? z[q] + F.Yh * a + F.Ye * bb : 0.0;
  }
  D<N> x;
  x.v = out[0];
  if constexpr (N > 0) { ABCMB_DLOOP(N) x.d[i_] = out[1 + i_]; }
  return x;
}

__device__ __forceinline__ double rms_term(double e, double y0, double y1, double rtol, double atol) {
  double m = fmax(fabs(y0), fabs(y1));
  double v = e / (atol + rtol * m);
  return v * v;
}

// initial conditions (abcmb/perturbations.py:161-199 and species y_ini)
template <int N>
__device__ __noinline__ D<N> ic_row(const Cosmo<N>& P, const Layout& L, double k, D<N> ti, int r) {
  const D<N> om = P(F_OM), Rn = P(F_R_NU), Rb = P(F_R_B);
  D<N> kt = k * ti;
  D<N> dg = -kt * kt / 3. * (1. - om * ti / 5.);
  D<N> thb = -k * k * k * k * ti * ti * ti / 36. * (1. - 3. * (1. + 5. * Rb - Rn) / 20. / (1. - Rn) * om * ti);
  D<N> tn = -k * kt * kt * kt / 36. / (4. * Rn + 15.) *
            (4. * Rn + 11. + 12. - 3. * (8. * Rn * Rn + 50. * Rn + 275.) / 20. / (2. * Rn + 15.) * ti * om);
  D<N> sn = kt * kt / (45. + 12. * Rn) * 2. * (1. + (4. * Rn - 5.) / 4. / (2. * Rn + 15.) * ti * om);
  if (r == 0)
    return 1. - k * k * ti * ti / 12. / (15. + 4. * Rn) *
                    (5. + 4. * Rn - (16. * Rn * Rn + 280. * Rn + 325.) / 10. / (2. * Rn + 15.) * ti * om);
  if (r == 1 || r == 2) return -kt * kt / 4. * (1. - om * ti / 5.);
  if (r == 3) return thb;
  if (r == L.ig) return dg;
  if (r == L.ig + 1) return thb;
  if (r == L.iu) return dg;
  if (r == L.iu + 1) return tn;
  if (r == L.iu + 2) return sn;
  if (r >= L.inc && r < L.ifl) {
    int nl = L.ln + 1;
    int i = (r - L.inc) / nl, l = (r - L.inc) % nl;
    double q = L.nq == 5 ? Q5[i] : Q3[i];
    double f = q / (1. + exp(-q));
    if (l == 0) return dg / 4. * f;
    if (l == 1) return tn / 3. * f;
    if (l == 2) return sn / 2. * f;
  }
  if constexpr (fluid::kRows > 0) {
    if (r >= L.ifl && r < L.n) return fluid::ic<N>(r - L.ifl, P, k, ti);
  }
  return D<N>(0.0);
}

// output-table fields at one saved point from the state in LDS (ys[c][row])
// (abcmb/perturbations.py:318-404; output_perturbations in abcmb/species.py)
// b: the background at the saved point (precomputed once per cosmology)
// The output fields read 12 + 3 n_q state rows: eta, delta_c, delta_b,
// theta_b, photon F_0..F_2, polarisation G_0 and G_2, massless-nu F_0..F_2 and
// F_0..F_2 of every massive-nu momentum bin.  Slot s of that list holds row
// out_row(L, s); out_slot is the inverse (-1 for the other rows).
__device__ __forceinline__ int out_rows(const Layout& L) { return 12 + 3 * L.nq; }

__device__ __forceinline__ int out_slot(const Layout& L, int r) {
  if (r < 4) return r;
  if (r >= L.ig && r < L.ig + 3) return 4 + r - L.ig;
  if (r == L.ip) return 7;
  if (r == L.ip + 2) return 8;
  if (r >= L.iu && r < L.iu + 3) return 9 + r - L.iu;
  if (r >= L.inc && r < L.ifl) {
    const int b = (r - L.inc) / (L.ln + 1), m = (r - L.inc) % (L.ln + 1);
    if (m < 3) return 12 + 3 * b + m;
  }
  return -1;
}

// the rows as columns of the ODE kernel's LDS state
template <int N, int LD>
struct RowsLds {
  const double (*ys)[LD];
  const Layout& L;
  __device__ __forceinline__ D<N> at(int i) const {
    D<N> x(ys[0][i]);
    if constexpr (N > 0) { ABCMB_DLOOP(N) x.d[i_] = ys[1 + i_][i]; }
    return x;
  }
  __device__ __forceinline__ D<N> eta() const { return at(0); }
  __device__ __forceinline__ D<N> dc() const { return at(1); }
  __device__ __forceinline__ D<N> db() const { return at(2); }
  __device__ __forceinline__ D<N> tb() const { return at(3); }
  __device__ __forceinline__ D<N> g(int m) const { return at(L.ig + m); }
  __device__ __forceinline__ D<N> pol(int m) const { return at(L.ip + m); }
  __device__ __forceinline__ D<N> ur(int m) const { return at(L.iu + m); }
  __device__ __forceinline__ D<N> nc(int b, int m) const { return at(L.inc + b * (L.ln + 1) + m); }
  __device__ __forceinline__ D<N> fl(int j) const { return at(L.ifl + j); }
};

// the rows in registers, by slot (bins beyond the layout's read as zero)
template <int N>
struct RowsReg {
  D<N> Y[27];
  D<N> F[fluid::kRows > 0 ? fluid::kRows : 1];  // user-fluid rows
  __device__ __forceinline__ D<N> eta() const { return Y[0]; }
  __device__ __forceinline__ D<N> dc() const { return Y[1]; }
  __device__ __forceinline__ D<N> db() const { return Y[2]; }
  __device__ __forceinline__ D<N> tb() const { return Y[3]; }
  __device__ __forceinline__ D<N> g(int m) const { return Y[4 + m]; }
  __device__ __forceinline__ D<N> pol(int m) const { return m == 0 ? Y[7] : Y[8]; }
  __device__ __forceinline__ D<N> ur(int m) const { return Y[9 + m]; }
  __device__ __forceinline__ D<N> nc(int b, int m) const { return Y[12 + 3 * b + m]; }
  __device__ __forceinline__ D<N> fl(int j) const { return F[j]; }
};

// every field of the perturbation table at one saved point
// (abcmb/perturbations.py:318-404 and the species output_perturbations),
// handed to emit(field, value)
template <int N, typename Rows, typename Emit>
__device__ __forceinline__ void all_fields(const Cosmo<N>& P, const Layout& L, double k, const Bg<N>& b,
                                           const Rows& y, Emit emit) {
  const double a = b.a;
  const D<N> aH = b.aH;
  D<N> nrd(0.0), nrt(0.0), nrs(0.0);
  if (L.nq > 0) {
    D<N> s0(0.0), s1(0.0), s2(0.0);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      if (i < L.nq) {
        double q = L.nq == 5 ? Q5[i] : Q3[i];
        double wq = L.nq == 5 ? W5[i] : W3[i];
        D<N> e = dsqrt(q * q + b.xn * b.xn);
        s0 += wq * (1.0 + exp(-q)) * e / (q * q) * y.nc(i, 0);
        s1 += wq * (1.0 + exp(-q)) / q * y.nc(i, 1);
        s2 += wq * (1.0 + exp(-q)) / e * y.nc(i, 2);
      }
    }
    D<N> pre = P(F_N_NCDM) * b.Tn * b.Tn * b.Tn * b.Tn / (PI * PI) / (HBAR * HBAR * HBAR) / (C_CM * C_CM * C_CM);
    nrd = s0 * 4.0 * pre; nrt = s1 * 4.0 * pre; nrs = s2 * 8.0 / 3.0 * pre;
  }
  emit(ABCMB_PT_THETA_B_PRIME, -y.tb() + b.cs2 / aH * (k * k * y.db()) + b.R / aH / b.tauc * (y.g(1) - y.tb()));
  emit(ABCMB_PT_ETA, y.eta());
  emit(ABCMB_PT_CDM_DELTA, y.dc());
  emit(ABCMB_PT_B_DELTA, y.db());
  emit(ABCMB_PT_B_THETA, y.tb());
  emit(ABCMB_PT_G_DELTA, y.g(0));
  emit(ABCMB_PT_G_THETA, y.g(1));
  emit(ABCMB_PT_G_SIGMA, y.g(2));
  emit(ABCMB_PT_G_G0, y.pol(0));
  emit(ABCMB_PT_G_G2, y.pol(2));
  emit(ABCMB_PT_UR_DELTA, y.ur(0));
  emit(ABCMB_PT_UR_THETA, y.ur(1));
  emit(ABCMB_PT_UR_SIGMA, y.ur(2));
  emit(ABCMB_PT_NCDM_DELTA, L.nq > 0 ? nrd / b.rn : D<N>(0.0));
  emit(ABCMB_PT_NCDM_THETA, L.nq > 0 ? nrt / (b.rn + b.pn) : D<N>(0.0));
  emit(ABCMB_PT_NCDM_SIGMA, L.nq > 0 ? nrs / (b.rn + b.pn) : D<N>(0.0));
  D<N> srd = b.rc * y.dc() + b.rb * y.db() + b.rg * y.g(0) + b.ru * y.ur(0);
  D<N> srt = b.rb * y.tb() + (b.rg + b.rg / 3.) * y.g(1) + (b.ru + b.ru / 3.) * y.ur(1);
  D<N> srs = (b.rg + b.rg / 3.) * y.g(2) + (b.ru + b.ru / 3.) * y.ur(2);
  D<N> srdm = b.rc * y.dc() + b.rb * y.db(), srm = b.rc + b.rb;
  if (L.nq > 0) { srd += nrd; srt += nrt; srs += nrs; srdm += nrd; srm += b.rn; }
  if constexpr (fluid::kActive) fluid::sums<N>(P, log(a), a, y, srd, srt, srs, srdm, srm);
  emit(ABCMB_PT_DELTA_M, srdm / srm);
  double KG = 4. * PI * GN * a * a / (C_MPC * C_MPC);
  D<N> hp = 2. / (aH * aH) * (k * k * y.eta() + KG * srd);
  D<N> ep = KG / aH * srt / (k * k);
  emit(ABCMB_PT_H_PRIME, hp);
  emit(ABCMB_PT_ETA_PRIME, ep);
  D<N> al = aH * (hp + 6. * ep) / 2. / (k * k);
  emit(ABCMB_PT_ALPHA, al);
  emit(ABCMB_PT_ALPHA_PRIME, y.eta() / aH - 2. * al - 3. * KG / aH * srs / (k * k));
}

// field f at one saved point from the LDS state (the in-kernel output path)
template <int N, int LD>
__device__ __noinline__ D<N> output_field(const Cosmo<N>& P, const Layout& L, double k, const Bg<N>& b,
                                          const double (*ys)[LD], int f) {
  D<N> out(0.0);
  all_fields(P, L, k, b, RowsLds<N, LD>{ys, L}, [&](int g, const D<N>& v) {
    if (g == f) out = v;
  });
  return out;
}

// The perturbation table in place: where the layout's output rows fit the
// table's field slots (12 + 3 n_q <= 21, i.e. up to three momentum bins) the
// ODE kernel stores the saved rows themselves, slot s in field slot s, with
// no workgroup synchronisation; after the launch every (k-mode, saved point)
// is turned into the table fields, one thread each (fields_at).  The rows of
// a saved point are read into registers first, so the fields can overwrite
// them in any order.
template <int N>
__device__ __forceinline__ void fields_at(const AbcmbCosmo& c, const AbcmbTangent* T, const AbcmbOdeSettings& st,
                                          const Bg<N>& bg, double k, int ik, int isave) {
  const Cosmo<N> P{c, T};
  const Layout L = make_layout(st);
  const int nr = out_rows(L);
  const int n_save = c.n_save;
  const size_t dstride = (size_t)ABCMB_N_PT_FIELDS * n_save * c.nk;
  auto at = [&](int s) { return ((size_t)s * n_save + isave) * c.nk + ik; };
  RowsReg<N> y;
#pragma unroll
  for (int s = 0; s < 27; ++s) {
    y.Y[s] = D<N>(0.0);
    if (s < ABCMB_N_PT_FIELDS && s < nr) {
      const size_t o = at(s);
      y.Y[s].v = c.pt[o];
      if constexpr (N > 0) { ABCMB_DLOOP(N) y.Y[s].d[i_] = T->pt[i_ * dstride + o]; }
    }
  }
  if constexpr (fluid::kRows > 0) {
    const size_t fstride = (size_t)fluid::kRows * n_save * c.nk;
#pragma unroll
    for (int j = 0; j < fluid::kRows; ++j) {
      const size_t o = ((size_t)j * n_save + isave) * c.nk + ik;
      y.F[j] = D<N>(c.pt_fluid[o]);
      if constexpr (N > 0) { ABCMB_DLOOP(N) y.F[j].d[i_] = T->pt_fluid[i_ * fstride + o]; }
    }
  }
  all_fields(P, L, k, bg, y, [&](int f, const D<N>& v) {
    const size_t o = at(f);
    c.pt[o] = v.v;
    if constexpr (N > 0) { ABCMB_DLOOP(N) T->pt[i_ * dstride + o] = v.d[i_]; }
  });
}

// how many of a launch's successful items fall on each cosmology (cnt zeroed
// first by a fill kernel, so that the sequence can sit in a captured graph)
__global__ void zero_counts_kernel(int32_t* __restrict__ cnt, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cnt[i] = 0;
}
__global__ void count_items_kernel(const AbcmbPertItem* __restrict__ items, int n_items,
                                   const int32_t* __restrict__ status, int32_t* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_items && status[i] == ABCMB_K_OK) atomicAdd(&cnt[items[i].cosmo], 1);
}

// cosmologies whose every k-mode is in the launch and integrated without
// failure (items are distinct (cosmology, k) pairs): one block per
// (cosmology, saved point), threads along k, so that the table's
// k-contiguous rows are read and written coalesced
template <int N>
__global__ void __launch_bounds__(64) fields_cosmo_kernel(const AbcmbCosmo* __restrict__ cosmos,
                                                          const AbcmbTangent* __restrict__ tangents,
                                                          AbcmbOdeSettings st, const Bg<N>* __restrict__ bgsave,
                                                          int n_save_max, const int32_t* __restrict__ cnt) {
  const int ci = blockIdx.x / n_save_max, isave = blockIdx.x % n_save_max;
  const AbcmbCosmo& c = cosmos[ci];
  if (isave >= c.n_save || cnt[ci] != c.nk) return;
  const AbcmbTangent* T = nullptr;
  if constexpr (N > 0) T = tangents + ci;
  const Bg<N>& bg = bgsave[(size_t)ci * n_save_max + isave];
  for (int ik = threadIdx.x; ik < c.nk; ik += 64) fields_at<N>(c, T, st, bg, c.k[ik], ik, isave);
}

// the other items (cosmologies covered in part or with a failed item), one
// thread per (item, saved point); a failed item's columns are zeroed, so that
// work enqueued behind the launch reads defined values until the caller has
// seen the status words
template <int N>
__global__ void __launch_bounds__(64) fields_kernel(const AbcmbCosmo* __restrict__ cosmos,
                                                    const AbcmbTangent* __restrict__ tangents,
                                                    const AbcmbPertItem* __restrict__ items, int n_items,
                                                    AbcmbOdeSettings st, const Bg<N>* __restrict__ bgsave,
                                                    int n_save_max, const int32_t* __restrict__ status,
                                                    const int32_t* __restrict__ cnt) {
  const long long idx = (long long)blockIdx.x * 64 + threadIdx.x;
  if (idx >= (long long)n_items * n_save_max) return;
  const int item = (int)(idx / n_save_max), isave = (int)(idx % n_save_max);
  const AbcmbPertItem I = items[item];
  const AbcmbCosmo& c = cosmos[I.cosmo];
  if (isave >= c.n_save || cnt[I.cosmo] == c.nk) return;
  const AbcmbTangent* T = nullptr;
  if constexpr (N > 0) T = tangents + I.cosmo;
  if (status[item] != ABCMB_K_OK) {
    const size_t dstride = (size_t)ABCMB_N_PT_FIELDS * c.n_save * c.nk;
    for (int f = 0; f < ABCMB_N_PT_FIELDS; ++f) {
      const size_t o = ((size_t)f * c.n_save + isave) * c.nk + I.ik;
      c.pt[o] = 0.0;
      if constexpr (N > 0) { ABCMB_DLOOP(N) T->pt[i_ * dstride + o] = 0.0; }
    }
    return;
  }
  fields_at<N>(c, T, st, bgsave[(size_t)I.cosmo * n_save_max + isave], I.k, I.ik, isave);
}

// the background (with its tangents) at every saved point of every cosmology,
// shared by all the cosmology's k-modes: bgsave[cosmo][i], i < n_save_max
template <int N>
__global__ void bg_save_kernel(const AbcmbCosmo* __restrict__ cosmos, const AbcmbTangent* __restrict__ tangents,
                               int n_save_max, Bg<N>* __restrict__ bgsave) {
  const int ci = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const AbcmbCosmo& c = cosmos[ci];
  if (i >= c.n_save || i >= n_save_max) return;
  const AbcmbTangent* T = nullptr;
  if constexpr (N > 0) T = tangents + ci;
  const Cosmo<N> P{c, T};
  Bg<N> b;
  bg_eval(P, c.lna_save[i], b);
  bgsave[(size_t)ci * n_save_max + i] = b;
}

// Stages 1..6 of one Kvaerno5 step from (y, f0) over hh, with the step's chord
// factors FI (ok: the factorisation succeeded).  On return sm.Fs holds the
// stage derivatives F_0..F_6 of the value and z the last stage (the new state);
// false when a stage's Newton iteration failed.  zs(s, z) sees every converged
// stage s >= 1 (the adjoint's recomputation stores them; a no-op elsewhere).
template <int NT, int BMAX, int N, typename ZHook>
__device__ __forceinline__ bool step_stages(Smem<NT, BMAX, N>& sm, const RowState& R, int& rs, const Inv<BMAX>& FI,
                                            const AbcmbOdeSettings& st, bool ok, double hh, double rtol,
                                            double atol, double inv_n, const D<N>& y, const D<N>& f0, D<N>& z,
                                            int& iters, ZHook zs) {
  const int r = threadIdx.x;
  const int n = R.n;
  const double g = hh * KV_G;
  const double inv_g = 1.0 / g;
  put_F(sm, 0, r, f0);
  z = y;
#pragma unroll 1
  for (int s = 1; s < 7; ++s) {
    if (!ok) break;
    D<N> acc(0.0);
    // all six stage derivatives read at once (one wait), the first s summed in order
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const D<N> fj = get_F(sm, j, r);
      if (j < s) acc += KV_A[s][j] * fj;
    }
    const D<N> base = y + hh * acc;
    z = base + g * get_F(sm, s - 1, r);
    const int slot = STAGE_SLOT[s];
    double ms[2 * (1 + N)];  // metric sums of z, kept up to date by the updates
    double zn[N > 0 ? N : 1];  // the predictor's scaled tangent norms (the tangents' yardstick)
    {
      constexpr int NM = 2 * (1 + N);
      double sv[NM + N];
      put_z(sm, r, z);
      metric_terms(sm, R, slot, z, sv);
      if constexpr (N > 0) {
        const double w0 = r < n ? 1.0 / (atol + rtol * fabs(z.v)) : 0.0;
        ABCMB_DLOOP(N) sv[NM + i_] = (z.d[i_] * w0) * (z.d[i_] * w0);
      }
      wg_sum(sm, rs, sv);
#pragma unroll
      for (int q = 0; q < NM; ++q) ms[q] = sv[q];
      if constexpr (N > 0) { ABCMB_DLOOP(N) zn[i_] = sv[NM + i_]; }
    }
    bool conv = false;      // value converged
    bool tconv = N == 0;    // tangents converged
    double prev = 0.0;
    double rate = 1.0;      // the value's latest contraction-rate estimate
    for (int it = 0;; ++it) {
      if (!conv && it >= st.max_newton) break;
      ++iters;
      D<N> fz = rhs_gather(sm, R, slot, ms);
      D<N> dz = solve(sm, R, rs, FI, -(z - base - g * fz));
      if (conv) dz.v = 0.0;  // the value is final: iterate the tangents alone
      z += dz;
      put_z(sm, r, z);  // for the next gather, ordered by the sum below
      // the Newton norms, then the metric-sum increments of dz
      constexpr int NQ = 1 + N;
      double v[NQ + 2 * (1 + N)];
      // one reciprocal per row; the decisions below compare products, so the
      // iteration spends two divisions (this one and the rate)
      const double w = r < n ? 1.0 / (atol + rtol * fabs(z.v)) : 0.0;
      v[0] = conv ? 0.0 : (dz.v * w) * (dz.v * w);
      if constexpr (N > 0) { ABCMB_DLOOP(N) v[1 + i_] = (dz.d[i_] * w) * (dz.d[i_] * w); }
      metric_terms(sm, R, slot, dz, v + NQ);
      wg_sum(sm, rs, v);
#pragma unroll
      for (int q = 0; q < 2 * (1 + N); ++q) ms[q] += v[NQ + q];
      if (!conv) {
        double nrm = sqrt(v[0] * inv_n);
        if (!isfinite(nrm)) break;
        if (nrm < NEWTON_SMALL) {
          conv = true;
        } else if (it > 0) {
          double th = nrm / prev;
          if (th > 2.0) break;
          rate = th;
          // nrm th / (1 - th) in (0, kappa), with nrm, th > 0
          if (th < 1.0 && nrm * th < st.newton_kappa * (1.0 - th)) conv = true;
        }
        prev = nrm;
      }
      if constexpr (N > 0) {
        // remaining error ~ update * amp, amp = rate / (1 - rate) when the
        // iteration contracts (else 1): converged when |dz| amp <= TANGENT_RTOL |z|
        const double an = rate < 1.0 ? rate : 1.0, ad = rate < 1.0 ? 1.0 - rate : 1.0;
        const double lhs = an * an, rhs = (TANGENT_RTOL * ad) * (TANGENT_RTOL * ad);
        tconv = true;
        ABCMB_DLOOP(N) tconv = tconv && v[1 + i_] * lhs <= rhs * zn[i_];
      }
      if (conv && (tconv || it + 1 >= st.max_newton + TANGENT_EXTRA)) break;
    }
    if (!conv) { ok = false; break; }
    put_F(sm, s, r, (z - base) * inv_g);
    zs(s, z);
  }
  return ok;
}

// what ode_item does besides integrating: NoRecord for the forward and
// tangent-linear launches (the saved points go to the table); the adjoint
// launch records the accepted steps and checkpoints instead (adjoint.hip)
struct NoRecord {
  static constexpr bool kSave = true;
  template <int N>
  __device__ __forceinline__ void accept(double, double, const D<N>&, const D<N>&) {}
  template <int N>
  __device__ __forceinline__ void finish(int, const D<N>&, const D<N>&) {}
};

struct NoStage {
  template <int N>
  __device__ __forceinline__ void operator()(int, const D<N>&) const {}
};

// One work item (one k-mode of one cosmology) on the workgroup.
template <int NT, int BMAX, int N, typename Rec = NoRecord>
__device__ __forceinline__ void ode_item(const int item, Smem<NT, BMAX, N>& sm, const AbcmbCosmo* __restrict__ cosmos,
                                         const AbcmbTangent* __restrict__ tangents,
                                         const AbcmbPertItem* __restrict__ items, const AbcmbOdeSettings& st,
                                         const Bg<N>* __restrict__ bgsave, int n_save_max,
                                         int32_t* __restrict__ status_out, int32_t* __restrict__ steps_out,
                                         int32_t* __restrict__ iters_out, double* __restrict__ state_out,
                                         Rec rec = Rec()) {
  const int r = threadIdx.x;
  const AbcmbPertItem I = items[item];
  const AbcmbCosmo& c = cosmos[I.cosmo];
  const AbcmbTangent* T = nullptr;
  if constexpr (N > 0) T = tangents + I.cosmo;
  const Cosmo<N> P{c, T};
  const double k = I.k;
  const Layout L = make_layout(st);
  const int n = L.n;
  RowState R;
  if constexpr (N > 0) row_desc(L, k, r, R.w);
  else row_desc_call(L, k, r, R.w);
  R.B = row_block(L, r);
  R.n = n;
#pragma unroll
  for (int j = 0; j < 5; ++j) R.lc[j] = R.w.col[j] - R.B.start;
  const int maxb = max(max(L.iu - 2, L.nfl), max(L.lu + 1, L.nq > 0 ? L.ln + 1 : 1));
  // every hierarchy block inside one wavefront (rows lo..hi)
  auto one_wave = [](int lo, int hi) { return (lo >> 6) == (hi >> 6); };
  bool wl = one_wave(2, L.iu - 1) && one_wave(L.iu, L.inc - 1);
  for (int i = 0; i < L.nq; ++i) wl = wl && one_wave(L.inc + i * (L.ln + 1), L.inc + (i + 1) * (L.ln + 1) - 1);
  if (L.nfl > 0) wl = wl && one_wave(L.ifl, L.n - 1);
  for (int q = 0; q < 1 + N; ++q) {
    sm.zl[q][r] = 0.0; sm.rl[r * (1 + N) + q] = 0.0;
    if (r < BMAX) { sm.zl[q][NT + r] = 0.0; sm.rl[(NT + r) * (1 + N) + q] = 0.0; }
  }
  sm.uh[r] = r < n ? R.w.uh : 0.0;
  sm.ue[r] = r < n ? R.w.ue : 0.0;
  if (r < BMAX) { sm.uh[NT + r] = 0.0; sm.ue[NT + r] = 0.0; }

  const double rtol = I.rtol, atol = I.atol;
  const double inv_n = 1.0 / n;
  const int n_save = c.n_save;
  const double* __restrict__ lna_save = c.lna_save;
  if (n_save > n_save_max) {  // the background table has no room for this cosmology's saves
    if (r == 0) { status_out[item] = ABCMB_K_BAD_INPUT; steps_out[item] = 0; if (iters_out) iters_out[item] = 0; }
    return;
  }
  const Bg<N>* __restrict__ bgs = bgsave + (size_t)I.cosmo * n_save_max;

  double t = I.lna_start;
  const D<N> ti = fast_interp<N>(t, ABCMB_TAU_TAB_LO, 0.0, c.tau_tab, P.dtau(), ABCMB_TAU_TAB_N);
  D<N> y = r < n ? ic_row(P, L, k, ti, r) : D<N>(0.0);
  int rs = 0;  // reduction buffer half
  set_times(sm, P, L, k, t, 0.0, 1, st.jac_offset);
  D<N> f0 = rhs_eval(sm, R, rs, 0, y);
  double h = st.dt0;
  const double b1 = st.pcoeff + st.icoeff + st.dcoeff, b2 = -(st.pcoeff + 2.0 * st.dcoeff);
  double prev_inv = 1.0;
  int isave = 0, steps = 0, iters = 0, status = ABCMB_K_OK;
  Inv<BMAX> FI;

  // with the output rows fitting the table's field slots, a saved point is the
  // rows themselves (fields_kernel turns them into fields after the launch)
  const bool rows_out = out_rows(L) <= ABCMB_N_PT_FIELDS;
  const int oslot = rows_out && r < n ? out_slot(L, r) : -1;
  auto save_point = [&](const D<N>& ys) {
    if constexpr (fluid::kRows > 0) {  // the user-fluid rows, as they are
      if (r >= L.ifl && r < n) {
        const size_t o = ((size_t)(r - L.ifl) * n_save + isave) * c.nk + I.ik;
        c.pt_fluid[o] = ys.v;
        if constexpr (N > 0) {
          const size_t fstride = (size_t)fluid::kRows * n_save * c.nk;
          ABCMB_DLOOP(N) T->pt_fluid[i_ * fstride + o] = ys.d[i_];
        }
      }
    }
    if (rows_out) {
      if (state_out && r < n) state_out[((size_t)item * n_save + isave) * n + r] = ys.v;
      if (oslot >= 0) {
        const size_t o = ((size_t)oslot * n_save + isave) * c.nk + I.ik;
        c.pt[o] = ys.v;
        if constexpr (N > 0) {
          const size_t dstride = (size_t)ABCMB_N_PT_FIELDS * n_save * c.nk;
          ABCMB_DLOOP(N) T->pt[i_ * dstride + o] = ys.d[i_];
        }
      }
      return;
    }
    __syncthreads();
    put_z(sm, r, ys);
    __syncthreads();
    if (state_out && r < n) state_out[((size_t)item * n_save + isave) * n + r] = ys.v;
    if (r < ABCMB_N_PT_FIELDS) {
      D<N> v = output_field<N, Smem<NT, BMAX, N>::LDZ>(P, L, k, bgs[isave], sm.zl, r);
      const size_t o = ((size_t)r * n_save + isave) * c.nk + I.ik;
      c.pt[o] = v.v;
      if constexpr (N > 0) {
        const size_t dstride = (size_t)ABCMB_N_PT_FIELDS * n_save * c.nk;
        ABCMB_DLOOP(N) T->pt[i_ * dstride + o] = v.d[i_];
      }
    }
  };

  if constexpr (Rec::kSave) {
    while (isave < n_save && lna_save[isave] <= t) { save_point(y); ++isave; }
  }

  while (t < 0.0) {
    if (steps >= st.max_steps) { status = ABCMB_K_MAX_STEPS; break; }
    double tn = t + h;
    if (tn > 0.0) tn = 0.0;
    const double hh = tn - t;
    if (!(hh > 0.0)) { status = ABCMB_K_STEP_UNDERFLOW; break; }
    ++steps;
    const double g = hh * KV_G;
    set_times(sm, P, L, k, t, hh, NTIMES, st.jac_offset);
    bool ok = factor(sm, R, rs, maxb, wl, g, FI);
    D<N> z;
    ok = step_stages(sm, R, rs, FI, st, ok, hh, rtol, atol, inv_n, y, f0, z, iters, NoStage());
    double enorm = INFINITY;
    const D<N> ynew = z;
    if (ok) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 7; ++j) acc += KV_E[j] * sm.Fs[j][0][r];
      double v[1] = {r < n ? rms_term(hh * acc, y.v, ynew.v, rtol, atol) : 0.0};
      wg_sum(sm, rs, v);
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
      rec.accept(t, hh, y, f0);
      D<N> fnew = rhs_eval(sm, R, rs, 5, ynew);
      if constexpr (Rec::kSave) {
        while (isave < n_save && lna_save[isave] <= tn) {
          double th = (lna_save[isave] - t) / hh;
          double h00 = 2 * th * th * th - 3 * th * th + 1, h10 = th * th * th - 2 * th * th + th;
          double h01 = -2 * th * th * th + 3 * th * th, h11 = th * th * th - th * th;
          D<N> ys = h00 * y + h10 * hh * f0 + h01 * ynew + h11 * hh * fnew;
          save_point(ys);
          ++isave;
        }
      }
      t = tn;
      y = ynew;
      f0 = fnew;
      prev_inv = inv;
    }
    h = hh * fac;
  }
  rec.finish(status, y, f0);
  if (r == 0) {
    status_out[item] = status;
    steps_out[item] = steps;
    if (iters_out) iters_out[item] = iters;
  }
}

// host checks of a launch's settings shared by the entry points; sets the block bound
int check_settings(const AbcmbOdeSettings& st, int& maxb) {
  ABCMB_REQUIRE(st.l_g >= 5 && st.l_pol >= 4 && st.l_ur >= 5, "hierarchy cut-offs too small");
  ABCMB_REQUIRE(st.n_q == 0 || ((st.n_q == 3 || st.n_q == 5) && st.l_ncdm >= 4), "bad massive-nu layout");
  ABCMB_REQUIRE(st.n_fluid == fluid::kRows,
                "settings.n_fluid=%d but this library's fluid code has %d rows (signature '%s')", st.n_fluid,
                fluid::kRows, fluid::kSignature);
  int n_expect = 4 + (st.l_g + 1) + (st.l_pol + 1) + (st.l_ur + 1) + st.n_q * (st.l_ncdm + 1) + st.n_fluid;
  ABCMB_REQUIRE(st.n == n_expect, "settings.n=%d does not match the layout (%d)", st.n, n_expect);
  ABCMB_REQUIRE(st.max_newton > 0 && st.max_steps > 0, "bad iteration limits");
  maxb = 2 + (st.l_g + 1) + (st.l_pol + 1);
  if (st.l_ur + 1 > maxb) maxb = st.l_ur + 1;
  if (st.n_q > 0 && st.l_ncdm + 1 > maxb) maxb = st.l_ncdm + 1;
  if (st.n_fluid > maxb) maxb = st.n_fluid;
  ABCMB_REQUIRE(maxb <= 64, "a hierarchy block of %d rows exceeds 64", maxb);
  ABCMB_REQUIRE(st.n <= 192, "state size %d exceeds 192", st.n);
  return ABCMB_OK;
}

}  // namespace
}  // namespace abcmb
