// Host-side background producer for the perturbation + LOS path.
//
// Produces, per cosmology and in FP64, everything the GPU hot path reads
// about the homogeneous universe: the conformal-time table, the free-electron
// fraction and matter temperature histories, reionization, the visibility
// function, and the discrete grid choices the reference makes
// (lna_rec, lna_transfer_start, per-k start times).
//
// Reference behaviour followed:
//   closed-form rho/P/H/aH          abcmb/species.py (rho/P of each fluid),
//                                   abcmb/background.py:101-208
//   tau(ln a) table                 abcmb/background.py:252-347
//   x_e, T_m, tau_c lookups         abcmb/background.py:568-680 (incl. the
//                                   fast_interp span quirk, :599-601, :653-655)
//   reionization                    abcmb/background.py:939-1017
//   kappa, exp(-kappa), visibility  abcmb/background.py:682-768
//   lna_rec, lna_transfer_start     abcmb/background.py:555-564
//   start time per k                abcmb/perturbations.py:119-159, :279
//   recombination (x_e, T_m)        abcmb/hyrex/*.py, restated in hyrex.cpp,
//                                   on the tables' reference layout
#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>

#include "../../include/abcmb.h"
#include "hyrex.h"

namespace {

constexpr double C_CM = 29979245800.0;
constexpr double C_MPC = 9.71561e-15;
constexpr double H0_OVER_H = 3.24078e-18;
constexpr double HBAR = 6.582119569509e-16;
constexpr double GN = 1.18980205e-40;
constexpr double KB = 8.617343e-5;
constexpr double MP = 938271999.0;
constexpr double ME = 510998.9461;
constexpr double MH = MP + ME + 13.5982860719383;
constexpr double MHE = 3.72839e9;
constexpr double MU_E = MP * ME / MH;
constexpr double E21 = 10.198714553953742;
constexpr double RYD = 13.598286071938324;
constexpr double SIGMA_T = 6.652458734e-25;
constexpr double R2S1S = 8.2206;
const double PI = 3.14159265358979323846;
const double STEF = PI * PI / (60.0 * HBAR * HBAR * HBAR * C_CM * C_CM);
const double LYA_FREQ = RYD * 0.75 / (2.0 * PI * HBAR);
const double RHO_UNIT = 3.0 * H0_OVER_H * H0_OVER_H / 8.0 / PI / GN;
const double HC3 = (C_CM * HBAR) * (C_CM * HBAR) * (C_CM * HBAR);

const double Q5[5] = {0.583165, 2.0, 4.0, 7.26582, 13.0};
const double W5[5] = {0.0081201, 0.689407, 2.8063, 2.05156, 0.12681};
const double EQ5[5] = {std::exp(-Q5[0]), std::exp(-Q5[1]), std::exp(-Q5[2]), std::exp(-Q5[3]), std::exp(-Q5[4])};

#ifndef ABCMB_FLUID_HOST
#define ABCMB_FLUID_HOST "fluid_none_host.h"
#endif
#include ABCMB_FLUID_HOST

struct Rho {
  double de, cdm, b, g, ur, ncdm, p_ncdm;
  double fl, p_fl;  // user fluids (abcmb_amd/fluidgen.py), summed
};

inline Rho rho_all(const AbcmbCosmo& c, double lna) {
  Rho r;
  double a = std::exp(lna);
  double a3 = a * a * a, a4 = a3 * a;
  r.de = (c.species_off & ABCMB_NO_LAMBDA) ? 0.0 : c.omega_Lambda * RHO_UNIT;
  r.cdm = (c.species_off & ABCMB_NO_CDM) ? 0.0 : c.omega_cdm * RHO_UNIT / a3;
  r.fl = r.p_fl = 0.0;
  if (fluid_host::kActive) fluid_host::bg(c, lna, a, r.fl, r.p_fl);
  r.b = c.omega_b * RHO_UNIT / a3;
  double t04 = c.TCMB0 * c.TCMB0 * c.TCMB0 * c.TCMB0;
  r.g = PI * PI / 15.0 * t04 / a4 / HC3;
  double tu4 = c.T_ur * c.T_ur * c.T_ur * c.T_ur;
  r.ur = c.N_ur * 2.0 * 7.0 / 8.0 * PI * PI / 30.0 * tu4 * t04 / a4 / HC3;
  r.ncdm = 0.0;
  r.p_ncdm = 0.0;
  if (c.has_ncdm) {
    double T = c.T_ncdm * c.TCMB0 / a;
    double x = c.m_ncdm / T;
    double ir = 0.0, ip = 0.0;
    for (int i = 0; i < 5; ++i) {
      double q = Q5[i];
      double eps = std::sqrt(q * q + x * x);
      ir += W5[i] * (1.0 + EQ5[i]) / (q * q) * eps;
      ip += W5[i] * (1.0 + EQ5[i]) / eps;
    }
    double T4 = T * T * T * T;
    r.ncdm = c.N_ncdm * 4.0 * T4 / (PI * PI) * ir / (HBAR * HBAR * HBAR) / (C_CM * C_CM * C_CM);
    r.p_ncdm = c.N_ncdm * 4.0 / 3.0 * T4 / (PI * PI) * ip / (HBAR * HBAR * HBAR) / (C_CM * C_CM * C_CM);
  }
  return r;
}

inline double rho_tot(const Rho& r) { return r.de + r.cdm + r.b + r.g + r.ur + r.ncdm + r.fl; }
inline double p_tot(const Rho& r) { return -r.de + r.g / 3.0 + r.ur / 3.0 + r.p_ncdm + r.p_fl; }

inline double hubble(const AbcmbCosmo& c, double lna) {
  return std::sqrt(8.0 * PI * GN * rho_tot(rho_all(c, lna)) / 3.0);
}
inline double aH(const AbcmbCosmo& c, double lna) { return std::exp(lna) * hubble(c, lna) / C_MPC; }
inline double aH_prime(const AbcmbCosmo& c, double lna) {
  Rho r = rho_all(c, lna);
  double a = std::exp(lna);
  double ah = a * std::sqrt(8.0 * PI * GN * rho_tot(r) / 3.0) / C_MPC;
  return -4.0 * PI * GN * a * a / 3.0 / ah * (rho_tot(r) + 3.0 * p_tot(r)) / (C_MPC * C_MPC);
}
inline double nH(const AbcmbCosmo& c, double lna) {
  double a = std::exp(lna);
  return (1.0 - c.YHe) * 3.0 * c.omega_b * H0_OVER_H * H0_OVER_H / 8.0 / PI / GN / MH / (a * a * a);
}

// fast_interp of abcmb/ABCMBTools.py:268-308 (index clip to [eps, n-1-eps]).
inline double fast_interp(double x, double lo, double hi, const double* f, int n) {
  double i = (x - lo) / (hi - lo) * (n - 1);
  const double eps = 1e-6;
  if (i < eps) i = eps;
  if (i > n - 1.0 - eps) i = n - 1.0 - eps;
  int il = (int)std::floor(i);
  int iu = std::min(il + 1, n - 1);
  double wu = i - il;
  return (1.0 - wu) * f[il] + wu * f[iu];
}
// slope d/dx of the same interpolant (0 where the index is clipped)
inline double fast_interp_slope(double x, double lo, double hi, const double* f, int n) {
  double i = (x - lo) / (hi - lo) * (n - 1);
  const double eps = 1e-6;
  if (i < eps || i > n - 1.0 - eps) return 0.0;
  int il = (int)std::floor(i);
  int iu = std::min(il + 1, n - 1);
  return (f[iu] - f[il]) * (n - 1) / (hi - lo);
}

inline double tau_of(const AbcmbCosmo& c, double lna) {
  return fast_interp(lna, ABCMB_TAU_TAB_LO, 0.0, c.tau_tab, ABCMB_TAU_TAB_N);
}
inline double xe_of(const AbcmbCosmo& c, double lna) {
  if (lna < c.xe_x0) return c.xe_tab[0];
  if (lna >= c.xe_lna_last) return c.xe_last;
  return fast_interp(lna, c.xe_x0, c.xe_x0 + c.xe_n * c.xe_dx, c.xe_tab, c.xe_n);
}
inline double dxe_dlna(const AbcmbCosmo& c, double lna) {
  if (lna < c.xe_x0 || lna >= c.xe_lna_last) return 0.0;
  return fast_interp_slope(lna, c.xe_x0, c.xe_x0 + c.xe_n * c.xe_dx, c.xe_tab, c.xe_n);
}
inline double tau_c_of(const AbcmbCosmo& c, double lna) {
  double a = std::exp(lna);
  double ne = nH(c, lna) * xe_of(c, lna);
  return 1.0 / a / ne / SIGMA_T / C_CM * C_MPC;
}
inline double gamma_compton(double xe, double T, double YHe) {
  double fhe = YHe * MH / MHE / (1.0 - YHe);
  return xe / (1.0 + fhe + xe) * 8.0 * SIGMA_T * (4.0 * STEF) * T * T * T * T / (3.0 * ME);
}
inline double tm_of(const AbcmbCosmo& c, double lna) {
  if (lna < c.tm_x0) {
    double T = c.TCMB0 / std::exp(lna);
    double xe = xe_of(c, lna);
    return T * (1.0 - hubble(c, lna) / gamma_compton(xe, T, c.YHe));
  }
  if (lna >= c.tm_lna_last) return c.tm_last;
  return fast_interp(lna, c.tm_x0, c.tm_x0 + c.tm_n * c.tm_dx, c.tm_tab, c.tm_n);
}
inline double cs2_of(const AbcmbCosmo& c, double lna) {
  Rho r = rho_all(c, lna);
  double tm = tm_of(c, lna);
  double tg = c.TCMB0 / std::exp(lna);
  double xe = xe_of(c, lna);
  double mu = MH / ((1.0 + xe) * (1.0 - c.YHe) + MH / MHE * c.YHe);
  double R = 4.0 * r.g / 3.0 / r.b;
  return tm / mu * (5.0 / 3.0 - 2.0 / 3.0 * mu * R / ME / aH(c, lna) / tau_c_of(c, lna) * (tg / tm - 1.0));
}

// 8-point Gauss-Legendre on [-1, 1]
const double GLX[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290, -0.1834346424956498,
                       0.1834346424956498, 0.5255324099163290, 0.7966664774136267, 0.9602898564975363};
const double GLW[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873, 0.3626837833783620,
                       0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763};

const double GL4X[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
const double GL4W[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};

template <class F>
double gl4(F f, double a, double b) {
  double m = 0.5 * (a + b), hw = 0.5 * (b - a), s = 0.0;
  for (int i = 0; i < 4; ++i) s += GL4W[i] * f(m + hw * GL4X[i]);
  return s * hw;
}

template <class F>
double gl8(F f, double a, double b) {
  double m = 0.5 * (a + b), hw = 0.5 * (b - a), s = 0.0;
  for (int i = 0; i < 8; ++i) s += GLW[i] * f(m + hw * GLX[i]);
  return s * hw;
}

// ---- conformal time table (abcmb/background.py:252-319) ----
void tau_table(AbcmbCosmo& c, double* tab) {
  const double lna_cut = -16.1;
  auto approx = [&](double l) { return std::exp(l) / (H0_OVER_H / C_MPC) / std::sqrt(c.omega_r); };
  auto inv_aH = [&](double l) { return 1.0 / aH(c, l); };
  const int n = ABCMB_TAU_TAB_N;
  double prev_l = lna_cut, prev_tau = approx(lna_cut);
  for (int i = 0; i < n; ++i) {
    double l = ABCMB_TAU_TAB_LO + (0.0 - ABCMB_TAU_TAB_LO) * i / (n - 1);
    if (i == n - 1) l = 0.0;
    if (l <= lna_cut) {
      tab[i] = approx(l);
      continue;
    }
    // integrate from prev_l to l in sub-intervals of at most 2e-3
    int nsub = std::max(1, (int)std::ceil((l - prev_l) / 4e-3));
    double h = (l - prev_l) / nsub, t = prev_tau;
    for (int s = 0; s < nsub; ++s) t += gl4(inv_aH, prev_l + s * h, prev_l + (s + 1) * h);
    tab[i] = t;
    prev_l = l;
    prev_tau = t;
  }
}

// ---- recombination: HyRex (hyrex.cpp) on its grid ----
int rec_points() { return hyrex::XE_ALLOC; }

// ---- reionization (abcmb/background.py:939-1017) ----
double xe_reion(double lna, double z_re, double dz, double zHe, double dzHe, double ex, double YHe) {
  double fHe = YHe / 4.0 / (1.0 - YHe);
  double z = 1.0 / std::exp(lna) - 1.0;
  double y = std::pow(1.0 + z, ex);
  double yre = std::pow(1.0 + z_re, ex);
  double dy = ex * std::pow(1.0 + z_re, ex - 1.0) * dz;
  double xH = (1.0 + fHe) / 2.0 * (1.0 + std::tanh((yre - y) / dy));
  double xHe = fHe / 2.0 * (1.0 + std::tanh((zHe - z) / dzHe));
  return xH + xHe;
}
// tau_reion(z_re) on a fixed ln a grid; the pieces of the integrand that do
// not depend on z_re (n_H, a H, (1 + z)^ex, the helium term) are evaluated
// once per cosmology, and each root-finder call only redoes the hydrogen
// step -- the same operations, in the same order, as a direct evaluation
struct ReionGrid {
  static constexpr int n = 2000;
  std::vector<double> nh, ea, ah, y, xhe;
  double fHe;
  void build(const AbcmbCosmo& c, double zHe, double dzHe, double ex) {
    nh.resize(n); ea.resize(n); ah.resize(n); y.resize(n); xhe.resize(n);
    fHe = c.YHe / 4.0 / (1.0 - c.YHe);
    for (int i = 0; i < n; ++i) {
      double l = -5.0 + 5.0 * i / (n - 1);
      double z = 1.0 / std::exp(l) - 1.0;
      nh[i] = nH(c, l);
      ea[i] = std::exp(l);
      ah[i] = aH(c, l);
      y[i] = std::pow(1.0 + z, ex);
      xhe[i] = fHe / 2.0 * (1.0 + std::tanh((zHe - z) / dzHe));
    }
  }
  double tau(double z_re, double dz, double ex) const {
    double yre = std::pow(1.0 + z_re, ex);
    double dy = ex * std::pow(1.0 + z_re, ex - 1.0) * dz;
    double s = 0.0, prev = 0.0;
    for (int i = 0; i < n; ++i) {
      double xH = (1.0 + fHe) / 2.0 * (1.0 + std::tanh((yre - y[i]) / dy));
      double ne = nh[i] * (xH + xhe[i]);
      double gam = ea[i] * ne * SIGMA_T * C_CM / C_MPC;
      double f = gam / ah[i];
      if (i > 0) s += 0.5 * (f + prev) * (5.0 / (n - 1));
      prev = f;
    }
    return s;
  }
};

// kappa(lna) = int_lna^0 dl / (tau_c aH): composite GL on a node grid
struct Kappa {
  std::vector<double> nodes, cum;  // cum[i] = kappa(nodes[i])
  const AbcmbCosmo* c;
  double integrand(double l) const { return 1.0 / tau_c_of(*c, l) / aH(*c, l); }
  void build(const AbcmbCosmo& cc) {
    c = &cc;
    const double lo = -10.0;
    const int n = 20001;  // 5e-4 spacing
    nodes.resize(n);
    cum.assign(n, 0.0);
    for (int i = 0; i < n; ++i) nodes[i] = lo + (0.0 - lo) * i / (n - 1);
    nodes[n - 1] = 0.0;
    for (int i = n - 2; i >= 0; --i)
      cum[i] = cum[i + 1] + gl4([&](double l) { return integrand(l); }, nodes[i], nodes[i + 1]);
  }
  double operator()(double l) const {
    const int n = (int)nodes.size();
    if (l <= nodes[0]) return cum[0];
    if (l >= 0.0) return 0.0;
    double fi = (l - nodes[0]) / (nodes[1] - nodes[0]);
    int i = std::min((int)fi, n - 2);
    return cum[i + 1] + gl4([&](double x) { return integrand(x); }, l, nodes[i + 1]);
  }
};

double expmkappa(const Kappa& K, double l) { return l < -10.0 ? 0.0 : std::exp(-K(l)); }

}  // namespace

extern "C" int abcmb_host_recomb_points(void) { return rec_points(); }
extern "C" int abcmb_host_tm_points(void) { return hyrex::TM_ALLOC; }

extern "C" const char* abcmb_host_fluid_signature(void) { return fluid_host::kSignature; }

namespace {
int background(AbcmbCosmo* c, double tau_reion, double z_reion_in, double Delta_z_reion, double z_reion_He,
               double Delta_z_reion_He, double exp_reion, double n_eff, int32_t* rec_switch, bool skip_reion) {
  if (!c || !c->tau_tab || !c->xe_tab || !c->tm_tab) return ABCMB_EINVAL;
  if (c->xe_n < hyrex::XE_ALLOC || c->tm_n < hyrex::TM_ALLOC) return ABCMB_EINVAL;
  double* tau = const_cast<double*>(c->tau_tab);
  double* xe = const_cast<double*>(c->xe_tab);
  double* tm = const_cast<double*>(c->tm_tab);
  tau_table(*c, tau);
  c->tau0 = fast_interp(0.0, ABCMB_TAU_TAB_LO, 0.0, c->tau_tab, ABCMB_TAU_TAB_N);
  // HyRex's inputs on its grid (abcmb/background.py:88-99)
  const double x0 = -std::log(1.0 + hyrex::Z0), dx = hyrex::DLNA;
  std::vector<double> Tg(hyrex::N_GRID), nhg(hyrex::N_GRID), Hg(hyrex::N_GRID);
  for (int j = 0; j < hyrex::N_GRID; ++j) {
    const double l = x0 + j * dx;
    Tg[j] = c->TCMB0 / std::exp(l);
    nhg[j] = nH(*c, l);
    Hg[j] = hubble(*c, l);
  }
  int32_t sw[hyrex::N_STAGES];
  for (int s = 0; s < hyrex::N_STAGES; ++s) sw[s] = rec_switch ? rec_switch[s] : -1;
  hyrex::History hist;
  const int rc = hyrex::recombination(*c, n_eff, Tg.data(), nhg.data(), Hg.data(), hyrex::N_GRID, xe, c->xe_n, tm,
                                      c->tm_n, sw, hist);
  if (rc != ABCMB_OK) return rc;
  if (rec_switch) for (int s = 0; s < hyrex::N_STAGES; ++s) rec_switch[s] = sw[s];
  const int N = hist.n_xe;
  c->xe_x0 = x0;
  c->xe_dx = c->tm_dx = dx;
  c->tm_x0 = hist.tm_x0;
  c->xe_lna_last = x0 + (N - 1) * dx;
  c->tm_lna_last = hist.tm_x0 + (hist.n_tm - 1) * dx;
  // reionization: solve tau_reion(z) = target (Newton/bisection), or take z
  double zr = z_reion_in;
  if (tau_reion > 0.0 && !skip_reion) {
    ReionGrid G;
    G.build(*c, z_reion_He, Delta_z_reion_He, exp_reion);
    auto f = [&](double z) { return G.tau(z, Delta_z_reion, exp_reion) - tau_reion; };
    double lo = 0.0, hi = 50.0, flo = f(lo), fhi = f(hi);
    for (int it = 0; it < 8; ++it) {  // bracket
      double mid = 0.5 * (lo + hi), fm = f(mid);
      if ((fm < 0) == (flo < 0)) { lo = mid; flo = fm; } else { hi = mid; fhi = fm; }
    }
    double a = lo, fa = flo, b = hi, fb = fhi;  // safeguarded secant (Illinois)
    for (int it = 0; it < 60 && std::fabs(b - a) > 1e-12; ++it) {
      double z = b - fb * (b - a) / (fb - fa);
      double fz = f(z);
      if ((fz < 0) == (fb < 0)) { fa *= 0.5; } else { a = b; fa = fb; }
      b = z; fb = fz;
      if (fz == 0.0) break;
    }
    zr = b;
  }
  c->z_reion = zr;
  if (!skip_reion) {
    for (int j = 0; j < N; ++j) {
      double l = x0 + j * dx;
      xe[j] += xe_reion(l, zr, Delta_z_reion, z_reion_He, Delta_z_reion_He, exp_reion, c->YHe);
    }
  }
  for (int j = N; j < c->xe_n; ++j) xe[j] = xe[N - 1];
  c->xe_last = xe[N - 1];
  c->tm_last = tm[hist.n_tm - 1];
  // a caller that freezes the grids at another cosmology's (both set, as
  // negative ln a, on input) keeps them and skips the two scans below
  if (c->lna_rec < 0.0 && c->lna_transfer_start < 0.0) return ABCMB_OK;
  // lna_rec: argmax of the visibility on linspace(-8, -4, 1500)
  Kappa K;
  K.build(*c);
  {
    int best = 0;
    double bv = -1.0;
    for (int i = 0; i < 1500; ++i) {
      double l = -8.0 + 4.0 * i / 1499.0;
      double g = expmkappa(K, l) / tau_c_of(*c, l);
      if (g > bv) { bv = g; best = i; }
    }
    c->lna_rec = -8.0 + 4.0 * best / 1499.0;
  }
  {
    int best = 0;
    double bv = 1e300;
    for (int i = 0; i < 5000; ++i) {
      double l = -15.0 + 9.0 * i / 4999.0;
      double v = aH(*c, l) * tau_c_of(*c, l) - 0.008;
      v = v * v;
      if (v < bv) { bv = v; best = i; }
    }
    c->lna_transfer_start = -15.0 + 9.0 * best / 4999.0;
  }
  return ABCMB_OK;
}

}  // namespace

extern "C" int abcmb_host_background(AbcmbCosmo* c, double tau_reion, double z_reion_in,
                                     double Delta_z_reion, double z_reion_He,
                                     double Delta_z_reion_He, double exp_reion, double n_eff,
                                     int32_t* rec_switch) {
  return background(c, tau_reion, z_reion_in, Delta_z_reion, z_reion_He, Delta_z_reion_He, exp_reion, n_eff,
                    rec_switch, false);
}

// The same tables without the reionization: x_e is the recombination history
// alone and z_reion is left 0; the scans for lna_rec and lna_transfer_start
// do not see reionization (no free electrons from it at z > 50, and its
// constant factor e^-tau_reio does not move the visibility's peak).
extern "C" int abcmb_host_recombination(AbcmbCosmo* c, double n_eff, int32_t* rec_switch) {
  return background(c, 0.0, 0.0, 0.5, 3.5, 0.5, 1.5, n_eff, rec_switch, true);
}

extern "C" int abcmb_host_los_background(const AbcmbCosmo* c, const double* lna, int n, double* out) {
  if (!c || !lna || !out || n <= 0) return ABCMB_EINVAL;
  Kappa K;
  K.build(*c);
  for (int i = 0; i < n; ++i) {
    double l = lna[i];
    double tc = tau_c_of(*c, l);
    double ah = aH(*c, l);
    double em = expmkappa(K, l);
    double g = em / tc;
    // d g / d lna: exp(-kappa) gains 1/(tau_c aH); tau_c ~ a^2 / x_e
    double xe = xe_of(*c, l);
    double dlntc = 2.0 - dxe_dlna(*c, l) / xe;
    double gp = (l < -10.0) ? 0.0 : g * (1.0 / (tc * ah) - dlntc);
    out[0 * n + i] = tau_of(*c, l);
    out[1 * n + i] = g;
    out[2 * n + i] = gp;
    out[3 * n + i] = ah;
    out[4 * n + i] = em;
    out[5 * n + i] = aH_prime(*c, l) * ah;
  }
  return ABCMB_OK;
}

extern "C" int abcmb_host_lna_start(const AbcmbCosmo* c, const double* k, int nk, double R_tc,
                                    double R_large, double* out) {
  if (!c || !k || !out) return ABCMB_EINVAL;
  const int n = 10000;
  std::vector<double> grid(n), f1(n), ah(n);
  for (int i = 0; i < n; ++i) {
    grid[i] = -20.0 + 10.0 * i / (n - 1);
    ah[i] = aH(*c, grid[i]);
    f1[i] = tau_c_of(*c, grid[i]) * ah[i];
  }
  // numpy/jax interp: clamp outside, linear inside (xp ascending); xp(i) is
  // evaluated where the search probes it
  auto interp = [&](double x, auto xp) {
    if (x <= xp(0)) return grid[0];
    if (x >= xp(n - 1)) return grid[n - 1];
    int lo = 0, hi = n;  // first index with xp > x
    while (lo < hi) {
      const int m = (lo + hi) / 2;
      if (xp(m) > x) hi = m; else lo = m + 1;
    }
    const int i = lo - 1;
    double t = (x - xp(i)) / (xp(i + 1) - xp(i));
    return grid[i] + t * (grid[i + 1] - grid[i]);
  };
  double l1 = interp(R_tc, [&](int i) { return f1[i]; });
  for (int j = 0; j < nk; ++j) {
    const double kj = k[j];
    double l2 = interp(R_large, [&](int i) { return kj / ah[i]; });
    double l = std::min(l1, l2);
    out[j] = std::min(l, -10.0);
  }
  return ABCMB_OK;
}

extern "C" int abcmb_host_eval(const AbcmbCosmo* c, const double* lna, int n, double* out) {
  if (!c || !lna || !out) return ABCMB_EINVAL;
  for (int i = 0; i < n; ++i) {
    double l = lna[i];
    Rho r = rho_all(*c, l);
    out[0 * n + i] = aH(*c, l);
    out[1 * n + i] = tau_of(*c, l);
    out[2 * n + i] = tau_c_of(*c, l);
    out[3 * n + i] = cs2_of(*c, l);
    out[4 * n + i] = xe_of(*c, l);
    out[5 * n + i] = tm_of(*c, l);
    out[6 * n + i] = 4.0 * r.g / 3.0 / r.b;
  }
  return ABCMB_OK;
}

// Baryon drag epoch and the sound horizon there (abcmb/background.py:774-936):
// kappa_d(ln a) = int_{ln a}^{0} dln a' / (tau_c aH R), R = 3 rho_b / (4 rho_g);
// z_d where kappa_d = 1; r_s(ln a) = r_s(-33) + int_{-33}^{ln a} c_s / aH dln a',
// c_s = 1 / sqrt(3 (1 + R)), r_s(-33) = 1 / (sqrt(3) aH(-33)).  The reference
// integrates both with an adaptive Runge-Kutta at rtol 1e-3; here Gauss-Legendre
// panels converge them to ~1e-12, so the two agree to the reference's tolerance.
extern "C" int abcmb_host_drag(const AbcmbCosmo* c, double* z_d, double* rs_d) {
  if (!c || !z_d || !rs_d) return ABCMB_EINVAL;
  auto R = [&](double l) {
    Rho r = rho_all(*c, l);
    return 3.0 * r.b / (4.0 * r.g);
  };
  auto dkap = [&](double l) { return 1.0 / (tau_c_of(*c, l) * aH(*c, l) * R(l)); };
  const double step = 1e-3;
  double hi = 0.0, kap = 0.0, lna_d = 0.0;
  bool found = false;
  while (hi > -12.0) {
    const double lo = hi - step;
    const double seg = gl8(dkap, lo, hi);
    if (kap + seg >= 1.0) {
      // kappa_d(l) = kap + int_l^hi: bisect for kappa_d = 1 inside the panel
      double a = lo, b = hi;
      for (int it = 0; it < 80; ++it) {
        const double m = 0.5 * (a + b);
        if (kap + gl8(dkap, m, hi) >= 1.0) a = m; else b = m;
      }
      lna_d = 0.5 * (a + b);
      found = true;
      break;
    }
    kap += seg;
    hi = lo;
  }
  if (!found) return ABCMB_EINVAL;
  auto drs = [&](double l) { return 1.0 / std::sqrt(3.0 * (1.0 + R(l))) / aH(*c, l); };
  const double l0 = ABCMB_TAU_TAB_LO;
  double rs = 1.0 / std::sqrt(3.0) / aH(*c, l0);
  const int n = (int)std::ceil((lna_d - l0) / 0.01);
  const double h = (lna_d - l0) / n;
  for (int i = 0; i < n; ++i) rs += gl8(drs, l0 + i * h, l0 + (i + 1) * h);
  *z_d = std::exp(-lna_d) - 1.0;
  *rs_d = rs;
  return ABCMB_OK;
}
