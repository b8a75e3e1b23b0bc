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
#include <type_traits>
#include <utility>

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

// The producer's physics is written over the cosmology type C: AbcmbCosmo
// for values, or CosmoT<Tan> for values with their derivatives along the
// tangent directions (abcmb_host_background_tangent).  Sc<C> is its scalar.
using abcmb_ad::dbl;
using hyrex::Tan;
using std::exp;
using std::log;
using std::pow;
using std::sqrt;
using std::tanh;

// One cosmology over scalar S: AbcmbCosmo's parameters and results with their
// tangents, its tables over S; the grid metadata stays double.
template <class S>
struct CosmoT {
  S h, omega_b, omega_cdm, omega_Lambda;
  S A_s, n_s, k_pivot, TCMB0, YHe;
  S N_ur, T_ur, N_ncdm, T_ncdm, m_ncdm;
  S R_nu, R_b, om, omega_r, omega_m;
  S tau0;
  double xe_x0, xe_dx, xe_lna_last;
  S xe_last;
  double tm_x0, tm_dx, tm_lna_last;
  S tm_last;
  double lna_rec, lna_transfer_start;
  S z_reion;
  S fluid_par[ABCMB_FLUID_PARAMS];
  int32_t has_ncdm, n_q, xe_n, tm_n, species_off;
  const S* tau_tab;
  const S* xe_tab;
  const S* tm_tab;
};

template <class C>
using Sc = std::decay_t<decltype(std::declval<const C&>().h)>;

template <class S>
struct Rho {
  S de, cdm, b, g, ur, ncdm, p_ncdm;
  S fl, p_fl;  // user fluids (abcmb_amd/fluidgen.py), summed
};

template <class C>
inline Rho<Sc<C>> rho_all(const C& c, double lna) {
  using S = Sc<C>;
  Rho<S> r;
  double a = std::exp(lna);
  double a3 = a * a * a, a4 = a3 * a;
  r.de = (c.species_off & ABCMB_NO_LAMBDA) ? S(0.0) : c.omega_Lambda * RHO_UNIT;
  r.cdm = (c.species_off & ABCMB_NO_CDM) ? S(0.0) : c.omega_cdm * RHO_UNIT / a3;
  r.fl = r.p_fl = S(0.0);
  if (fluid_host::kActive) fluid_host::bg(c, lna, a, r.fl, r.p_fl);
  r.b = c.omega_b * RHO_UNIT / a3;
  S t04 = c.TCMB0 * c.TCMB0 * c.TCMB0 * c.TCMB0;
  r.g = PI * PI / 15.0 * t04 / a4 / HC3;
  S tu4 = c.T_ur * c.T_ur * c.T_ur * c.T_ur;
  r.ur = c.N_ur * 2.0 * 7.0 / 8.0 * PI * PI / 30.0 * tu4 * t04 / a4 / HC3;
  r.ncdm = S(0.0);
  r.p_ncdm = S(0.0);
  if (c.has_ncdm) {
    S T = c.T_ncdm * c.TCMB0 / a;
    S x = c.m_ncdm / T;
    S ir = S(0.0), ip = S(0.0);
    for (int i = 0; i < 5; ++i) {
      double q = Q5[i];
      S eps = sqrt(q * q + x * x);
      ir += W5[i] * (1.0 + EQ5[i]) / (q * q) * eps;
      ip += W5[i] * (1.0 + EQ5[i]) / eps;
    }
    S T4 = T * T * T * T;
    r.ncdm = c.N_ncdm * 4.0 * T4 / (PI * PI) * ir / (HBAR * HBAR * HBAR) / (C_CM * C_CM * C_CM);
    r.p_ncdm = c.N_ncdm * 4.0 / 3.0 * T4 / (PI * PI) * ip / (HBAR * HBAR * HBAR) / (C_CM * C_CM * C_CM);
  }
  return r;
}

template <class S> inline S rho_tot(const Rho<S>& r) { return r.de + r.cdm + r.b + r.g + r.ur + r.ncdm + r.fl; }
template <class S> inline S p_tot(const Rho<S>& r) { return -r.de + r.g / 3.0 + r.ur / 3.0 + r.p_ncdm + r.p_fl; }

template <class C>
inline Sc<C> hubble(const C& c, double lna) {
  return sqrt(8.0 * PI * GN * rho_tot(rho_all(c, lna)) / 3.0);
}
template <class C>
inline Sc<C> aH(const C& c, double lna) { return std::exp(lna) * hubble(c, lna) / C_MPC; }
template <class C>
inline Sc<C> aH_prime(const C& c, double lna) {
  using S = Sc<C>;
  Rho<S> r = rho_all(c, lna);
  double a = std::exp(lna);
  S ah = a * sqrt(8.0 * PI * GN * rho_tot(r) / 3.0) / C_MPC;
  return -4.0 * PI * GN * a * a / 3.0 / ah * (rho_tot(r) + 3.0 * p_tot(r)) / (C_MPC * C_MPC);
}
template <class C>
inline Sc<C> nH(const C& c, double lna) {
  double a = std::exp(lna);
  return (1.0 - c.YHe) * 3.0 * c.omega_b * H0_OVER_H * H0_OVER_H / 8.0 / PI / GN / MH / (a * a * a);
}

// fast_interp of abcmb/ABCMBTools.py:268-308 (index clip to [eps, n-1-eps]).
template <class S>
inline S fast_interp(double x, double lo, double hi, const S* f, int n) {
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
template <class S>
inline S fast_interp_slope(double x, double lo, double hi, const S* f, int n) {
  double i = (x - lo) / (hi - lo) * (n - 1);
  const double eps = 1e-6;
  if (i < eps || i > n - 1.0 - eps) return S(0.0);
  int il = (int)std::floor(i);
  int iu = std::min(il + 1, n - 1);
  return (f[iu] - f[il]) * (n - 1) / (hi - lo);
}

template <class C>
inline Sc<C> tau_of(const C& c, double lna) {
  return fast_interp(lna, ABCMB_TAU_TAB_LO, 0.0, c.tau_tab, ABCMB_TAU_TAB_N);
}
template <class C>
inline Sc<C> xe_of(const C& c, double lna) {
  if (lna < c.xe_x0) return c.xe_tab[0];
  if (lna >= c.xe_lna_last) return c.xe_last;
  return fast_interp(lna, c.xe_x0, c.xe_x0 + c.xe_n * c.xe_dx, c.xe_tab, c.xe_n);
}
template <class C>
inline Sc<C> dxe_dlna(const C& c, double lna) {
  if (lna < c.xe_x0 || lna >= c.xe_lna_last) return Sc<C>(0.0);
  return fast_interp_slope(lna, c.xe_x0, c.xe_x0 + c.xe_n * c.xe_dx, c.xe_tab, c.xe_n);
}
template <class C>
inline Sc<C> tau_c_of(const C& c, double lna) {
  using S = Sc<C>;
  double a = std::exp(lna);
  S ne = nH(c, lna) * xe_of(c, lna);
  return 1.0 / a / ne / SIGMA_T / C_CM * C_MPC;
}
template <class S>
inline S gamma_compton(S xe, S T, S YHe) {
  S fhe = YHe * MH / MHE / (1.0 - YHe);
  return xe / (1.0 + fhe + xe) * 8.0 * SIGMA_T * (4.0 * STEF) * T * T * T * T / (3.0 * ME);
}
template <class C>
inline Sc<C> tm_of(const C& c, double lna) {
  using S = Sc<C>;
  if (lna < c.tm_x0) {
    S T = c.TCMB0 / std::exp(lna);
    S xe = xe_of(c, lna);
    return T * (1.0 - hubble(c, lna) / gamma_compton(xe, T, c.YHe));
  }
  if (lna >= c.tm_lna_last) return c.tm_last;
  return fast_interp(lna, c.tm_x0, c.tm_x0 + c.tm_n * c.tm_dx, c.tm_tab, c.tm_n);
}
template <class C>
inline Sc<C> cs2_of(const C& c, double lna) {
  using S = Sc<C>;
  Rho<S> r = rho_all(c, lna);
  S tm = tm_of(c, lna);
  S tg = c.TCMB0 / std::exp(lna);
  S xe = xe_of(c, lna);
  S mu = MH / ((1.0 + xe) * (1.0 - c.YHe) + MH / MHE * c.YHe);
  S R = 4.0 * r.g / 3.0 / r.b;
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
auto gl4(F f, double a, double b) {
  double m = 0.5 * (a + b), hw = 0.5 * (b - a);
  decltype(f(a)) s = 0.0;
  for (int i = 0; i < 4; ++i) s += GL4W[i] * f(m + hw * GL4X[i]);
  return s * hw;
}

template <class F>
auto gl8(F f, double a, double b) {
  double m = 0.5 * (a + b), hw = 0.5 * (b - a);
  decltype(f(a)) s = 0.0;
  for (int i = 0; i < 8; ++i) s += GLW[i] * f(m + hw * GLX[i]);
  return s * hw;
}

// ---- conformal time table (abcmb/background.py:252-319) ----
template <class C>
void tau_table(const C& c, Sc<C>* tab) {
  using S = Sc<C>;
  const double lna_cut = -16.1;
  auto approx = [&](double l) { return std::exp(l) / (H0_OVER_H / C_MPC) / sqrt(c.omega_r); };
  auto inv_aH = [&](double l) { return 1.0 / aH(c, l); };
  const int n = ABCMB_TAU_TAB_N;
  double prev_l = lna_cut;
  S prev_tau = approx(lna_cut);
  for (int i = 0; i < n; ++i) {
    double l = ABCMB_TAU_TAB_LO + (0.0 - ABCMB_TAU_TAB_LO) * i / (n - 1);
    if (i == n - 1) l = 0.0;
    if (l <= lna_cut) {
      tab[i] = approx(l);
      continue;
    }
    // integrate from prev_l to l in sub-intervals of at most 2e-3
    int nsub = std::max(1, (int)std::ceil((l - prev_l) / 4e-3));
    double h = (l - prev_l) / nsub;
    S t = prev_tau;
    for (int s = 0; s < nsub; ++s) t += gl4(inv_aH, prev_l + s * h, prev_l + (s + 1) * h);
    tab[i] = t;
    prev_l = l;
    prev_tau = t;
  }
}

// ---- recombination: HyRex (hyrex.cpp) on its grid ----
int rec_points() { return hyrex::XE_ALLOC; }

// ---- reionization (abcmb/background.py:939-1017) ----
template <class S, class Z>
auto xe_reion(double lna, Z z_re, double dz, double zHe, double dzHe, double ex, S YHe) {
  S fHe = YHe / 4.0 / (1.0 - YHe);
  double z = 1.0 / std::exp(lna) - 1.0;
  double y = std::pow(1.0 + z, ex);
  Z yre = pow(1.0 + z_re, ex);
  Z dy = ex * pow(1.0 + z_re, ex - 1.0) * dz;
  auto xH = (1.0 + fHe) / 2.0 * (1.0 + tanh((yre - y) / dy));
  S xHe = fHe / 2.0 * (1.0 + std::tanh((zHe - z) / dzHe));
  return xH + xHe;
}
// tau_reion(z_re) on a fixed ln a grid; the pieces of the integrand that do
// not depend on z_re (n_H, a H, (1 + z)^ex, the helium term) are evaluated
// once per cosmology, and each root-finder call only redoes the hydrogen
// step -- the same operations, in the same order, as a direct evaluation
template <class S>
struct ReionGrid {
  static constexpr int n = 2000;
  std::vector<S> nh, ah, xhe;
  std::vector<double> ea, y;
  S fHe;
  template <class C>
  void build(const C& c, double zHe, double dzHe, double ex) {
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
  // the same grid's values alone
  ReionGrid<double> values() const {
    ReionGrid<double> g;
    auto v = [](const std::vector<S>& a) {
      std::vector<double> o(a.size());
      for (size_t i = 0; i < a.size(); ++i) o[i] = dbl(a[i]);
      return o;
    };
    g.nh = v(nh); g.ah = v(ah); g.xhe = v(xhe); g.ea = ea; g.y = y; g.fHe = dbl(fHe);
    return g;
  }
  template <class Z>
  auto tau(Z z_re, double dz, double ex) const {
    Z yre = pow(1.0 + z_re, ex);
    Z dy = ex * pow(1.0 + z_re, ex - 1.0) * dz;
    using R = decltype(nh[0] * yre);
    R s = 0.0, prev = 0.0;
    for (int i = 0; i < n; ++i) {
      auto xH = (1.0 + fHe) / 2.0 * (1.0 + tanh((yre - y[i]) / dy));
      R ne = nh[i] * (xH + xhe[i]);
      R gam = ea[i] * ne * SIGMA_T * C_CM / C_MPC;
      R f = gam / ah[i];
      if (i > 0) s = s + 0.5 * (f + prev) * (5.0 / (n - 1));
      prev = f;
    }
    return s;
  }
};

// z_reion where tau_reion(z) = target: bracketing, then a safeguarded secant
// (Illinois) to 1e-12
double solve_zreion(const ReionGrid<double>& G, double target, double Delta_z, double ex) {
  auto f = [&](double z) { return G.tau(z, Delta_z, ex) - target; };
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
  return b;
}

// with derivatives: the root of the values, and dz/dtheta from the implicit
// function theorem, -(d tau/d theta - d target/d theta) / (d tau / dz)
Tan solve_zreion(const ReionGrid<Tan>& G, Tan target, double Delta_z, double ex) {
  const double z0 = solve_zreion(G.values(), target.v, Delta_z, ex);
  const Tan F = G.tau(Tan(z0), Delta_z, ex) - target;
  const double dFdz = G.values().tau(abcmb_ad::Dl<double>(z0, 1.0), Delta_z, ex).d;
  Tan z(z0);
  for (int i = 0; i < ABCMB_TANGENT_DIRS; ++i) z.d[i] = -F.d[i] / dFdz;
  return z;
}

// kappa(lna) = int_lna^0 dl / (tau_c aH): composite GL on a node grid
template <class C>
struct Kappa {
  using S = Sc<C>;
  std::vector<double> nodes;
  std::vector<S> cum;  // cum[i] = kappa(nodes[i])
  const C* c;
  S integrand(double l) const { return 1.0 / tau_c_of(*c, l) / aH(*c, l); }
  void build(const C& cc) {
    c = &cc;
    const double lo = -10.0;
    const int n = 20001;  // 5e-4 spacing
    nodes.resize(n);
    cum.assign(n, S(0.0));
    for (int i = 0; i < n; ++i) nodes[i] = lo + (0.0 - lo) * i / (n - 1);
    nodes[n - 1] = 0.0;
    for (int i = n - 2; i >= 0; --i)
      cum[i] = cum[i + 1] + gl4([&](double l) { return integrand(l); }, nodes[i], nodes[i + 1]);
  }
  S operator()(double l) const {
    const int n = (int)nodes.size();
    if (l <= nodes[0]) return cum[0];
    if (l >= 0.0) return S(0.0);
    double fi = (l - nodes[0]) / (nodes[1] - nodes[0]);
    int i = std::min((int)fi, n - 2);
    return cum[i + 1] + gl4([&](double x) { return integrand(x); }, l, nodes[i + 1]);
  }
};

template <class C>
Sc<C> expmkappa(const Kappa<C>& K, double l) { return l < -10.0 ? Sc<C>(0.0) : exp(-K(l)); }

// The tables of one cosmology.  C: AbcmbCosmo, or CosmoT<Tan> with every grid
// choice (stage ends, lna_rec, lna_transfer_start) given.  tau, xe, tm: the
// table storage; n_eff and tau_reion over the same scalar.
template <class C>
int background(C* c, Sc<C>* tau, Sc<C>* xe, Sc<C>* tm, Sc<C> tau_reion, double z_reion_in, double Delta_z_reion,
               double z_reion_He, double Delta_z_reion_He, double exp_reion, Sc<C> n_eff, int32_t* rec_switch,
               bool skip_reion) {
  using S = Sc<C>;
  if (c->xe_n < hyrex::XE_ALLOC || c->tm_n < hyrex::TM_ALLOC) return ABCMB_EINVAL;
  tau_table(*c, tau);
  c->tau0 = fast_interp(0.0, ABCMB_TAU_TAB_LO, 0.0, tau, ABCMB_TAU_TAB_N);
  // HyRex's inputs on its grid (abcmb/background.py:88-99)
  const double x0 = -std::log(1.0 + hyrex::Z0), dx = hyrex::DLNA;
  std::vector<S> Tg(hyrex::N_GRID), nhg(hyrex::N_GRID), Hg(hyrex::N_GRID);
  for (int j = 0; j < hyrex::N_GRID; ++j) {
    const double l = x0 + j * dx;
    Tg[j] = c->TCMB0 / std::exp(l);
    nhg[j] = nH(*c, l);
    Hg[j] = hubble(*c, l);
  }
  int32_t sw[hyrex::N_STAGES];
  for (int s = 0; s < hyrex::N_STAGES; ++s) sw[s] = rec_switch ? rec_switch[s] : -1;
  hyrex::History hist;
  const int rc = hyrex::recombination<S>(c->YHe, c->omega_b, c->omega_cdm, n_eff, Tg.data(), nhg.data(), Hg.data(),
                                         hyrex::N_GRID, xe, c->xe_n, tm, c->tm_n, sw, hist);
  if (rc != ABCMB_OK) return rc;
  if (rec_switch) for (int s = 0; s < hyrex::N_STAGES; ++s) rec_switch[s] = sw[s];
  const int N = hist.n_xe;
  c->xe_x0 = x0;
  c->xe_dx = c->tm_dx = dx;
  c->tm_x0 = hist.tm_x0;
  c->xe_lna_last = x0 + (N - 1) * dx;
  c->tm_lna_last = hist.tm_x0 + (hist.n_tm - 1) * dx;
  // reionization: solve tau_reion(z) = target, or take z
  S zr = S(z_reion_in);
  if (dbl(tau_reion) > 0.0 && !skip_reion) {
    ReionGrid<S> G;
    G.build(*c, z_reion_He, Delta_z_reion_He, exp_reion);
    zr = solve_zreion(G, tau_reion, Delta_z_reion, exp_reion);
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
  if constexpr (!std::is_same_v<C, AbcmbCosmo>) {
    return ABCMB_EINVAL;  // derivatives need the grids held
  } else {
    // lna_rec: argmax of the visibility on linspace(-8, -4, 1500)
    Kappa<C> K;
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
}

// the LOS background vectors on lna: tau, g, g', aH, exp(-kappa), aH' aH
template <class C>
void los_vectors(const C& c, const double* lna, int n, Sc<C>* out) {
  using S = Sc<C>;
  Kappa<C> K;
  K.build(c);
  for (int i = 0; i < n; ++i) {
    double l = lna[i];
    S tc = tau_c_of(c, l);
    S ah = aH(c, l);
    S em = expmkappa(K, l);
    S g = em / tc;
    // d g / d lna: exp(-kappa) gains 1/(tau_c aH); tau_c ~ a^2 / x_e
    S xe = xe_of(c, l);
    S dlntc = 2.0 - dxe_dlna(c, l) / xe;
    S gp = (l < -10.0) ? S(0.0) : g * (1.0 / (tc * ah) - dlntc);
    out[0 * n + i] = tau_of(c, l);
    out[1 * n + i] = g;
    out[2 * n + i] = gp;
    out[3 * n + i] = ah;
    out[4 * n + i] = em;
    out[5 * n + i] = aH_prime(c, l) * ah;
  }
}

// aH, tau, tau_c, cs2, x_e, T_m, R on lna
template <class C>
void eval_all(const C& c, const double* lna, int n, Sc<C>* out) {
  using S = Sc<C>;
  for (int i = 0; i < n; ++i) {
    double l = lna[i];
    Rho<S> r = rho_all(c, l);
    out[0 * n + i] = aH(c, l);
    out[1 * n + i] = tau_of(c, l);
    out[2 * n + i] = tau_c_of(c, l);
    out[3 * n + i] = cs2_of(c, l);
    out[4 * n + i] = xe_of(c, l);
    out[5 * n + i] = tm_of(c, l);
    out[6 * n + i] = 4.0 * r.g / 3.0 / r.b;
  }
}

// ---- the tangent producer's cosmology: AbcmbCosmo's doubles (h .. z_reion,
// then fluid_par, in declaration order) with their derivatives dpar[d][f]
#define ABCMB_COSMO_SCALARS(X) \
  X(h) X(omega_b) X(omega_cdm) X(omega_Lambda) X(A_s) X(n_s) X(k_pivot) X(TCMB0) X(YHe) X(N_ur) X(T_ur) \
  X(N_ncdm) X(T_ncdm) X(m_ncdm) X(R_nu) X(R_b) X(om) X(omega_r) X(omega_m) X(tau0) X(xe_x0) X(xe_dx)    \
  X(xe_lna_last) X(xe_last) X(tm_x0) X(tm_dx) X(tm_lna_last) X(tm_last) X(lna_rec) X(lna_transfer_start) \
  X(z_reion)

template <class T> void set_field(T& dst, const double* dpar, int f, double v);
template <> void set_field<double>(double& dst, const double*, int, double v) { dst = v; }
template <> void set_field<Tan>(Tan& dst, const double* dpar, int f, double v) {
  dst = Tan(v);
  for (int d = 0; d < ABCMB_TANGENT_DIRS; ++d) dst.d[d] = dpar[d * ABCMB_COSMO_DOUBLES + f];
}
inline void get_field(double v, double* val, double* dpar, int f) {
  *val = v;
  for (int d = 0; d < ABCMB_TANGENT_DIRS; ++d) dpar[d * ABCMB_COSMO_DOUBLES + f] = 0.0;
}
inline void get_field(const Tan& v, double* val, double* dpar, int f) {
  *val = v.v;
  for (int d = 0; d < ABCMB_TANGENT_DIRS; ++d) dpar[d * ABCMB_COSMO_DOUBLES + f] = v.d[d];
}

CosmoT<Tan> with_tangents(const AbcmbCosmo& c, const double* dpar) {
  CosmoT<Tan> t;
  int f = 0;
#define X(name) set_field(t.name, dpar, f++, c.name);
  ABCMB_COSMO_SCALARS(X)
#undef X
  for (int i = 0; i < ABCMB_FLUID_PARAMS; ++i) set_field(t.fluid_par[i], dpar, f++, c.fluid_par[i]);
  t.has_ncdm = c.has_ncdm;
  t.n_q = c.n_q;
  t.xe_n = c.xe_n;
  t.tm_n = c.tm_n;
  t.species_off = c.species_off;
  t.tau_tab = t.xe_tab = t.tm_tab = nullptr;
  return t;
}

void store_tangents(const CosmoT<Tan>& t, AbcmbCosmo& c, double* dpar) {
  int f = 0;
#define X(name) get_field(t.name, &c.name, dpar, f++);
  ABCMB_COSMO_SCALARS(X)
#undef X
  for (int i = 0; i < ABCMB_FLUID_PARAMS; ++i) get_field(t.fluid_par[i], &c.fluid_par[i], dpar, f++);
}

// the cosmology of c with tables tau, xe, tm (values and their derivatives)
CosmoT<Tan> tangent_cosmo(const AbcmbCosmo& c, const double* dpar, const double* dtau, const double* dxe,
                          const double* dtm, std::vector<Tan>& tau, std::vector<Tan>& xe, std::vector<Tan>& tm) {
  auto load = [](const double* v, const double* d, int n, std::vector<Tan>& out) {
    out.resize(n);
    for (int i = 0; i < n; ++i) {
      out[i] = Tan(v[i]);
      for (int k = 0; k < ABCMB_TANGENT_DIRS; ++k) out[i].d[k] = d[(size_t)k * n + i];
    }
  };
  load(c.tau_tab, dtau, ABCMB_TAU_TAB_N, tau);
  load(c.xe_tab, dxe, c.xe_n, xe);
  load(c.tm_tab, dtm, c.tm_n, tm);
  CosmoT<Tan> t = with_tangents(c, dpar);
  t.tau_tab = tau.data();
  t.xe_tab = xe.data();
  t.tm_tab = tm.data();
  return t;
}

void split(const std::vector<Tan>& a, double* val, double* d) {
  const int n = (int)a.size();
  for (int i = 0; i < n; ++i) {
    val[i] = a[i].v;
    for (int k = 0; k < ABCMB_TANGENT_DIRS; ++k) d[(size_t)k * n + i] = a[i].d[k];
  }
}

}  // namespace

extern "C" int abcmb_host_recomb_points(void) { return rec_points(); }
extern "C" int abcmb_host_tm_points(void) { return hyrex::TM_ALLOC; }

extern "C" const char* abcmb_host_fluid_signature(void) { return fluid_host::kSignature; }


extern "C" int abcmb_host_background(AbcmbCosmo* c, double tau_reion, double z_reion_in,
                                     double Delta_z_reion, double z_reion_He,
                                     double Delta_z_reion_He, double exp_reion, double n_eff,
                                     int32_t* rec_switch) {
  if (!c || !c->tau_tab || !c->xe_tab || !c->tm_tab) return ABCMB_EINVAL;
  return background(c, const_cast<double*>(c->tau_tab), const_cast<double*>(c->xe_tab),
                    const_cast<double*>(c->tm_tab), tau_reion, z_reion_in, Delta_z_reion, z_reion_He,
                    Delta_z_reion_He, exp_reion, n_eff, rec_switch, false);
}

// The same tables without the reionization: x_e is the recombination history
// alone and z_reion is left 0; the scans for lna_rec and lna_transfer_start
// do not see reionization (no free electrons from it at z > 50, and its
// constant factor e^-tau_reio does not move the visibility's peak).
extern "C" int abcmb_host_recombination(AbcmbCosmo* c, double n_eff, int32_t* rec_switch) {
  if (!c || !c->tau_tab || !c->xe_tab || !c->tm_tab) return ABCMB_EINVAL;
  return background(c, const_cast<double*>(c->tau_tab), const_cast<double*>(c->xe_tab),
                    const_cast<double*>(c->tm_tab), 0.0, 0.0, 0.5, 3.5, 0.5, 1.5, n_eff, rec_switch, true);
}

extern "C" int abcmb_host_los_background(const AbcmbCosmo* c, const double* lna, int n, double* out) {
  if (!c || !lna || !out || n <= 0) return ABCMB_EINVAL;
  los_vectors(*c, lna, n, out);
  return ABCMB_OK;
}

// ---- derivatives of the producer along the tangent directions (forward mode)

extern "C" int abcmb_host_background_tangent(AbcmbCosmo* c, double tau_reion, const double* dtau_reion,
                                             double z_reion_in, double Delta_z_reion, double z_reion_He,
                                             double Delta_z_reion_He, double exp_reion, double n_eff,
                                             const double* dn_eff, const int32_t* rec_switch, int reionization,
                                             double* dpar, double* dtau, double* dxe, double* dtm) {
  if (!c || !c->tau_tab || !c->xe_tab || !c->tm_tab || !rec_switch || !dpar || !dtau || !dxe || !dtm ||
      !dtau_reion || !dn_eff)
    return ABCMB_EINVAL;
  if (!(c->lna_rec < 0.0 && c->lna_transfer_start < 0.0)) return ABCMB_EINVAL;
  int32_t sw[hyrex::N_STAGES];
  for (int s = 0; s < hyrex::N_STAGES; ++s) {
    if (rec_switch[s] < 0) return ABCMB_EINVAL;
    sw[s] = rec_switch[s];
  }
  CosmoT<Tan> t = with_tangents(*c, dpar);
  std::vector<Tan> tau(ABCMB_TAU_TAB_N), xe(c->xe_n), tm(c->tm_n);
  t.tau_tab = tau.data();
  t.xe_tab = xe.data();
  t.tm_tab = tm.data();
  Tan tr(tau_reion), ne(n_eff);
  for (int d = 0; d < ABCMB_TANGENT_DIRS; ++d) {
    tr.d[d] = dtau_reion[d];
    ne.d[d] = dn_eff[d];
  }
  const int rc = reionization
                     ? background(&t, tau.data(), xe.data(), tm.data(), tr, z_reion_in, Delta_z_reion, z_reion_He,
                                  Delta_z_reion_He, exp_reion, ne, sw, false)
                     : background(&t, tau.data(), xe.data(), tm.data(), Tan(0.0), 0.0, 0.5, 3.5, 0.5, 1.5, ne, sw,
                                  true);
  if (rc != ABCMB_OK) return rc;
  store_tangents(t, *c, dpar);
  split(tau, const_cast<double*>(c->tau_tab), dtau);
  split(xe, const_cast<double*>(c->xe_tab), dxe);
  split(tm, const_cast<double*>(c->tm_tab), dtm);
  return ABCMB_OK;
}

extern "C" int abcmb_host_los_background_tangent(const AbcmbCosmo* c, const double* dpar, const double* dtau,
                                                 const double* dxe, const double* dtm, const double* lna, int n,
                                                 double* out, double* dout) {
  if (!c || !dpar || !dtau || !dxe || !dtm || !lna || !out || !dout || n <= 0) return ABCMB_EINVAL;
  std::vector<Tan> tau, xe, tm, o(6 * (size_t)n);
  const CosmoT<Tan> t = tangent_cosmo(*c, dpar, dtau, dxe, dtm, tau, xe, tm);
  los_vectors(t, lna, n, o.data());
  split(o, out, dout);
  return ABCMB_OK;
}

extern "C" int abcmb_host_eval_tangent(const AbcmbCosmo* c, const double* dpar, const double* dtau,
                                       const double* dxe, const double* dtm, const double* lna, int n, double* out,
                                       double* dout) {
  if (!c || !dpar || !dtau || !dxe || !dtm || !lna || !out || !dout || n <= 0) return ABCMB_EINVAL;
  std::vector<Tan> tau, xe, tm, o(7 * (size_t)n);
  const CosmoT<Tan> t = tangent_cosmo(*c, dpar, dtau, dxe, dtm, tau, xe, tm);
  eval_all(t, lna, n, o.data());
  split(o, out, dout);
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
  eval_all(*c, lna, n, out);
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
