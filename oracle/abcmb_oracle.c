/*
 * abcmb_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of ABCMB's perturbation + line-of-sight hot path,
 * used solely as the checker for the HIP library (tests/, the smoke check
 * and the CPU-baseline leg of bench.py).  Nothing in the product links it.
 *
 * Each function follows the reference function cited next to it (paths
 * relative to the reference repository root).  The integrator follows the
 * published Kvaerno5 ESDIRK (Kvaerno 2004, BIT 44:489) with a PID step
 * controller and chord Newton as diffrax is configured by the reference
 * (abcmb/perturbations.py:285-312); diffrax itself is absent, so the
 * details below are this project's and are stated in the design notes:
 *   - error order 5 in the PID exponents, safety 0.9, factor in
 *     [0.2 (rejected) or 1 (accepted), 10];
 *   - chord Newton on each implicit stage with the matrix
 *     I - h*gamma*J(t_n + jac_offset*h) (jac_offset = 0.6: the middle of the
 *     stage abscissae 0..1.23), J the exact (linear) Jacobian, stage
 *     predictor base + h*gamma*F_{i-1}, convergence
 *     theta/(1-theta)*|dz| < kappa after two or more iterations (or
 *     |dz| < 1e-9, the rounding floor), divergence theta > 2;
 *   - stage derivative F_i = (z_i - base_i) / (h*gamma);
 *   - cubic Hermite dense output with exact end derivatives.
 * Linear solves use LU with partial pivoting (the GPU uses Gauss-Jordan
 * inverses: an independent route to the same iterates).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <omp.h>

#include "../include/abcmb.h"

#define PI 3.14159265358979323846
static const double C_CM = 29979245800.0;
static const double C_MPC = 9.71561e-15;
static const double H0_OVER_H = 3.24078e-18;
static const double HBAR = 6.582119569509e-16;
static const double GN = 1.18980205e-40;
static const double MP = 938271999.0;
static const double ME = 510998.9461;
static const double MHE = 3.72839e9;
static const double SIGMA_T = 6.652458734e-25;

static double MH(void) { return MP + ME + 13.5982860719383; }
static double RHO_UNIT(void) { return 3.0 * H0_OVER_H * H0_OVER_H / 8.0 / PI / GN; }
static double STEF(void) { return PI * PI / (60.0 * HBAR * HBAR * HBAR * C_CM * C_CM); }

static const double Q3[3] = {0.913201, 3.37517, 7.79184};
static const double W3[3] = {0.0687359, 3.31435, 2.29911};
static const double Q5[5] = {0.583165, 2.0, 4.0, 7.26582, 13.0};
static const double W5[5] = {0.0081201, 0.689407, 2.8063, 2.05156, 0.12681};

/* ---------- background functions (abcmb/background.py, abcmb/species.py) ---------- */

/* abcmb/ABCMBTools.py:268-308 */
static double fast_interp(double x, double lo, double hi, const double* f, int n) {
  double i = (x - lo) / (hi - lo) * (n - 1);
  double eps = 1e-6;
  if (i < eps) i = eps;
  if (i > n - 1.0 - eps) i = n - 1.0 - eps;
  int il = (int)floor(i);
  int iu = il + 1 < n - 1 ? il + 1 : n - 1;
  double wu = i - il;
  return (1.0 - wu) * f[il] + wu * f[iu];
}

typedef struct {
  double a, de, cdm, b, g, ur, ncdm, p_ncdm, T_nu, x_nu;
  double fl, p_fl, m_fl;  /* user fluids: density, pressure, matter density (summed) */
} Rho;

/* User-defined fluids (abcmb/species.py:11-249): the species methods are
 * evaluated as written by the caller (tests: the model's Fluid objects called
 * with numbers, their parameters read from the cosmology c) through these
 * callbacks; their rows are the last n_rows rows of the state.  NULL: the
 * standard species only. */
typedef struct OracleFluids {
  int n_rows;
  void* ctx;
  /* out[0..2]: summed rho, P and matter rho of the user fluids at ln a */
  void (*bg)(void* ctx, const AbcmbCosmo* c, double lna, double* out);
  /* out[0..3]: sum rho delta, (rho + P) theta, (rho + P) sigma and the matter
   * rho delta of the user fluids in state y (n rows) at ln a */
  void (*sums)(void* ctx, const AbcmbCosmo* c, double lna, const double* y, int n, double* out);
  /* dy[0..n_rows): y_prime of the user rows at (k, ln a) given aH, h', eta' */
  void (*rhs)(void* ctx, const AbcmbCosmo* c, double k, double lna, double aH, double hp, double ep,
              const double* y, int n, double* dy);
  /* y[0..n_rows): y_ini of the user rows at (k, tau) */
  void (*ic)(void* ctx, const AbcmbCosmo* c, double k, double tau, double* y);
} OracleFluids;
static const OracleFluids* g_fl = NULL;

/* installs (or, with NULL, removes) the user-fluid callbacks; the caller keeps
 * them alive and runs the oracle single-threaded while they are set */
int oracle_set_fluids(const OracleFluids* f) {
  g_fl = f;
  return 0;
}

/* rho of each species in species order (abcmb/species.py:454-1456) */
static Rho rho_all(const AbcmbCosmo* c, double lna) {
  Rho r;
  double a = exp(lna);
  double hc3 = pow(C_CM * HBAR, 3.0);
  r.a = a;
  r.de = (c->species_off & ABCMB_NO_LAMBDA) ? 0.0 : c->omega_Lambda * RHO_UNIT();
  r.cdm = (c->species_off & ABCMB_NO_CDM) ? 0.0 : c->omega_cdm * RHO_UNIT() / pow(a, 3.0);
  r.fl = r.p_fl = r.m_fl = 0.0;
  if (g_fl) {
    double o[3];
    g_fl->bg(g_fl->ctx, c, lna, o);
    r.fl = o[0]; r.p_fl = o[1]; r.m_fl = o[2];
  }
  r.b = c->omega_b * RHO_UNIT() / pow(a, 3.0);
  r.g = PI * PI / 15.0 * pow(c->TCMB0, 4.0) / pow(a, 4.0) / hc3;
  r.ur = c->N_ur * 2.0 * 7.0 / 8.0 * PI * PI / 30.0 * pow(c->T_ur, 4.0) * pow(c->TCMB0, 4.0) / pow(a, 4.0) / hc3;
  r.ncdm = r.p_ncdm = 0.0;
  r.T_nu = c->T_ncdm * c->TCMB0 / a;
  r.x_nu = c->m_ncdm / r.T_nu;
  if (c->has_ncdm) {
    double ir = 0.0, ip = 0.0;
    for (int i = 0; i < 5; ++i) {
      double q = Q5[i];
      ir += (1.0 + exp(-q)) / (q * q) * sqrt(q * q + r.x_nu * r.x_nu) * W5[i];
      ip += (1.0 + exp(-q)) / sqrt(q * q + r.x_nu * r.x_nu) * W5[i];
    }
    r.ncdm = c->N_ncdm * 4.0 * pow(r.T_nu, 4.0) / (PI * PI) * ir / pow(HBAR, 3.0) / pow(C_CM, 3.0);
    r.p_ncdm = c->N_ncdm * 4.0 / 3.0 * pow(r.T_nu, 4.0) / (PI * PI) * ip / pow(HBAR, 3.0) / pow(C_CM, 3.0);
  }
  return r;
}

static double rho_tot(const Rho* r) {
  double s = 0.0;
  s += r->de; s += r->cdm; s += r->b; s += r->g; s += r->ur; s += r->ncdm; s += r->fl;
  return s;
}

/* abcmb/background.py:147-187 */
static double bg_H(const AbcmbCosmo* c, double lna) {
  Rho r = rho_all(c, lna);
  return sqrt(8.0 * PI * GN * rho_tot(&r) / 3.0);
}
static double bg_aH(const AbcmbCosmo* c, double lna) { return exp(lna) * bg_H(c, lna) / C_MPC; }
/* abcmb/background.py:321-347 */
static double bg_tau(const AbcmbCosmo* c, double lna) {
  return fast_interp(lna, ABCMB_TAU_TAB_LO, 0.0, c->tau_tab, ABCMB_TAU_TAB_N);
}
/* abcmb/background.py:349-367 */
static double bg_nH(const AbcmbCosmo* c, double lna) {
  return (1 - c->YHe) * 3.0 * c->omega_b * H0_OVER_H * H0_OVER_H / 8 / PI / GN / MH() / pow(exp(lna), 3.0);
}
/* abcmb/background.py:568-603 (span quirk: x0 + len(arr)*dx) */
static double bg_xe(const AbcmbCosmo* c, double lna) {
  if (lna < c->xe_x0) return c->xe_tab[0];
  if (lna >= c->xe_lna_last) return c->xe_last;
  return fast_interp(lna, c->xe_x0, c->xe_x0 + c->xe_n * c->xe_dx, c->xe_tab, c->xe_n);
}
/* abcmb/background.py:659-680 */
static double bg_tau_c(const AbcmbCosmo* c, double lna) {
  double a = exp(lna);
  double ne = bg_nH(c, lna) * bg_xe(c, lna);
  return 1.0 / a / ne / SIGMA_T / C_CM * C_MPC;
}
/* abcmb/hyrex/recomb_functions.py:127-153 */
static double gamma_compton(double xe, double T, double YHe) {
  double fhe = YHe * MH() / MHE / (1 - YHe);
  return xe / (1.0 + fhe + xe) * 8.0 * SIGMA_T * (4.0 * STEF()) * pow(T, 4.0) / (3.0 * ME);
}
/* abcmb/background.py:605-657 */
static double bg_Tm(const AbcmbCosmo* c, double lna) {
  if (lna < c->tm_x0) {
    double T = c->TCMB0 / exp(lna);
    double xe = bg_xe(c, lna);
    return T * (1.0 - bg_H(c, lna) / gamma_compton(xe, T, c->YHe));
  }
  if (lna >= c->tm_lna_last) return c->tm_last;
  return fast_interp(lna, c->tm_x0, c->tm_x0 + c->tm_n * c->tm_dx, c->tm_tab, c->tm_n);
}
/* abcmb/species.py:1163-1220 */
static double bg_cs2(const AbcmbCosmo* c, double lna) {
  Rho r = rho_all(c, lna);
  double Tm = bg_Tm(c, lna);
  double Tg = c->TCMB0 / exp(lna);
  double denom = (1.0 + bg_xe(c, lna)) * (1.0 - c->YHe) + MH() / MHE * c->YHe;
  double mu = MH() / denom;
  double R = 4.0 * r.g / 3.0 / r.b;
  return Tm / mu * (5.0 / 3.0 - 2.0 / 3.0 * mu * R / ME / bg_aH(c, lna) / bg_tau_c(c, lna) * (Tg / Tm - 1.0));
}

/* ---------- layout ---------- */
typedef struct {
  const AbcmbCosmo* c;
  int lg, lp, lu, ln, nq, n;
  int ig, ip, iu, inc;
  double k;
  const double* q;
  const double* w;
} Ctx;

static void ctx_init(Ctx* x, const AbcmbCosmo* c, const AbcmbOdeSettings* s, double k) {
  x->c = c;
  x->lg = s->l_g; x->lp = s->l_pol; x->lu = s->l_ur; x->ln = s->l_ncdm; x->nq = s->n_q; x->n = s->n;
  x->ig = 4;
  x->ip = 4 + s->l_g + 1;
  x->iu = x->ip + s->l_pol + 1;
  x->inc = x->iu + s->l_ur + 1;
  x->k = k;
  x->q = s->n_q == 5 ? Q5 : Q3;
  x->w = s->n_q == 5 ? W5 : W3;
}

/* massive-nu sums (abcmb/species.py:979-1075) */
static void ncdm_sums(const Ctx* x, const Rho* r, const double* y, double* rd, double* rpt, double* rps) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  int nl = x->ln + 1;
  for (int i = 0; i < x->nq; ++i) {
    double q = x->q[i], w = x->w[i];
    double eps = sqrt(q * q + r->x_nu * r->x_nu);
    const double* P = y + x->inc + i * nl;
    s0 += w * (1.0 + exp(-q)) * eps / (q * q) * P[0];
    s1 += w * (1.0 + exp(-q)) / q * P[1];
    s2 += w * (1.0 + exp(-q)) / eps * P[2];
  }
  double pre = x->c->N_ncdm * pow(r->T_nu, 4.0) / (PI * PI) / pow(HBAR, 3.0) / pow(C_CM, 3.0);
  *rd = s0 * 4.0 * pre;
  *rpt = s1 * 4.0 * pre;
  *rps = s2 * 8.0 / 3.0 * pre;
}

/* PerturbationEvolver.get_derivatives + species y_prime
 * (abcmb/perturbations.py:201-248; abcmb/species.py:583-607, :707-753,
 *  :919-977, :1245-1285, :1390-1448) */
static void rhs(const Ctx* x, double lna, const double* y, double* dy) {
  const AbcmbCosmo* c = x->c;
  double k = x->k;
  Rho r = rho_all(c, lna);
  double a = r.a;
  double aH = bg_aH(c, lna);
  double tau = bg_tau(c, lna);
  double tau_c = bg_tau_c(c, lna);
  double srd = 0.0, srt = 0.0;
  double nrd = 0.0, nrt = 0.0, nrs = 0.0;
  if (c->has_ncdm && x->nq > 0) ncdm_sums(x, &r, y, &nrd, &nrt, &nrs);
  /* species order: DE, CDM, Baryon, Photon, Massless, [Massive] */
  srd += 0.0;
  srd += r.cdm * y[1];
  srd += r.b * y[2];
  srd += r.g * y[x->ig];
  srd += r.ur * y[x->iu];
  if (c->has_ncdm && x->nq > 0) srd += nrd;
  srt += 0.0;
  srt += (r.cdm + 0.0) * 0.0;
  srt += (r.b + 0.0) * y[3];
  srt += (r.g + r.g / 3.0) * y[x->ig + 1];
  srt += (r.ur + r.ur / 3.0) * y[x->iu + 1];
  if (c->has_ncdm && x->nq > 0) srt += nrt;
  if (g_fl && g_fl->n_rows > 0) {
    double o[4];
    g_fl->sums(g_fl->ctx, c, lna, y, x->n, o);
    srd += o[0];
    srt += o[1];
  }
  double hp = 2.0 / (aH * aH) * (k * k * y[0] + 4.0 * PI * GN * a * a / (C_MPC * C_MPC) * srd);
  double ep = 4.0 * PI * GN * a * a / aH / (k * k) * srt / (C_MPC * C_MPC);
  dy[0] = ep;
  if (g_fl && g_fl->n_rows > 0) g_fl->rhs(g_fl->ctx, c, k, lna, aH, hp, ep, y, x->n, dy + x->n - g_fl->n_rows);
  /* CDM */
  dy[1] = -0.5 * hp;
  /* Baryon */
  {
    double cs2 = bg_cs2(c, lna);
    double R = 4.0 * r.g / 3.0 / r.b;
    double db = y[2], tb = y[3], tg = y[x->ig + 1];
    dy[2] = -tb / aH - hp / 2.0;
    dy[3] = -tb + cs2 * k * k * db / aH + R / tau_c / aH * (tg - tb);
  }
  /* Photon */
  {
    const double* F = y + x->ig;
    const double* G = y + x->ip;
    double* dF = dy + x->ig;
    double* dG = dy + x->ip;
    int Fl = x->lg, Gl = x->lp;
    double tb = y[3];
    dF[0] = -4.0 / 3.0 / aH * F[1] - 2.0 / 3.0 * hp;
    dF[1] = k * k / aH * (F[0] / 4.0 - F[2]) + (tb - F[1]) / aH / tau_c;
    dF[2] = 4.0 / 15.0 / aH * F[1] - 3.0 / 10.0 * k / aH * F[3] + 2.0 / 15.0 * hp + 4.0 / 5.0 * ep
            - 9.0 / 10.0 / aH / tau_c * F[2] + (G[0] + G[2]) / 20.0 / aH / tau_c;
    dF[3] = k / 7.0 / aH * (6.0 * F[2] - 4.0 * F[4]) - F[3] / aH / tau_c;
    for (int l = 4; l < Fl; ++l)
      dF[l] = 1.0 / (2.0 * l + 1.0) * k / aH * (l * F[l - 1] - (l + 1) * F[l + 1]) - F[l] / aH / tau_c;
    dF[Fl] = k / aH * F[Fl - 1] - (Fl + 1) / aH / tau * F[Fl] - F[Fl] / aH / tau_c;
    double pi_src = (2.0 * F[2] + G[0] + G[2]) / 2.0 / aH / tau_c;
    for (int l = 0; l < Gl; ++l) {
      double gm1 = l > 0 ? G[l - 1] : 0.0;
      double wl = l == 0 ? 1.0 : (l == 2 ? 0.2 : 0.0);
      dG[l] = 1.0 / (2.0 * l + 1.0) * k / aH * (l * gm1 - (l + 1) * G[l + 1]) - G[l] / aH / tau_c + pi_src * wl;
    }
    dG[Gl] = k / aH * G[Gl - 1] - (Gl + 1) / aH / tau * G[Gl] - G[Gl] / aH / tau_c;
  }
  /* Massless neutrinos */
  {
    const double* N = y + x->iu;
    double* dN = dy + x->iu;
    int L = x->lu;
    dN[0] = -4.0 / 3.0 / aH * N[1] - 2.0 / 3.0 * hp;
    dN[1] = k * k / aH * (N[0] / 4.0 - N[2]);
    dN[2] = 4.0 / 15.0 / aH * N[1] - 3.0 / 10.0 * k / aH * N[3] + 2.0 / 15.0 * hp + 4.0 / 5.0 * ep;
    dN[3] = 1.0 / 7.0 * k / aH * (6.0 * N[2] - 4.0 * N[4]);
    for (int l = 4; l < L; ++l) dN[l] = 1.0 / (2.0 * l + 1.0) * k / aH * (l * N[l - 1] - (l + 1) * N[l + 1]);
    dN[L] = k / aH * N[L - 1] - (L + 1) / aH / tau * N[L];
  }
  /* Massive neutrinos */
  if (x->nq > 0) {
    int nl = x->ln + 1, L = x->ln;
    for (int i = 0; i < x->nq; ++i) {
      double q = x->q[i];
      double eps = sqrt(q * q + r.x_nu * r.x_nu);
      double dlf = -q / (1 + exp(-q));
      const double* P = y + x->inc + i * nl;
      double* dP = dy + x->inc + i * nl;
      dP[0] = -q / eps / aH * P[1] + hp / 6.0 * dlf;
      dP[1] = q * k * k / 3.0 / eps / aH * (P[0] - 2.0 * P[2]);
      dP[2] = q * k / 5.0 / eps / aH * (2.0 * P[1] / k - 3.0 * P[3]) - (hp / 15.0 + 2.0 * ep / 5.0) * dlf;
      for (int l = 3; l < L; ++l)
        dP[l] = q * k / eps / aH / (2 * l + 1) * (l * P[l - 1] - (l + 1) * P[l + 1]);
      dP[L] = q * k / aH / eps * P[L - 1] - (L + 1) / aH / tau * P[L];
    }
  }
}

/* initial_conditions_one_k + species y_ini (abcmb/perturbations.py:161-199;
 * abcmb/species.py:561-581, :677-705, :882-917, :1222-1243, :1367-1388) */
static void initial_conditions(const Ctx* x, double lna0, double* y) {
  const AbcmbCosmo* c = x->c;
  double k = x->k;
  double ti = bg_tau(c, lna0);
  double om = c->om, Rn = c->R_nu, Rb = c->R_b;
  memset(y, 0, sizeof(double) * x->n);
  y[0] = (1. - k * k * ti * ti / 12. / (15. + 4. * Rn) *
                   (5. + 4. * Rn - (16. * Rn * Rn + 280. * Rn + 325) / 10. / (2. * Rn + 15.) * ti * om));
  y[1] = -(k * ti) * (k * ti) / 4. * (1. - om * ti / 5.);
  y[2] = -(k * ti) * (k * ti) / 4. * (1. - om * ti / 5.);
  double thb = -pow(k, 4.0) * pow(ti, 3.0) / 36. * (1. - 3. * (1. + 5. * Rb - Rn) / 20. / (1. - Rn) * om * ti);
  y[3] = thb;
  y[x->ig] = -(k * ti) * (k * ti) / 3. * (1. - om * ti / 5.);
  y[x->ig + 1] = thb;
  double dn = -(k * ti) * (k * ti) / 3. * (1. - om * ti / 5.);
  double tn = -k * pow(k * ti, 3.0) / 36. / (4. * Rn + 15.) *
              (4. * Rn + 11. + 12. - 3. * (8. * Rn * Rn + 50. * Rn + 275.) / 20. / (2. * Rn + 15.) * ti * om);
  double sn = (k * ti) * (k * ti) / (45. + 12. * Rn) * 2. * (1. + (4. * Rn - 5.) / 4. / (2. * Rn + 15.) * ti * om);
  y[x->iu] = dn;
  y[x->iu + 1] = tn;
  y[x->iu + 2] = sn;
  for (int i = 0; i < x->nq; ++i) {
    double q = x->q[i];
    double* P = y + x->inc + i * (x->ln + 1);
    P[0] = dn / 4. * q / (1. + exp(-q));
    P[1] = tn / 3. * q / (1. + exp(-q));
    P[2] = sn / 2. * q / (1. + exp(-q));
  }
  if (g_fl && g_fl->n_rows > 0) g_fl->ic(g_fl->ctx, c, k, ti, y + x->n - g_fl->n_rows);
}

/* ---------- dense linear algebra ---------- */
static int lu_factor(double* A, int n, int* piv) {
  for (int p = 0; p < n; ++p) {
    int best = p;
    double bv = fabs(A[p * n + p]);
    for (int r = p + 1; r < n; ++r)
      if (fabs(A[r * n + p]) > bv) { bv = fabs(A[r * n + p]); best = r; }
    piv[p] = best;
    if (bv == 0.0) return -1;
    if (best != p)
      for (int j = 0; j < n; ++j) { double t = A[p * n + j]; A[p * n + j] = A[best * n + j]; A[best * n + j] = t; }
    double inv = 1.0 / A[p * n + p];
    for (int r = p + 1; r < n; ++r) {
      double f = A[r * n + p] * inv;
      A[r * n + p] = f;
      for (int j = p + 1; j < n; ++j) A[r * n + j] -= f * A[p * n + j];
    }
  }
  return 0;
}
static void lu_solve(const double* A, int n, const int* piv, double* b) {
  for (int p = 0; p < n; ++p) {
    int q = piv[p];
    if (q != p) { double t = b[p]; b[p] = b[q]; b[q] = t; }
  }
  for (int i = 0; i < n; ++i) {
    double s = b[i];
    for (int j = 0; j < i; ++j) s -= A[i * n + j] * b[j];
    b[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * b[j];
    b[i] = s / A[i * n + i];
  }
}

/* chord updates below this (in units of the error tolerance) are rounding noise:
 * u / rtol is ~2e-11 for rtol = 1e-5, so a smaller threshold could stall */
#define NEWTON_SMALL 1e-9

/* ---------- Kvaerno5 tableau (Kvaerno 2004) ---------- */
static const double KV_G = 0.26;
static const double KV_C[7] = {0.0, 0.52, 1.230333209967908, 0.8957659843500759, 0.43639360985864756, 1.0, 1.0};
static const double KV_A[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {0.26, 0, 0, 0, 0, 0},
    {0.13, 0.84033320996790809, 0, 0, 0, 0},
    {0.22371961478320505, 0.47675532319799699, -0.06470895363112615, 0, 0, 0},
    {0.16648564323248321, 0.10450018841591720, 0.03631482272098715, -0.13090704451073998, 0, 0},
    {0.13855640231268224, 0.0, -0.04245337201752043, 0.02446657898003141, 0.61943039072480676, 0},
    {0.13659751177640291, 0.0, -0.05496908796538376, -0.04118626728321046, 0.62993304899016403,
     0.06962479448202728}};
/* b - b_hat: b = row 7 with gamma, b_hat = row 6 with gamma */
static const double KV_E[7] = {0.13659751177640291 - 0.13855640231268224, 0.0,
                               -0.05496908796538376 + 0.04245337201752043,
                               -0.04118626728321046 - 0.02446657898003141,
                               0.62993304899016403 - 0.61943039072480676,
                               0.06962479448202728 - 0.26, 0.26};

void oracle_kvaerno5_tableau(double* A49, double* b7, double* bh7, double* c7) {
  for (int i = 0; i < 7; ++i) {
    for (int j = 0; j < 7; ++j) A49[i * 7 + j] = j < i ? KV_A[i][j] : (j == i && i > 0 ? KV_G : 0.0);
    c7[i] = KV_C[i];
  }
  for (int j = 0; j < 7; ++j) {
    b7[j] = A49[6 * 7 + j];
    bh7[j] = b7[j] - KV_E[j];
  }
}

/* ---------- one k-mode (abcmb/perturbations.py:250-316) ---------- */
static double rms_scaled(const double* e, const double* y0, const double* y1, int n, double rtol, double atol) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    double m = fabs(y0[i]);
    if (y1 && fabs(y1[i]) > m) m = fabs(y1[i]);
    double v = e[i] / (atol + rtol * m);
    s += v * v;
  }
  return sqrt(s / n);
}

/* out_states: [n_save][n] raw saved states; returns status */
static double* g_trace = NULL;  /* optional per-step trace (single-threaded debug) */
static int g_trace_cap = 0;

/* Step script: with mode SCRIPT_RECORD every attempted step writes
 * {hh, accepted, chord iterations of stages 1..6, t + hh} (9 doubles); with
 * SCRIPT_REPLAY the integration takes exactly those steps and iteration
 * counts, skipping error control and convergence tests.  Differencing
 * replays at nearby parameters gives the derivative of the discrete map at
 * a frozen step sequence, which is what a tangent-linear integrator computes. */
enum { SCRIPT_NONE = 0, SCRIPT_RECORD = 1, SCRIPT_REPLAY = 2, SCRIPT_REPLAY_CONVERGED = 3 };
#define SCRIPT_W 9
typedef struct {
  int mode, cap, len;
  double* rec;        /* [cap][SCRIPT_W] */
  const Ctx* jac;     /* replay: build the chord matrix from this cosmology (NULL: the integrated one) */
} Script;

/* Tangent-linear companion of the integration (forward mode along nd
 * parameter directions).  The parameter derivative of the right-hand side at
 * a fixed state, dA/dtheta_d y, is the central difference of rhs() between
 * the cosmologies xp[d] and xm[d] at theta +- h[d] (rhs is linear in y and
 * smooth in theta, so this is exact to O(h^2)); everything else is exact
 * linear algebra on the recorded step sequence: each stage tangent solves
 * (I - g A(t_s)) zdot = basedot + g dA/dtheta z to rounding (iterated with
 * the step's chord matrix), with z the value's stage.  The saved tangent
 * states go to ys_dot[nd][n_save][n]. */
typedef struct {
  int nd;
  const Ctx* xp;
  const Ctx* xm;
  const double* h;
  double* ys_dot;
} Tangent;

/* out = dA/dtheta_d y at lna */
static void rhs_dtheta(const Tangent* tg, int d, double lna, const double* y, double* out, double* work, int n) {
  rhs(&tg->xp[d], lna, y, out);
  rhs(&tg->xm[d], lna, y, work);
  for (int i = 0; i < n; ++i) out[i] = (out[i] - work[i]) / (2.0 * tg->h[d]);
}

static int ode_one_k_s(const Ctx* x, double lna0, double rtol, double atol, const AbcmbOdeSettings* st,
                       int n_save, const double* lna_save, double* out_states, int* nsteps, Script* sc,
                       const Tangent* tg);
static int ode_one_k(const Ctx* x, double lna0, double rtol, double atol, const AbcmbOdeSettings* st,
                     int n_save, const double* lna_save, double* out_states, int* nsteps) {
  return ode_one_k_s(x, lna0, rtol, atol, st, n_save, lna_save, out_states, nsteps, NULL, NULL);
}

static int ode_one_k_s(const Ctx* x, double lna0, double rtol, double atol, const AbcmbOdeSettings* st,
                       int n_save, const double* lna_save, double* out_states, int* nsteps, Script* sc,
                       const Tangent* tg) {
  const int converged = sc && sc->mode == SCRIPT_REPLAY_CONVERGED;
  const int replay = sc && (sc->mode == SCRIPT_REPLAY || converged);
  const Ctx* xj = (replay && sc->jac) ? sc->jac : x;
  const int record = sc && sc->mode == SCRIPT_RECORD;
  int n = x->n;
  double* mem = (double*)malloc(sizeof(double) * (size_t)(n * n + 20 * n));
  int* piv = (int*)malloc(sizeof(int) * n);
  double* M = mem;
  double* y = M + n * n;
  double* f0 = y + n;
  double* F = f0 + n;            /* 7 * n */
  double* base = F + 7 * n;
  double* z = base + n;
  double* r = z + n;
  double* ynew = r + n;
  double* fnew = ynew + n;
  double* err = fnew + n;
  double* e = err + n;
  double* tmp = e + n;
  initial_conditions(x, lna0, y);
  double t = lna0, h = st->dt0;
  double b1 = st->pcoeff + st->icoeff + st->dcoeff, b2 = -(st->pcoeff + 2.0 * st->dcoeff);
  double prev_inv = 1.0;
  rhs(x, t, y, f0);
  /* tangent state: per direction y', f0', F'[7], stage base', z', f_new' (12 n) */
  const int nd = tg ? tg->nd : 0;
  double* tmem = nd ? (double*)malloc(sizeof(double) * ((size_t)nd * 12 * n + 2 * (size_t)n)) : NULL;
  double* tw1 = tmem ? tmem + (size_t)nd * 12 * n : NULL;
  double* tw2 = tw1 ? tw1 + n : NULL;
#define TYD(d) (tmem + (size_t)(d) * 12 * n)
#define TF0(d) (TYD(d) + n)
#define TFS(d, s_) (TYD(d) + (size_t)(2 + (s_)) * n)
#define TBASE(d) (TYD(d) + 9 * n)
#define TZ(d) (TYD(d) + 10 * n)
#define TFN(d) (TYD(d) + 11 * n)
  for (int d = 0; d < nd; ++d) {
    /* initial conditions and their derivative at the (fixed) start time */
    initial_conditions(&tg->xp[d], lna0, TZ(d));
    initial_conditions(&tg->xm[d], lna0, TBASE(d));
    for (int i = 0; i < n; ++i) TYD(d)[i] = (TZ(d)[i] - TBASE(d)[i]) / (2.0 * tg->h[d]);
    rhs(x, t, TYD(d), TF0(d));
    rhs_dtheta(tg, d, t, y, tw1, tw2, n);
    for (int i = 0; i < n; ++i) TF0(d)[i] += tw1[i];
  }
  int isave = 0, steps = 0, status = ABCMB_K_OK;
  while (isave < n_save && lna_save[isave] <= t) {  /* save points at t0 */
    memcpy(out_states + (size_t)isave * n, y, sizeof(double) * n);
    for (int d = 0; d < nd; ++d)
      memcpy(tg->ys_dot + ((size_t)d * n_save + isave) * n, TYD(d), sizeof(double) * n);
    ++isave;
  }
  while (t < 0.0) {
    if (steps >= st->max_steps) { status = ABCMB_K_MAX_STEPS; break; }
    const double* srec = NULL;
    if (replay) {
      if (steps >= sc->len) { status = ABCMB_K_MAX_STEPS; break; }
      srec = sc->rec + (size_t)steps * SCRIPT_W;
      if (srec[1] == 0.0) { ++steps; continue; }  /* a rejected step leaves the state alone */
    }
    double tn = t + h;
    if (tn > 0.0) tn = 0.0;
    double hh = tn - t;
    if (replay) { hh = srec[0]; tn = srec[8]; }
    if (!(hh > 0.0)) { status = ABCMB_K_STEP_UNDERFLOW; break; }
    ++steps;
    int its[7] = {0, 0, 0, 0, 0, 0, 0};
    /* M = I - hh*gamma*J(t) with J columns from the linear RHS */
    for (int j = 0; j < n; ++j) {
      memset(e, 0, sizeof(double) * n);
      e[j] = 1.0;
      rhs(xj, t + st->jac_offset * hh, e, tmp);
      for (int i = 0; i < n; ++i) M[i * n + j] = (i == j ? 1.0 : 0.0) - hh * KV_G * tmp[i];
    }
    int ok = lu_factor(M, n, piv) == 0;
    memcpy(F, f0, sizeof(double) * n);
    for (int s = 1; s < 7 && ok; ++s) {
      double ts = t + KV_C[s] * hh;
      for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int j = 0; j < s; ++j) acc += KV_A[s][j] * F[j * n + i];
        base[i] = y[i] + hh * acc;
        z[i] = base[i] + hh * KV_G * F[(s - 1) * n + i];
      }
      int conv = 0;
      double prev = 0.0;
      const int nit = converged ? 200 : replay ? (int)srec[1 + s] : st->max_newton;
      for (int it = 0; it < nit; ++it) {
        ++its[s];
        rhs(x, ts, z, tmp);
        for (int i = 0; i < n; ++i) r[i] = -(z[i] - base[i] - hh * KV_G * tmp[i]);
        lu_solve(M, n, piv, r);
        for (int i = 0; i < n; ++i) z[i] += r[i];
        if (converged) {  /* to rounding: the stage equations solved exactly */
          double nrm = rms_scaled(r, z, NULL, n, rtol, atol);
          conv = 1;
          /* past the recorded iterations: stop when done or at the rounding floor */
          if (it + 1 >= (int)srec[1 + s] && (nrm < 1e-10 || nrm >= prev)) break;
          prev = nrm;
          continue;
        }
        if (replay) { conv = 1; continue; }
        double nrm = rms_scaled(r, z, NULL, n, rtol, atol);
        if (!isfinite(nrm)) break;
        if (nrm < NEWTON_SMALL) { conv = 1; break; }
        if (it > 0) {
          double th = nrm / prev;
          if (th > 2.0) break;
          double fac = nrm * th / (1.0 - th);
          if (fac > 0.0 && fac < st->newton_kappa) { conv = 1; break; }
        }
        prev = nrm;
      }
      if (!conv) { ok = 0; break; }
      for (int i = 0; i < n; ++i) F[s * n + i] = (z[i] - base[i]) / (hh * KV_G);
      for (int d = 0; d < nd; ++d) {
        double* bd = TBASE(d);
        double* zd = TZ(d);
        if (s == 1) memcpy(TFS(d, 0), TF0(d), sizeof(double) * n);
        for (int i = 0; i < n; ++i) {
          double acc = 0.0;
          for (int j = 0; j < s; ++j) acc += KV_A[s][j] * TFS(d, j)[i];
          bd[i] = TYD(d)[i] + hh * acc;
          zd[i] = bd[i] + hh * KV_G * TFS(d, s - 1)[i];
        }
        double* adz = tw1;   /* g dA/dtheta z, fixed during the iteration */
        rhs_dtheta(tg, d, ts, z, adz, tw2, n);
        double prevt = INFINITY;
        for (int it = 0; it < 400; ++it) {
          rhs(x, ts, zd, tmp);
          for (int i = 0; i < n; ++i) r[i] = -(zd[i] - bd[i] - hh * KV_G * (tmp[i] + adz[i]));
          lu_solve(M, n, piv, r);
          double nr = 0.0, nz = 0.0;
          for (int i = 0; i < n; ++i) {
            zd[i] += r[i];
            double w = 1.0 / (atol + rtol * fabs(z[i]));
            nr += (r[i] * w) * (r[i] * w);
            nz += (zd[i] * w) * (zd[i] * w);
          }
          if (nr <= 1e-30 * nz || (it > 2 && nr >= prevt)) break;  /* to rounding */
          prevt = nr;
        }
        for (int i = 0; i < n; ++i) TFS(d, s)[i] = (zd[i] - bd[i]) / (hh * KV_G);
      }
    }
    double enorm = INFINITY;
    if (ok) {
      for (int i = 0; i < n; ++i) {
        ynew[i] = z[i];
        double acc = 0.0;
        for (int j = 0; j < 7; ++j) acc += KV_E[j] * F[j * n + i];
        err[i] = hh * acc;
      }
      enorm = rms_scaled(err, y, ynew, n, rtol, atol);
      if (!isfinite(enorm)) enorm = INFINITY;
    }
    int accept = replay ? 1 : enorm < 1.0;
    if (record && steps - 1 < sc->cap) {
      double* w = sc->rec + (size_t)(steps - 1) * SCRIPT_W;
      w[0] = hh; w[1] = accept;
      for (int q = 1; q < 7; ++q) w[1 + q] = its[q];
      w[8] = tn;
      sc->len = steps;
    }
    if (g_trace && steps - 1 < g_trace_cap) {
      double* tr = g_trace + (size_t)(steps - 1) * 4;
      tr[0] = t; tr[1] = hh; tr[2] = enorm; tr[3] = ok;
    }
    double inv = enorm == 0.0 ? INFINITY : 1.0 / enorm;
    double fmin = accept ? 1.0 : st->factormin;
    double fac = st->safety * pow(inv, b1 / 5.0) * pow(prev_inv, b2 / 5.0);
    if (!(fac >= fmin)) fac = fmin;   /* also catches NaN */
    if (fac > st->factormax) fac = st->factormax;
    if (accept) {
      rhs(x, tn, ynew, fnew);
      for (int d = 0; d < nd; ++d) {  /* z' of the last stage is the new tangent state */
        rhs(x, tn, TZ(d), TFN(d));
        rhs_dtheta(tg, d, tn, ynew, tw1, tw2, n);
        for (int i = 0; i < n; ++i) TFN(d)[i] += tw1[i];
      }
      while (isave < n_save && lna_save[isave] <= tn) {
        double th = (lna_save[isave] - t) / hh;
        double h00 = 2 * th * th * th - 3 * th * th + 1, h10 = th * th * th - 2 * th * th + th;
        double h01 = -2 * th * th * th + 3 * th * th, h11 = th * th * th - th * th;
        double* o = out_states + (size_t)isave * n;
        for (int i = 0; i < n; ++i) o[i] = h00 * y[i] + h10 * hh * f0[i] + h01 * ynew[i] + h11 * hh * fnew[i];
        for (int d = 0; d < nd; ++d) {
          double* od = tg->ys_dot + ((size_t)d * n_save + isave) * n;
          for (int i = 0; i < n; ++i)
            od[i] = h00 * TYD(d)[i] + h10 * hh * TF0(d)[i] + h01 * TZ(d)[i] + h11 * hh * TFN(d)[i];
        }
        ++isave;
      }
      t = tn;
      memcpy(y, ynew, sizeof(double) * n);
      memcpy(f0, fnew, sizeof(double) * n);
      for (int d = 0; d < nd; ++d) {
        memcpy(TYD(d), TZ(d), sizeof(double) * n);
        memcpy(TF0(d), TFN(d), sizeof(double) * n);
      }
      prev_inv = inv;
      int bad = 0;
      for (int i = 0; i < n; ++i) if (!isfinite(y[i])) bad = 1;
      if (bad) { status = ABCMB_K_NONFINITE; break; }
    }
    h = hh * fac;
  }
  *nsteps = steps;
  free(mem);
  free(piv);
  free(tmem);
#undef TYD
#undef TF0
#undef TFS
#undef TBASE
#undef TZ
#undef TFN
  return status;
}

/* make_output_table + output_perturbations at one (lna, k)
 * (abcmb/perturbations.py:318-404; abcmb/species.py:609-610, 755-760, 1077-1090, 1287-1291, 1450-1456) */
static void output_fields(const Ctx* x, double lna, const double* y, double* o /* [21] */) {
  const AbcmbCosmo* c = x->c;
  double k = x->k;
  Rho r = rho_all(c, lna);
  double a = r.a;
  double aH = bg_aH(c, lna);
  double cs2 = bg_cs2(c, lna);
  double R = 4. * r.g / 3. / r.b;
  double tau_c = bg_tau_c(c, lna);
  double nrd = 0.0, nrt = 0.0, nrs = 0.0;
  int has = c->has_ncdm && x->nq > 0;
  if (has) ncdm_sums(x, &r, y, &nrd, &nrt, &nrs);
  double db = y[2], tb = y[3], tg = y[x->ig + 1];
  double tbp = -tb + cs2 / aH * (k * k * db) + R / aH / tau_c * (tg - tb);
  double srd = 0.0, srt = 0.0, srs = 0.0, srdm = 0.0, srm = 0.0;
  /* CDM */
  srd += r.cdm * y[1]; srt += (r.cdm + 0.0) * 0.0; srs += (r.cdm + 0.0) * 0.0;
  srdm += r.cdm * y[1]; srm += r.cdm;
  /* Baryon */
  srd += r.b * y[2]; srt += (r.b + 0.0) * y[3]; srs += (r.b + 0.0) * 0.0;
  srdm += r.b * y[2]; srm += r.b;
  /* Photon */
  srd += r.g * y[x->ig]; srt += (r.g + r.g / 3.) * y[x->ig + 1]; srs += (r.g + r.g / 3.) * y[x->ig + 2];
  /* Massless */
  srd += r.ur * y[x->iu]; srt += (r.ur + r.ur / 3.) * y[x->iu + 1]; srs += (r.ur + r.ur / 3.) * y[x->iu + 2];
  /* Massive */
  if (x->nq > 0) {
    srd += nrd; srt += nrt; srs += nrs;
    srdm += nrd; srm += r.ncdm;
  }
  /* user fluids */
  if (g_fl) {
    double u[4] = {0.0, 0.0, 0.0, 0.0};
    if (g_fl->n_rows > 0) g_fl->sums(g_fl->ctx, c, lna, y, x->n, u);
    srd += u[0]; srt += u[1]; srs += u[2];
    srdm += u[3]; srm += r.m_fl;
  }
  double dm = srdm / srm;
  double hp = 2. / (aH * aH) * (k * k * y[0] + 4. * PI * GN * a * a / (C_MPC * C_MPC) * srd);
  double ep = 4. * PI * GN * a * a / aH * srt / (C_MPC * C_MPC) / (k * k);
  double al = aH * (hp + 6. * ep) / 2. / (k * k);
  double alp = y[0] / aH - 2. * al - 12. * PI * GN * a * a / aH * srs / (C_MPC * C_MPC) / (k * k);
  o[ABCMB_PT_DELTA_M] = dm;
  o[ABCMB_PT_THETA_B_PRIME] = tbp;
  o[ABCMB_PT_ETA] = y[0];
  o[ABCMB_PT_H_PRIME] = hp;
  o[ABCMB_PT_ETA_PRIME] = ep;
  o[ABCMB_PT_ALPHA] = al;
  o[ABCMB_PT_ALPHA_PRIME] = alp;
  o[ABCMB_PT_CDM_DELTA] = y[1];
  o[ABCMB_PT_B_DELTA] = y[2];
  o[ABCMB_PT_B_THETA] = y[3];
  o[ABCMB_PT_G_DELTA] = y[x->ig];
  o[ABCMB_PT_G_THETA] = y[x->ig + 1];
  o[ABCMB_PT_G_SIGMA] = y[x->ig + 2];
  o[ABCMB_PT_G_G0] = y[x->ip];
  o[ABCMB_PT_G_G2] = y[x->ip + 2];
  o[ABCMB_PT_UR_DELTA] = y[x->iu];
  o[ABCMB_PT_UR_THETA] = y[x->iu + 1];
  o[ABCMB_PT_UR_SIGMA] = y[x->iu + 2];
  if (x->nq > 0) {
    o[ABCMB_PT_NCDM_DELTA] = nrd / r.ncdm;
    o[ABCMB_PT_NCDM_THETA] = nrt / (r.ncdm + r.p_ncdm);
    o[ABCMB_PT_NCDM_SIGMA] = nrs / (r.ncdm + r.p_ncdm);
  } else {
    o[ABCMB_PT_NCDM_DELTA] = o[ABCMB_PT_NCDM_THETA] = o[ABCMB_PT_NCDM_SIGMA] = 0.0;
  }
}

/* ---------- exported: perturbations ---------- */

/* Runs every item; pt output goes to cosmos[item.cosmo].pt ([21][n_save][nk]).
 * states (optional): [n_items][n_save][n]. */
int oracle_pert(const AbcmbCosmo* cosmos, const AbcmbPertItem* items, int n_items,
                const AbcmbOdeSettings* st, int32_t* status, int32_t* steps, double* states) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int it = 0; it < n_items; ++it) {
    const AbcmbPertItem* I = &items[it];
    const AbcmbCosmo* c = &cosmos[I->cosmo];
    Ctx x;
    ctx_init(&x, c, st, I->k);
    int ns = c->n_save;
    double* S = states ? states + (size_t)it * ns * st->n : (double*)malloc(sizeof(double) * (size_t)ns * st->n);
    int nst = 0;
    status[it] = ode_one_k(&x, I->lna_start, I->rtol, I->atol, st, ns, c->lna_save, S, &nst);
    steps[it] = nst;
    double o[ABCMB_N_PT_FIELDS];
    for (int s = 0; s < ns; ++s) {
      output_fields(&x, c->lna_save[s], S + (size_t)s * st->n, o);
      for (int f = 0; f < ABCMB_N_PT_FIELDS; ++f) c->pt[((size_t)f * ns + s) * c->nk + I->ik] = o[f];
      /* the user-fluid rows as they are */
      for (int j = 0; j < st->n_fluid && c->pt_fluid; ++j)
        c->pt_fluid[((size_t)j * ns + s) * c->nk + I->ik] = S[(size_t)s * st->n + st->n - st->n_fluid + j];
    }
    if (!states) free(S);
  }
  return 0;
}

/* oracle_pert with a step script per item: mode 1 records into
 * scripts[n_items][cap][9] (and lens[n_items]), mode 2 replays them with the
 * recorded chord-iteration counts, mode 3 replays the steps with every stage
 * solved to rounding.  jac_cosmos (may be NULL): replays build their chord
 * matrices from these cosmologies instead of the integrated ones. */
int oracle_pert_scripted(const AbcmbCosmo* cosmos, const AbcmbCosmo* jac_cosmos, const AbcmbPertItem* items,
                         int n_items, const AbcmbOdeSettings* st, int mode, double* scripts, int32_t* lens, int cap,
                         int32_t* status, int32_t* steps) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int it = 0; it < n_items; ++it) {
    const AbcmbPertItem* I = &items[it];
    const AbcmbCosmo* c = &cosmos[I->cosmo];
    Ctx x;
    ctx_init(&x, c, st, I->k);
    int ns = c->n_save;
    double* S = (double*)malloc(sizeof(double) * (size_t)ns * st->n);
    Ctx xj;
    if (jac_cosmos) ctx_init(&xj, &jac_cosmos[I->cosmo], st, I->k);
    Script sc;
    sc.jac = jac_cosmos ? &xj : NULL;
    sc.mode = mode;
    sc.cap = cap;
    sc.len = mode >= SCRIPT_REPLAY ? lens[it] : 0;
    sc.rec = scripts + (size_t)it * cap * SCRIPT_W;
    int nst = 0;
    status[it] = ode_one_k_s(&x, I->lna_start, I->rtol, I->atol, st, ns, c->lna_save, S, &nst, &sc, NULL);
    if (mode == SCRIPT_RECORD) {
      lens[it] = sc.len;
      if (nst > cap) status[it] = ABCMB_K_MAX_STEPS;
    }
    steps[it] = nst;
    double o[ABCMB_N_PT_FIELDS];
    for (int s = 0; s < ns; ++s) {
      output_fields(&x, c->lna_save[s], S + (size_t)s * st->n, o);
      for (int f = 0; f < ABCMB_N_PT_FIELDS; ++f) c->pt[((size_t)f * ns + s) * c->nk + I->ik] = o[f];
    }
    free(S);
  }
  return 0;
}

/* Tangent-linear perturbation tables of one cosmology: base->pt gets the
 * values, dpt[nd][21][n_save][nk] their derivatives along nd directions whose
 * cosmologies at theta +- h[d] are plus[d] / minus[d] (same grids as base).
 * The output fields are linear in the state, so their derivative is the
 * field of the tangent state plus the central difference of the field at the
 * fixed state. */
int oracle_pert_tangent(const AbcmbCosmo* base, const AbcmbCosmo* plus, const AbcmbCosmo* minus,
                        const double* h, int nd, const AbcmbPertItem* items, int n_items,
                        const AbcmbOdeSettings* st, double* dpt, int32_t* status, int32_t* steps) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int it = 0; it < n_items; ++it) {
    const AbcmbPertItem* I = &items[it];
    Ctx x, xp[8], xm[8];
    ctx_init(&x, base, st, I->k);
    for (int d = 0; d < nd && d < 8; ++d) { ctx_init(&xp[d], &plus[d], st, I->k); ctx_init(&xm[d], &minus[d], st, I->k); }
    int ns = base->n_save, n = st->n;
    double* S = (double*)malloc(sizeof(double) * (size_t)ns * n * (1 + nd));
    Tangent tg = {nd, xp, xm, h, S + (size_t)ns * n};
    int nst = 0;
    status[it] = nd <= 8 ? ode_one_k_s(&x, I->lna_start, I->rtol, I->atol, st, ns, base->lna_save, S, &nst, NULL, &tg)
                         : ABCMB_K_NONFINITE;
    steps[it] = nst;
    double o[ABCMB_N_PT_FIELDS], op[ABCMB_N_PT_FIELDS], om[ABCMB_N_PT_FIELDS];
    const size_t plane = (size_t)ABCMB_N_PT_FIELDS * ns * base->nk;
    for (int s = 0; s < ns; ++s) {
      const double* ys = S + (size_t)s * n;
      output_fields(&x, base->lna_save[s], ys, o);
      for (int f = 0; f < ABCMB_N_PT_FIELDS; ++f) base->pt[((size_t)f * ns + s) * base->nk + I->ik] = o[f];
      for (int d = 0; d < nd && d < 8; ++d) {
        output_fields(&x, base->lna_save[s], tg.ys_dot + ((size_t)d * ns + s) * n, o);
        output_fields(&xp[d], base->lna_save[s], ys, op);
        output_fields(&xm[d], base->lna_save[s], ys, om);
        for (int f = 0; f < ABCMB_N_PT_FIELDS; ++f)
          dpt[d * plane + ((size_t)f * ns + s) * base->nk + I->ik] = o[f] + (op[f] - om[f]) / (2.0 * h[d]);
      }
    }
    free(S);
  }
  return 0;
}

int oracle_trace(const AbcmbCosmo* c, const AbcmbOdeSettings* st, const AbcmbPertItem* I, double* trace, int cap) {
  Ctx x;
  ctx_init(&x, c, st, I->k);
  double* S = (double*)malloc(sizeof(double) * (size_t)c->n_save * st->n);
  g_trace = trace; g_trace_cap = cap;
  int nst = 0;
  ode_one_k(&x, I->lna_start, I->rtol, I->atol, st, c->n_save, c->lna_save, S, &nst);
  g_trace = NULL;
  free(S);
  return nst;
}

int oracle_rhs(const AbcmbCosmo* c, const AbcmbOdeSettings* st, double k, double lna, const double* y, double* dy) {
  Ctx x;
  ctx_init(&x, c, st, k);
  rhs(&x, lna, y, dy);
  return 0;
}
int oracle_ic(const AbcmbCosmo* c, const AbcmbOdeSettings* st, double k, double lna, double* y) {
  Ctx x;
  ctx_init(&x, c, st, k);
  initial_conditions(&x, lna, y);
  return 0;
}
int oracle_bg(const AbcmbCosmo* c, double lna, double* out /* aH, tau, tau_c, cs2, xe, Tm */) {
  out[0] = bg_aH(c, lna); out[1] = bg_tau(c, lna); out[2] = bg_tau_c(c, lna);
  out[3] = bg_cs2(c, lna); out[4] = bg_xe(c, lna); out[5] = bg_Tm(c, lna);
  return 0;
}

/* ---------- cubic spline, not-a-knot (interpax CubicSpline default, abcmb/spectrum.py:664) ---------- */
/* second derivatives M of the not-a-knot spline through (x, y), n >= 4 */
static void spline_m(const double* x, const double* y, int n, double* M, double* work) {
  double* hh = work;            /* n-1 */
  double* a = hh + n;           /* sub */
  double* b = a + n;            /* diag */
  double* cc = b + n;           /* super */
  double* d = cc + n;
  for (int i = 0; i < n - 1; ++i) hh[i] = x[i + 1] - x[i];
  int m = n - 2;                /* unknowns M_1..M_{n-2} */
  for (int i = 1; i <= n - 2; ++i) {
    int r = i - 1;
    a[r] = hh[i - 1];
    b[r] = 2.0 * (hh[i - 1] + hh[i]);
    cc[r] = hh[i];
    d[r] = 6.0 * ((y[i + 1] - y[i]) / hh[i] - (y[i] - y[i - 1]) / hh[i - 1]);
  }
  /* not-a-knot at x_1: M_0 = M_1 (1 + h0/h1) - (h0/h1) M_2 */
  b[0] = hh[0] * (1.0 + hh[0] / hh[1]) + 2.0 * (hh[0] + hh[1]);
  cc[0] = hh[1] - hh[0] * hh[0] / hh[1];
  /* at x_{n-2}: M_{n-1} = M_{n-2} (1 + h_{n-2}/h_{n-3}) - (h_{n-2}/h_{n-3}) M_{n-3} */
  {
    double h2 = hh[n - 2], h3 = hh[n - 3];
    a[m - 1] = h3 - h2 * h2 / h3;
    b[m - 1] = 2.0 * (h3 + h2) + h2 * (1.0 + h2 / h3);
  }
  if (m == 1) { /* n == 3 not supported by not-a-knot reduction */ }
  /* Thomas */
  for (int r = 1; r < m; ++r) {
    double w = a[r] / b[r - 1];
    b[r] -= w * cc[r - 1];
    d[r] -= w * d[r - 1];
  }
  M[m] = d[m - 1] / b[m - 1];
  for (int r = m - 2; r >= 0; --r) M[r + 1] = (d[r] - cc[r] * M[r + 2]) / b[r];
  M[0] = M[1] * (1.0 + hh[0] / hh[1]) - hh[0] / hh[1] * M[2];
  M[n - 1] = M[n - 2] * (1.0 + hh[n - 2] / hh[n - 3]) - hh[n - 2] / hh[n - 3] * M[n - 3];
}

static double spline_eval(const double* x, const double* y, const double* M, int n, double xq) {
  int j;
  if (xq <= x[0]) j = 0;
  else if (xq >= x[n - 1]) j = n - 2;
  else {
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) { int mid = (lo + hi) / 2; if (x[mid] <= xq) lo = mid; else hi = mid; }
    j = lo;
  }
  double h = x[j + 1] - x[j];
  double A = (x[j + 1] - xq) / h, B = (xq - x[j]) / h;
  return A * y[j] + B * y[j + 1] + ((A * A * A - A) * M[j] + (B * B * B - B) * M[j + 1]) * h * h / 6.0;
}

int oracle_spline(const double* x, const double* y, int n, const double* xq, int nq, double* out) {
  double* M = (double*)malloc(sizeof(double) * n);
  double* w = (double*)malloc(sizeof(double) * 5 * n);
  spline_m(x, y, n, M, w);
  for (int i = 0; i < nq; ++i) out[i] = spline_eval(x, y, M, n, xq[i]);
  free(M);
  free(w);
  return 0;
}

/* ---------- sources (abcmb/spectrum.py:640-696, weights :755-758) ---------- */
/* pt: [21][n_save][nk]; bgv: [6][n_save-1]; out: [4][n_save-1][nkt] = w_i/aH_i * S_X */
int oracle_sources(const double* pt, const double* k, int nk, int n_save, const double* bgv,
                   const double* lna_save, const double* kt, int nkt, const double* scales, double* out) {
  static const int fields[10] = {ABCMB_PT_G_DELTA, ABCMB_PT_B_THETA, ABCMB_PT_THETA_B_PRIME, ABCMB_PT_G_SIGMA,
                                 ABCMB_PT_G_G0, ABCMB_PT_G_G2, ABCMB_PT_ETA, ABCMB_PT_ETA_PRIME,
                                 ABCMB_PT_ALPHA, ABCMB_PT_ALPHA_PRIME};
  int nt = n_save - 1;
  double* lk = (double*)malloc(sizeof(double) * nk);
  double* lkt = (double*)malloc(sizeof(double) * nkt);
  for (int i = 0; i < nk; ++i) lk[i] = log10(k[i]);
  for (int i = 0; i < nkt; ++i) lkt[i] = log10(kt[i]);
  double dl = lna_save[n_save - 1] - lna_save[n_save - 2];
#pragma omp parallel
  {
    double* row = (double*)malloc(sizeof(double) * nk);
    double* M = (double*)malloc(sizeof(double) * nk);
    double* w = (double*)malloc(sizeof(double) * 5 * nk);
    double* sp = (double*)malloc(sizeof(double) * 10 * nkt);
#pragma omp for schedule(dynamic, 4)
    for (int s = 0; s < nt; ++s) {
      for (int f = 0; f < 10; ++f) {
        for (int i = 0; i < nk; ++i) row[i] = pt[((size_t)fields[f] * n_save + s) * nk + i];
        spline_m(lk, row, nk, M, w);
        for (int j = 0; j < nkt; ++j) sp[f * nkt + j] = spline_eval(lk, row, M, nk, lkt[j]);
      }
      double g = bgv[1 * nt + s], gp = bgv[2 * nt + s], aH = bgv[3 * nt + s], em = bgv[4 * nt + s],
             aHd = bgv[5 * nt + s];
      double wq = (s == 0 ? 0.5 * dl : dl) / aH;
      for (int j = 0; j < nkt; ++j) {
        double dg = sp[0 * nkt + j], tb = sp[1 * nkt + j], tbp = sp[2 * nkt + j], sg = sp[3 * nkt + j];
        double g0 = sp[4 * nkt + j], g2 = sp[5 * nkt + j], eta = sp[6 * nkt + j], etap = sp[7 * nkt + j];
        double al = sp[8 * nkt + j], alp = sp[9 * nkt + j];
        double kk = kt[j];
        double sT0 = scales[0] * g * (dg / 4. + aH * alp) +
                     scales[1] * (g * (eta - aH * alp - 2. * aH * al) +
                                  2. * em * (aH * etap - aHd * al - aH * aH * alp)) +
                     scales[2] * (aH * (g * ((tbp / (kk * kk)) + alp) + gp * ((tb / (kk * kk)) + al)));
        double sT1 = scales[1] * em * ((aH * alp + 2. * aH * al - eta) * kk);
        double sT2 = scales[3] * g * (2 * sg + g0 + g2) / 8.;
        double sE = sqrt(6.0) * g * (2 * sg + g0 + g2) / 8.;
        size_t o = (size_t)s * nkt + j;
        out[0 * (size_t)nt * nkt + o] = wq * sT0;
        out[1 * (size_t)nt * nkt + o] = wq * sT1;
        out[2 * (size_t)nt * nkt + o] = wq * sT2;
        out[3 * (size_t)nt * nkt + o] = wq * sE;
      }
    }
    free(row); free(M); free(w); free(sp);
  }
  free(lk);
  free(lkt);
  return 0;
}

/* ---------- Bessel lookups (abcmb/spectrum.py:48-109) ---------- */
static double j_asym(double l, double x) {
  double nu = l + 0.5;
  double Q = sqrt(x * x - nu * nu) - nu * PI / 2 + nu * asin(nu / x);
  double J = sqrt(2 / PI / sqrt(x * x - nu * nu)) * cos(Q - PI / 4);
  return sqrt(PI / 2 / x) * J;
}
/* xb picks the branch (zero below the table, the table, the asymptotic form
 * above it) and x is where the branch is evaluated; the reference uses xb = x.
 * A branch frozen at other points (xb from a base cosmology) takes the jumps
 * between table and asymptotic form out of finite differences, which then
 * converge to the pointwise derivative (what autodiff of the reference gives). */
static double bessel_phi_b(int which, double l, const double* tab, double xmin, double xmax, int nx, double x,
                           double xb) {
  if (xb < xmin) return 0.0;
  if (xb >= xmax) {
    if (which == 0) return j_asym(l, x);
    if (which == 1) return l / x * j_asym(l, x) - j_asym(l + 1, x);
    return ((3 * l * (l - 1) - 2 * x * x) * j_asym(l, x) + 6 * x * j_asym(l + 1, x)) / 2 / (x * x);
  }
  return fast_interp(x, xmin, xmax, tab, nx);
}
static double bessel_phi(int which, double l, const double* tab, double xmin, double xmax, int nx, double x) {
  return bessel_phi_b(which, l, tab, xmin, xmax, nx, x, x);
}

/* LOS + C_l (abcmb/spectrum.py:697-797).  src: [4][n_tau][nkt]; bes: [3][n_lt][nx];
 * rng: [3][n_lt][2]; cl out: [3][n_ell]; delta out (optional) [4][n_ell][nkt];
 * chi_branch (optional, [n_tau]): distances that pick the Bessel branches
 * (bessel_phi_b), NULL for the reference's own choice */
int oracle_los_cl_branch(const double* src, int n_tau, const double* tau_los, double tau0, const double* kt,
                         int nkt, const double* bes, const double* rng, const int32_t* bes_l, int n_lt, int nx,
                         const int32_t* ell_idx, int n_ell, double A_s, double n_s, double k_pivot, double* cl,
                         double* delta, const double* chi_branch) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int e = 0; e < n_ell; ++e) {
    int li = ell_idx[e];
    double l = bes_l[li];
    double eps_f = sqrt(3. / 8. * (l + 2) * (l + 1) * l * (l - 1));
    double* D = (double*)calloc(4 * (size_t)nkt, sizeof(double));
    for (int s = 0; s < n_tau; ++s) {
      double chi0 = tau0 - tau_los[s];
      double chib = chi_branch ? chi_branch[s] : chi0;
      for (int j = 0; j < nkt; ++j) {
        double x = chi0 * kt[j], xb = chib * kt[j];
        double p0 = bessel_phi_b(0, l, bes + ((size_t)0 * n_lt + li) * nx, rng[(0 * n_lt + li) * 2], rng[(0 * n_lt + li) * 2 + 1], nx, x, xb);
        double p1 = bessel_phi_b(1, l, bes + ((size_t)1 * n_lt + li) * nx, rng[(1 * n_lt + li) * 2], rng[(1 * n_lt + li) * 2 + 1], nx, x, xb);
        double p2 = bessel_phi_b(2, l, bes + ((size_t)2 * n_lt + li) * nx, rng[(2 * n_lt + li) * 2], rng[(2 * n_lt + li) * 2 + 1], nx, x, xb);
        double ep = p0 / (x * x) * eps_f;
        size_t o = (size_t)s * nkt + j;
        D[0 * nkt + j] += src[0 * (size_t)n_tau * nkt + o] * p0;
        D[1 * nkt + j] += src[1 * (size_t)n_tau * nkt + o] * p1;
        D[2 * nkt + j] += src[2 * (size_t)n_tau * nkt + o] * p2;
        D[3 * nkt + j] += src[3 * (size_t)n_tau * nkt + o] * ep;
      }
    }
    double tt = 0, te = 0, ee = 0;
    for (int j = 0; j + 1 < nkt; ++j) {
      double fa[3], fb[3];
      for (int q = 0; q < 2; ++q) {
        int jj = j + q;
        double kk = kt[jj];
        double T = D[0 * nkt + jj] + D[1 * nkt + jj] + D[2 * nkt + jj];
        double E = D[3 * nkt + jj];
        double P = 4. * PI * A_s * pow(kk / k_pivot, n_s - 1.) / kk;
        double* f = q == 0 ? fa : fb;
        f[0] = P * T * T; f[1] = P * T * E; f[2] = P * E * E;
      }
      double dk = kt[j + 1] - kt[j];
      tt += 0.5 * (fa[0] + fb[0]) * dk;
      te += 0.5 * (fa[1] + fb[1]) * dk;
      ee += 0.5 * (fa[2] + fb[2]) * dk;
    }
    cl[0 * n_ell + e] = tt;
    cl[1 * n_ell + e] = te;
    cl[2 * n_ell + e] = ee;
    if (delta)
      for (int q = 0; q < 4; ++q)
        for (int j = 0; j < nkt; ++j) delta[((size_t)q * n_ell + e) * nkt + j] = D[q * nkt + j];
    free(D);
  }
  return 0;
}

int oracle_los_cl(const double* src, int n_tau, const double* tau_los, double tau0, const double* kt, int nkt,
                  const double* bes, const double* rng, const int32_t* bes_l, int n_lt, int nx,
                  const int32_t* ell_idx, int n_ell, double A_s, double n_s, double k_pivot, double* cl,
                  double* delta) {
  return oracle_los_cl_branch(src, n_tau, tau_los, tau0, kt, nkt, bes, rng, bes_l, n_lt, nx, ell_idx, n_ell, A_s,
                              n_s, k_pivot, cl, delta, NULL);
}

/* ---------- P(k) (abcmb/spectrum.py:250-302) ---------- */
static double interp_clamped(double x, const double* xp, const double* fp, int n) {
  if (x <= xp[0]) return fp[0];
  if (x >= xp[n - 1]) return fp[n - 1];
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) { int mid = (lo + hi) / 2; if (xp[mid] <= x) lo = mid; else hi = mid; }
  double t = (x - xp[lo]) / (xp[lo + 1] - xp[lo]);
  return fp[lo] + t * (fp[lo + 1] - fp[lo]);
}

static double primordial(double k, double A_s, double n_s, double kp) {
  return A_s * pow(k / kp, n_s - 1.) * (2 * PI * PI / (k * k * k));
}

/* delta_m at lna (linear in lna over the save grid, then in k), times P_R */
static double pk_at(const double* dm /* [n_save][nk] */, const double* lna_save, int n_save, const double* k, int nk,
                    double lna, double kq, double A_s, double n_s, double kp, double* col) {
  for (int i = 0; i < nk; ++i) {
    /* jnp.interp over PT.lna for column i */
    double v;
    if (lna <= lna_save[0]) v = dm[i];
    else if (lna >= lna_save[n_save - 1]) v = dm[(size_t)(n_save - 1) * nk + i];
    else {
      int lo = 0, hi = n_save - 1;
      while (hi - lo > 1) { int mid = (lo + hi) / 2; if (lna_save[mid] <= lna) lo = mid; else hi = mid; }
      double t = (lna - lna_save[lo]) / (lna_save[lo + 1] - lna_save[lo]);
      v = dm[(size_t)lo * nk + i] + t * (dm[(size_t)(lo + 1) * nk + i] - dm[(size_t)lo * nk + i]);
    }
    col[i] = v;
  }
  double d = interp_clamped(kq, k, col, nk);
  return d * d * primordial(kq, A_s, n_s, kp);
}

/* P(k, z) of the total matter (kind 0, SpectrumSolver.Pk_lin) or of CDM + baryons
 * (kind 1, SpectrumSolver.Pk_cb, abcmb/spectrum.py:304-347: (omega_b delta_b +
 * omega_cdm delta_cdm) / omega_m) at ln a = lna */
int oracle_pk_z(const double* pt, const double* k, int nk, int n_save, const double* lna_save, const double* kq,
                int nq, double A_s, double n_s, double kp, double lna, int kind, double omega_b, double omega_cdm,
                double omega_m, double* out) {
  const size_t fs = (size_t)n_save * nk;
  double* col = (double*)malloc(sizeof(double) * nk);
  double* dcb = NULL;
  const double* f = pt + (size_t)ABCMB_PT_DELTA_M * fs;
  if (kind == 1) {
    dcb = (double*)malloc(sizeof(double) * fs);
    for (size_t i = 0; i < fs; ++i)
      dcb[i] = (omega_b * pt[(size_t)ABCMB_PT_B_DELTA * fs + i] + omega_cdm * pt[(size_t)ABCMB_PT_CDM_DELTA * fs + i]) /
               omega_m;
    f = dcb;
  }
  for (int i = 0; i < nq; ++i) out[i] = pk_at(f, lna_save, n_save, k, nk, lna, kq[i], A_s, n_s, kp, col);
  free(col);
  free(dcb);
  return 0;
}

int oracle_pk(const double* pt, const double* k, int nk, int n_save, const double* lna_save, const double* kq,
              int nq, double A_s, double n_s, double kp, double* out) {
  double* col = (double*)malloc(sizeof(double) * nk);
  for (int i = 0; i < nq; ++i)
    out[i] = pk_at(pt + (size_t)ABCMB_PT_DELTA_M * n_save * nk, lna_save, n_save, k, nk, 0.0, kq[i], A_s, n_s, kp, col);
  free(col);
  return 0;
}

/* ---------- Limber C_l^phiphi (abcmb/spectrum.py:349-435) ---------- */
/* lna_ax: 500 points linspace(lna_rec, 0); aH_ax, tau_ax at those points; ells: n_l */
int oracle_clpp(const AbcmbCosmo* c, const double* pt, const double* lna_ax, const double* aH_ax,
                const double* tau_ax, int n_lna, double chi_rec, const double* ells, int n_l, double* out) {
  int ns = c->n_save, nk = c->nk;
  const double* dm = pt + (size_t)ABCMB_PT_DELTA_M * ns * nk;
  double Om = c->omega_m / (c->h * c->h), OL = c->omega_Lambda / (c->h * c->h);
  double lna_floor = lna_ax[n_lna - 2];
  /* background at the floor point (substituted at lna = 0) */
  int ifl = n_lna - 2;
#pragma omp parallel
  {
    double* col = (double*)malloc(sizeof(double) * nk);
    double* integ = (double*)malloc(sizeof(double) * n_lna);
#pragma omp for schedule(dynamic, 8)
    for (int il = 0; il < n_l; ++il) {
      double ell = ells[il];
      for (int i = 0; i < n_lna; ++i) {
        double lna = lna_ax[i];
        if (!(lna < 0.0)) { integ[i] = 0.0; continue; }
        int ii = lna < 0.0 ? i : ifl;
        double ls = lna < 0.0 ? lna : lna_floor;
        double chi = c->tau0 - tau_ax[ii];
        double kk = (ell + 0.5) / chi;
        double win = (chi_rec - chi) / chi_rec / chi;
        double a = exp(ls);
        double z = 1. / a - 1.;
        double aH = aH_ax[ii];
        double omz = (Om * pow(1. + z, 3.0)) / ((Om * pow(1. + z, 3.0)) + OL);
        double pk = pk_at(dm, c->lna_save, ns, c->k, nk, -log(1. + z), kk, c->A_s, c->n_s, c->k_pivot, col);
        double pl = 9. / 8. / (PI * PI) * omz * omz * pow(aH, 4.0) * pk / kk;
        integ[i] = chi / aH * win * win * pl;
      }
      double s = 0.0;
      for (int i = 0; i + 1 < n_lna; ++i) s += 0.5 * (integ[i] + integ[i + 1]) * (lna_ax[i + 1] - lna_ax[i]);
      out[il] = 8. * PI * PI / pow(ell + 0.5, 3.0) * s;
    }
    free(col);
    free(integ);
  }
  return 0;
}

/* ---------- Wigner d (abcmb/ABCMBTools.py:16-185) ---------- */
static double lfact(int n) { return lgamma(n + 1.0); }

/* d^l_{mn}(mu) for l = l0 .. l0+nl-1 (l0 >= m); m >= |n| */
static void wigner_d_row(int m, int n, double mu, int nl, double* out) {
  double beta = acos(mu);
  double norm = sqrt((2 * m + 1) / 2.0) * exp(0.5 * (lfact(2 * m) - lfact(m + n) - lfact(m - n)));
  double dl = norm * pow(cos(beta / 2.), m + n) * pow(-sin(beta / 2.), m - n);
  double dlm1 = 0.0;
  for (int i = 0; i < nl; ++i) {
    double l = m + i;
    out[i] = dl * sqrt(2. / (2. * l + 1.));
    double nA = sqrt((2 * l + 3) / (2 * l + 1));
    double nC = sqrt((2 * l + 3) / (2 * l - 1));
    double den = sqrt((l + 1) * (l + 1) - m * m) * sqrt((l + 1) * (l + 1) - n * n);
    double A = nA * (l + 1) * (2 * l + 1) / den;
    double B = (l == 0) ? 0.0 : -A * m * n / l / (l + 1);
    double Cc = (l == 0) ? 0.0 : -nC * sqrt(l * l - m * m) * sqrt(l * l - n * n) / den * (l + 1) / l;
    if (!isfinite(A)) A = 0.0;
    if (!isfinite(B)) B = 0.0;
    if (!isfinite(Cc)) Cc = 0.0;
    double dn = A * mu * dl + B * dl + Cc * dlm1;
    dlm1 = dl;
    dl = dn;
  }
}

/* d^l_{mn}(mu) on ells 2..lmax style grid: out[n_mu][n_l] for l = 2..n_l+1, zero for l < max(m,|n|) */
int oracle_wigner(int m, int n, const double* mu, int n_mu, int n_l, double* out) {
  int l0 = m;  /* start of the recursion */
  int nl = n_l + 2 - l0;
  double* row = (double*)malloc(sizeof(double) * (nl > 0 ? nl : 1));
  for (int i = 0; i < n_mu; ++i) {
    wigner_d_row(m, n, mu[i], nl, row);
    for (int j = 0; j < n_l; ++j) {
      int l = j + 2;
      out[(size_t)i * n_l + j] = l < l0 ? 0.0 : row[l - l0];
    }
  }
  free(row);
  return 0;
}

/* lensed_Cls (abcmb/spectrum.py:437-572) + BB from the same correlation functions.
 * cl_in: [3][n_l] unlensed TT, TE, EE on l = 2..n_l+1; clpp [n_l]; out [4][n_l] */
int oracle_lens(const double* cl_in, const double* clpp, int n_l, const double* mu, const double* w, int n_mu,
                double* out) {
  const double* TT = cl_in;
  const double* TE = cl_in + n_l;
  const double* EE = cl_in + 2 * n_l;
  double* cgl = (double*)malloc(sizeof(double) * n_mu);
  double* cgl2 = (double*)malloc(sizeof(double) * n_mu);
  double* ksi = (double*)malloc(sizeof(double) * n_mu * 4);
  /* pass 1: Cgl, Cgl2 */
#pragma omp parallel
  {
    double* d11 = (double*)malloc(sizeof(double) * n_l);
    double* d1m1 = (double*)malloc(sizeof(double) * n_l);
#pragma omp for schedule(static)
    for (int i = 0; i < n_mu; ++i) {
      oracle_wigner(1, 1, mu + i, 1, n_l, d11);
      oracle_wigner(1, -1, mu + i, 1, n_l, d1m1);
      double s1 = 0.0, s2 = 0.0;
      for (int j = 0; j < n_l; ++j) {
        double l = j + 2;
        double f = (2. * l + 1) * l * (l + 1) * clpp[j];
        s1 += f * d11[j];
        s2 += f * d1m1[j];
      }
      cgl[i] = s1 / 4. / PI;
      cgl2[i] = s2 / 4. / PI;
    }
    free(d11);
    free(d1m1);
  }
  double cgl_top = cgl[n_mu - 1];
#pragma omp parallel
  {
    double* d[12];
    for (int q = 0; q < 12; ++q) d[q] = (double*)malloc(sizeof(double) * n_l);
#pragma omp for schedule(static)
    for (int i = 0; i < n_mu; ++i) {
      /* d00 d11 d1m1 d2m2 d22 d31 d40 d3m3 d4m4 d20 d3m1 d4m2 */
      static const int mn[12][2] = {{0, 0}, {1, 1}, {1, -1}, {2, -2}, {2, 2}, {3, 1},
                                    {4, 0}, {3, -3}, {4, -4}, {2, 0}, {3, -1}, {4, -2}};
      for (int q = 0; q < 12; ++q) oracle_wigner(mn[q][0], mn[q][1], mu + i, 1, n_l, d[q]);
      double s2 = cgl_top - cgl[i];
      double C2 = cgl2[i];
      double a0 = 0, ap = 0, am = 0, ax = 0;
      for (int j = 0; j < n_l; ++j) {
        double l = j + 2;
        double llp1 = l * (l + 1);
        double X000 = exp(-llp1 * s2 / 4);
        double X000p = -llp1 / 4. * X000;
        double X220 = 1. / 4. * sqrt((l + 2) * (l - 1) * l * (l + 1)) * exp(-(llp1 - 2) * s2 / 4.);
        double X022 = exp(-(llp1 - 4) * s2 / 4);
        double X022p = -(llp1 - 4) / 4 * X022;
        double X121 = -1. / 2. * sqrt((l + 2) * (l - 1)) * exp(-(llp1 - 8. / 3.) * s2 / 4.);
        double X132 = -1. / 2. * sqrt((l + 3) * (l - 2)) * exp(-(llp1 - 20. / 3.) * s2 / 4.);
        double X242 = 1. / 4. * sqrt((l + 4) * (l + 3) * (l - 2) * (l - 3)) * exp(-(llp1 - 10.) * s2 / 4.);
        double d00 = d[0][j], d11 = d[1][j], d1m1 = d[2][j], d2m2 = d[3][j], d22 = d[4][j], d31 = d[5][j];
        double d40 = d[6][j], d3m3 = d[7][j], d4m4 = d[8][j], d20 = d[9][j], d3m1 = d[10][j], d4m2 = d[11][j];
        double pre = 2. * l + 1;
        a0 += pre * TT[j] * (X000 * X000 * d00 + 8. / l / (l + 1) * C2 * X000p * X000p * d1m1 +
                             C2 * C2 * (X000p * X000p * d00 + X220 * X220 * d2m2));
        ap += pre * EE[j] * (X022 * X022 * d22 + 2 * C2 * X132 * X121 * d31 +
                             C2 * C2 * (X022p * X022p * d22 + X242 * X220 * d40));
        am += pre * EE[j] * (X022 * X022 * d2m2 + C2 * (X121 * X121 * d1m1 + X132 * X132 * d3m3) +
                             1. / 2. * C2 * C2 * (2 * X022p * X022p * d2m2 + X220 * X220 * d00 + X242 * X242 * d4m4));
        ax += pre * TE[j] * (X022 * X000 * d20 + C2 * 2 * X000p / sqrt(llp1) * (X121 * d11 + X132 * d3m1) +
                             1. / 2. * C2 * C2 * ((2 * X022p * X000p + X220 * X220) * d20 + X220 * X242 * d4m2));
      }
      ksi[i * 4 + 0] = a0 / 4. / PI;
      ksi[i * 4 + 1] = ap / 4. / PI;
      ksi[i * 4 + 2] = am / 4. / PI;
      ksi[i * 4 + 3] = ax / 4. / PI;
    }
    for (int q = 0; q < 12; ++q) free(d[q]);
  }
  /* back-transform: sum over mu */
  for (int q = 0; q < 4 * n_l; ++q) out[q] = 0.0;
  {
    double* d00 = (double*)malloc(sizeof(double) * n_l);
    double* d20 = (double*)malloc(sizeof(double) * n_l);
    double* d22 = (double*)malloc(sizeof(double) * n_l);
    double* d2m2 = (double*)malloc(sizeof(double) * n_l);
    for (int i = 0; i < n_mu; ++i) {
      oracle_wigner(0, 0, mu + i, 1, n_l, d00);
      oracle_wigner(2, 0, mu + i, 1, n_l, d20);
      oracle_wigner(2, 2, mu + i, 1, n_l, d22);
      oracle_wigner(2, -2, mu + i, 1, n_l, d2m2);
      double wi = w[i];
      for (int j = 0; j < n_l; ++j) {
        out[0 * n_l + j] += ksi[i * 4 + 0] * d00[j] * wi;
        out[1 * n_l + j] += ksi[i * 4 + 3] * d20[j] * wi;
        out[2 * n_l + j] += (ksi[i * 4 + 1] * d22[j] + ksi[i * 4 + 2] * d2m2[j]) * wi;
        out[3 * n_l + j] += (ksi[i * 4 + 1] * d22[j] - ksi[i * 4 + 2] * d2m2[j]) * wi;
      }
    }
    for (int j = 0; j < n_l; ++j) {
      out[0 * n_l + j] *= 2 * PI;
      out[1 * n_l + j] *= 2 * PI;
      out[2 * n_l + j] *= 1. / 2. * 2 * PI;
      out[3 * n_l + j] *= 1. / 2. * 2 * PI;
    }
    free(d00); free(d20); free(d22); free(d2m2);
  }
  free(cgl);
  free(cgl2);
  free(ksi);
  return 0;
}

/* OpenMP threads of the oracle's parallel loops (the CPU-baseline timing runs
 * it once on every host core and once on one). */
int oracle_set_threads(int n) {
  if (n < 1) return 1;
  omp_set_num_threads(n);
  return 0;
}
int oracle_max_threads(void) { return omp_get_max_threads(); }
