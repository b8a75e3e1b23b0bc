// Discrete adjoint (reverse mode) of the perturbation ODE, FP64.
//
// The transpose of the tangent-linear kernel's map (abcmb_pert_tangent): for
// one cotangent dL/dT of every cosmology's perturbation table it returns
// dL/d(AbcmbCosmo double fields) and dL/d(entries of the tau(ln a), x_e and
// T_m tables), which a caller chains through the background and
// recombination the way the reference's jax.vjp does (Model(adjoint=
// RecursiveCheckpointAdjoint), abcmb/main.py:111, through
// abcmb/perturbations.py:75-117, :250-404).
//
// One work item (one k-mode of one cosmology) per workgroup, persistent grid
// fed from a device work queue.  Per item:
//   1. the forward kernel's integration (ode_item, value only) without the
//      table writes, recording every accepted step's (t, h), a checkpoint
//      (y, f) every SEG = 64 accepted steps and the final state, in the
//      workgroup's slice of the workspace;
//   2. the accepted steps backwards, one segment of SEG at a time, last
//      first: the segment is recomputed from its checkpoint with the recorded
//      steps (the same stage solves, bit for bit), its states, first stage
//      derivatives and stages kept in the workgroup's slice, then every step
//      of it is transposed (below);
//   3. the transposed initial conditions.
//
// One accepted step from (y, f0 = A(t) y) over h with stages z_s, s = 1..6,
// z_s = b_s + g A_s z_s, b_s = y + h sum_{j<s} a_sj F_j, F_s = A_s z_s,
// y' = z_6, f' = A(t + h) y', and the saved points inside it
// y_save = h00 y + h10 h f0 + h01 y' + h11 h f' (cubic Hermite), is
// linearised like the tangent kernel: the step sizes fixed, each stage the
// implicit-function derivative dz_s = M_s^-1 (db_s + g dA_s z_s), M_s =
// I - g A_s, dF_s = A_s dz_s + dA_s z_s.  Its transpose, with cotangents
// (ly, lf) of (y', f'):
//   saved points: ly_n += h00 ls, lF_0 += h10 h ls, ly += h01 ls, lf += h11 h ls
//   f':           ly += A(t + h)^T lf
//   stage s = 6..1: mu = M_s^-T ([s = 6] ly + A_s^T lF_s),
//                 ly_n += mu, lF_j += h a_sj mu (j < s),
//                 the scalars of A_s get (lF_s + g mu)^T (dA_s/dscalar) z_s
//   carried to the step before: (ly, lf) <- (ly_n, lF_0).
// M_s^-T x is iterated with the step's chord factors P = I - g A(t_jac),
// mu <- mu + P^-T (x - M_s^T mu), until the update times rate / (1 - rate)
// is below 1e-6 of mu (weighted by atol + rtol |z_s|).  P^-T comes from the
// same block Gauss-Jordan and Woodbury pieces as P^-1: with S = I - g B
// (block diagonal), U = g (u_h, u_e) and V = (v_h, v_e),
//   P^-T x = S^-T x + S^-T V C^-T U^T S^-T x,  C = I - V^T S^-1 U,
// S^-T read from the rows of S^-1 staged in LDS.
//
// Background scalars: A depends on the parameters only through the scalars
// sc[slot](ln a) of each evaluation time.  Their cotangents, a workgroup sum
// per time, are pulled back to the 15 inputs of the background evaluation (12
// cosmology fields and the interpolated tau, x_e and T_m values) through the
// Jacobian d sc / d input, which one-direction dual lanes evaluate with the
// forward kernel's own background code; the value cotangents go to the table
// entries through the interpolation weights (atomic adds).  The table fields
// at a saved point depend on the state rows and on the background there: the
// row cotangents come from a one-direction dual evaluation per output row,
// the background inputs' from one per input.
#include "pert_core.h"

namespace abcmb {
namespace {

constexpr int SEG = 64;       // accepted steps per checkpoint
constexpr int NSTORE = 8;     // rows kept per recomputed step: y, F_0, z_1..z_6
constexpr int NBG = 15;       // inputs of the background scalars
constexpr int NBG_FIELDS = 12;
constexpr int DIR_TAU = 12, DIR_XE = 13, DIR_TM = 14;
// the cosmology fields the background evaluation reads (bg_eval, fill_scalars, all_fields)
__constant__ int BG_FIELD[NBG_FIELDS] = {F_OMEGA_B, F_OMEGA_CDM, F_OMEGA_LAMBDA, F_TCMB0, F_YHE, F_N_UR,
                                         F_T_UR, F_N_NCDM, F_T_NCDM, F_M_NCDM, F_XE_LAST, F_TM_LAST};
constexpr int TMAX = 8;           // terms per column of the block-diagonal part of A
constexpr double ADJ_RTOL = 1e-6;
constexpr int UNIT_AT = 30;       // seeds.unit[UNIT_AT] = 1, the rest 0

// the seed arrays in the workspace: unit (64 doubles), ones and zeros (tab_max each)
struct Seeds {
  const double* unit;
  const double* ones;
  const double* zeros;
};

// a one-direction tangent of the inputs: cosmology field f (its par row is a
// shifted window of the unit vector), or none (f < 0)
__device__ __forceinline__ AbcmbTangent seed_field(const Seeds& S, int f) {
  AbcmbTangent t;
  t.par = f >= 0 ? S.unit + UNIT_AT - f : S.zeros;
  t.tau_tab = S.zeros;
  t.xe_tab = S.zeros;
  t.tm_tab = S.zeros;
  t.pt = nullptr;
  return t;
}

// background input d: field BG_FIELD[d], or the tau / x_e / T_m table value
// (every entry's tangent 1, so an interpolated value moves by 1)
__device__ __forceinline__ AbcmbTangent seed_bg(const Seeds& S, int d) {
  AbcmbTangent t = seed_field(S, d < NBG_FIELDS ? BG_FIELD[d] : -1);
  if (d == DIR_TAU) t.tau_tab = S.ones;
  if (d == DIR_XE) t.xe_tab = S.ones;
  if (d == DIR_TM) t.tm_tab = S.ones;
  return t;
}

// lam times the interpolation weights of fast_interp(x) into tab (atomic adds)
__device__ __forceinline__ void tab_scatter(double* tab, int n, double x, double lo, double hi, double lam) {
  double i = (x - lo) / (hi - lo) * (n - 1);
  const double eps = 1e-6;
  i = fmin(fmax(i, eps), n - 1.0 - eps);
  const int il = (int)floor(i);
  const int iu = min(il + 1, n - 1);
  const double wu = i - il;
  atomicAdd(tab + il, (1.0 - wu) * lam);
  atomicAdd(tab + iu, wu * lam);
}

// the cotangent lam of table value d (DIR_TAU, DIR_XE, DIR_TM) at ln a into
// the table entries it was interpolated from (the lookups of bg_eval / bg_xe)
__device__ __forceinline__ void scatter_value(const AbcmbCosmo& c, const AbcmbCotangent& C, int d, double lna,
                                              double lam) {
  if (lam == 0.0) return;
  if (d == DIR_TAU) {
    tab_scatter(C.tau_tab, ABCMB_TAU_TAB_N, lna, ABCMB_TAU_TAB_LO, 0.0, lam);
  } else if (d == DIR_XE) {
    if (lna < c.xe_x0) atomicAdd(C.xe_tab, lam);
    else if (lna < c.xe_lna_last) tab_scatter(C.xe_tab, c.xe_n, lna, c.xe_x0, c.xe_x0 + c.xe_n * c.xe_dx, lam);
  } else if (d == DIR_TM) {
    if (lna >= c.tm_x0 && lna < c.tm_lna_last)
      tab_scatter(C.tm_tab, c.tm_n, lna, c.tm_x0, c.tm_x0 + c.tm_n * c.tm_dx, lam);
  }
}

// d sc[slot] / d (background input d) at ln a, into out[SC_COUNT]
__device__ __noinline__ void jac_lane(const AbcmbCosmo& c, const Layout& L, double k, const Seeds& S, int d,
                                      double lna, double* out) {
  const AbcmbTangent td = seed_bg(S, d);
  const Cosmo<1> P1{c, &td};
  Bg<1> b;
  bg_eval(P1, lna, b);
  fill_scalars(P1, L, k, b, nullptr, out);
}

// the rows of the state in LDS as one-direction duals, seeded at state row
// `seed` (none for seed < 0)
template <int LD>
struct RowsSeed {
  const double (*ys)[LD];
  const Layout& L;
  int seed;
  __device__ __forceinline__ D<1> at(int i) const {
    D<1> x(ys[0][i]);
    x.d[0] = i == seed ? 1.0 : 0.0;
    return x;
  }
  __device__ __forceinline__ D<1> eta() const { return at(0); }
  __device__ __forceinline__ D<1> dc() const { return at(1); }
  __device__ __forceinline__ D<1> db() const { return at(2); }
  __device__ __forceinline__ D<1> tb() const { return at(3); }
  __device__ __forceinline__ D<1> g(int m) const { return at(L.ig + m); }
  __device__ __forceinline__ D<1> pol(int m) const { return at(L.ip + m); }
  __device__ __forceinline__ D<1> ur(int m) const { return at(L.iu + m); }
  __device__ __forceinline__ D<1> nc(int b, int m) const { return at(L.inc + b * (L.ln + 1) + m); }
  __device__ __forceinline__ D<1> fl(int j) const { return at(L.ifl + j); }
};

// sum_f lam_f d field_f / d (state row `seed`) at a saved point (background b0
// held fixed); lam_f = lp[f * fs]
template <int LD>
__device__ __noinline__ double fields_row_adj(const AbcmbCosmo& c, const Seeds& S, const Layout& L, double k,
                                              const Bg<0>& b0, const double (*ys)[LD], int seed,
                                              const double* __restrict__ lp, size_t fs) {
  const AbcmbTangent tz = seed_field(S, -1);
  const Cosmo<1> P{c, &tz};
  Bg<1> b;
  b.a = b0.a;
  b.aH = D<1>(b0.aH.v); b.tau = D<1>(b0.tau.v); b.tauc = D<1>(b0.tauc.v); b.cs2 = D<1>(b0.cs2.v);
  b.R = D<1>(b0.R.v); b.rc = D<1>(b0.rc.v); b.rb = D<1>(b0.rb.v); b.rg = D<1>(b0.rg.v);
  b.ru = D<1>(b0.ru.v); b.rn = D<1>(b0.rn.v); b.pn = D<1>(b0.pn.v); b.Tn = D<1>(b0.Tn.v);
  b.xn = D<1>(b0.xn.v);
  double acc = 0.0;
  all_fields(P, L, k, b, RowsSeed<LD>{ys, L, seed}, [&](int f, const D<1>& v) { acc += lp[f * fs] * v.d[0]; });
  return acc;
}

// sum_f lam_f d field_f / d (background input d) at a saved point (rows fixed)
template <int LD>
__device__ __noinline__ double fields_bg_adj(const AbcmbCosmo& c, const Seeds& S, const Layout& L, double k,
                                             double lna, const double (*ys)[LD], int d,
                                             const double* __restrict__ lp, size_t fs) {
  const AbcmbTangent td = seed_bg(S, d);
  const Cosmo<1> P{c, &td};
  Bg<1> b;
  bg_eval(P, lna, b);
  double acc = 0.0;
  all_fields(P, L, k, b, RowsSeed<LD>{ys, L, -1}, [&](int f, const D<1>& v) { acc += lp[f * fs] * v.d[0]; });
  return acc;
}

// d y0_r / d (om, R_nu, R_b, tau(t_start)) times lam, for q = 0..3
__device__ __noinline__ double ic_adj(const AbcmbCosmo& c, const Seeds& S, const Layout& L, double k, double tiv,
                                      int r, int q, double lam) {
  const AbcmbTangent tq = seed_field(S, q == 0 ? F_OM : (q == 1 ? F_R_NU : (q == 2 ? F_R_B : -1)));
  const Cosmo<1> P{c, &tq};
  D<1> ti(tiv);
  ti.d[0] = q == 3 ? 1.0 : 0.0;
  return ic_row(P, L, k, ti, r).d[0] * lam;
}

// the column of B (the block-diagonal part of A) that this thread's row is:
// B[r][c] = cs sc[slot] for the entries (r, cs, slot); row | slot << 16 packed
struct ColT {
  int rs[TMAX];
  double cs[TMAX];
};

__device__ __forceinline__ bool col_terms(const Layout& L, double k, int c, ColT& T) {
#pragma unroll
  for (int j = 0; j < TMAX; ++j) { T.rs[j] = (SC_ZERO << 16); T.cs[j] = 0.0; }
  int cnt = 0;
  if (c < L.n) {
    const Blk B = row_block(L, c);
    for (int r = B.start; r < B.start + B.size; ++r) {
      Row w;
      row_desc(L, k, r, w);
#pragma unroll
      for (int j = 0; j < 5; ++j) {
        if (w.cs[j] != 0.0 && w.col[j] == c) {
#pragma unroll
          for (int q = 0; q < TMAX; ++q)
            if (q == cnt) { T.rs[q] = r | (w.si[j] << 16); T.cs[q] = w.cs[j]; }
          ++cnt;
        }
      }
    }
  }
  return cnt <= TMAX;
}

// this row's part of P^-T: X = S^-T (v_h, v_e)
struct TInv {
  double Xh, Xe;
};

template <int NT, int BMAX>
struct AdjSmem {
  Smem<NT, BMAX, 0> f;                  // the forward kernel's LDS (value only)
  double jac[5][NBG][SC_COUNT];         // d sc / d input at the step's stage times (slots 1..5)
  double sinv[NT][BMAX];                // rows of S^-1 (block-local columns)
  double vw[2][NT];                     // v_h, v_e at the chord time
  double lam[2][NT];                    // vectors gathered by A^T products (alternating)
  double zth[2][NT];                    // states gathered by the scalar cotangents (alternating)
  double wred[2][SC_COUNT * (NT / 64)];  // the scalar cotangents' workgroup sums
  int sh[2];                            // accepted steps, forward status
};

// what the adjoint records during the forward integration (ode_item's hook)
struct AdjRecord {
  static constexpr bool kSave = false;
  double* rt;
  double* rh;
  double* cky;
  double* ckf;
  double* fy;
  double* ff;
  int* sh;
  int ld;
  int nacc;
  template <int N>
  __device__ __forceinline__ void accept(double t, double hh, const D<N>& y, const D<N>& f0) {
    const int r = threadIdx.x;
    if (r == 0) { rt[nacc] = t; rh[nacc] = hh; }
    if (nacc % SEG == 0) {
      cky[(size_t)(nacc / SEG) * ld + r] = y.v;
      ckf[(size_t)(nacc / SEG) * ld + r] = f0.v;
    }
    ++nacc;
  }
  template <int N>
  __device__ __forceinline__ void finish(int status, const D<N>& y, const D<N>& f0) {
    const int r = threadIdx.x;
    fy[r] = y.v;
    ff[r] = f0.v;
    if (r == 0) { sh[0] = nacc; sh[1] = status; }
  }
};

// stores the converged stages of a recomputed step
struct StoreStage {
  double* zrow;
  int ld;
  __device__ __forceinline__ void operator()(int s, const D<0>& z) const { zrow[(1 + s) * ld + threadIdx.x] = z.v; }
};

// value scalars at the step's seven times (as set_times) and d sc / d input at
// the stage times t + c_s h, slots 1..5 (lanes NTIMES .. NTIMES + 5 NBG - 1)
template <int NT, int BMAX>
__device__ void set_times_adj(AdjSmem<NT, BMAX>& sm, const Cosmo<0>& P, const Layout& L, double k, const Seeds& S,
                              double t, double hh, double jac_offset) {
  static_assert(NTIMES + 5 * NBG <= NT, "set_times_adj: one lane per time and input");
  __syncthreads();
  const int r = threadIdx.x;
  if (r < NTIMES) time_scalars(P, L, k, t + (r == JAC_SLOT ? jac_offset : TIME_C[r]) * hh, sm.f.sc[r]);
  const int q = r - NTIMES;
  if (q >= 0 && q < 5 * NBG) jac_lane(P.c, L, k, S, q % NBG, t + TIME_C[1 + q / NBG] * hh, sm.jac[q / NBG][q % NBG]);
  __syncthreads();
}

// x -> A(slot)^T x for this thread's row; om receives (u_h . x, u_e . x); z
// (when put_z) goes to zth[tb] for a gather after the same barrier
template <int NT, int BMAX, int NX>
__device__ __forceinline__ double at_apply(AdjSmem<NT, BMAX>& sm, const RowState& R, const ColT& T, int& rs, int& lb,
                                           int slot, double x, double (&ex)[NX], double* om) {
  const int r = threadIdx.x;
  sm.lam[lb][r] = r < R.n ? x : 0.0;
  double v[2 + NX];
  v[0] = R.w.uh * x;
  v[1] = R.w.ue * x;
#pragma unroll
  for (int i = 0; i < NX; ++i) v[2 + i] = ex[i];
  wg_sum(sm.f, rs, v);
#pragma unroll
  for (int i = 0; i < NX; ++i) ex[i] = v[2 + i];
  if (om) { om[0] = v[0]; om[1] = v[1]; }
  const double* sc = sm.f.sc[slot][0];
  const double* lm = sm.lam[lb];
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < TMAX; ++j) acc += T.cs[j] * sc[T.rs[j] >> 16] * lm[T.rs[j] & 0xffff];
  acc += R.w.whs * sc[R.w.whi] * v[0] + R.w.wes * sc[R.w.wei] * v[1];
  lb ^= 1;
  return r < R.n ? acc : 0.0;
}

// P^-T b for this thread's row
template <int NT, int BMAX>
__device__ __forceinline__ double pt_solve(AdjSmem<NT, BMAX>& sm, const RowState& R, int& rs, const Inv<BMAX>& F,
                                           const TInv& X, double g, double b) {
  const int r = threadIdx.x;
  sm.f.rl[r] = r < R.n ? b : 0.0;
  __syncthreads();
  double w = 0.0;
#pragma unroll
  for (int cc = 0; cc < BMAX; ++cc)
    if (cc < R.B.size) w += sm.sinv[R.B.start + cc][R.B.lr] * sm.f.rl[R.B.start + cc];
  double v[2] = {g * sm.f.uh[r] * w, g * sm.f.ue[r] * w};
  wg_sum(sm.f, rs, v);
  return r < R.n ? w + X.Xh * (F.c00 * v[0] + F.c10 * v[1]) + X.Xe * (F.c01 * v[0] + F.c11 * v[1]) : 0.0;
}

// the rows of S^-1 and the column weights into LDS, then X = S^-T V
template <int NT, int BMAX>
__device__ __forceinline__ TInv transposed_setup(AdjSmem<NT, BMAX>& sm, const RowState& R, const Inv<BMAX>& F) {
  const int r = threadIdx.x;
  __syncthreads();  // earlier readers of sinv / vw are done
#pragma unroll
  for (int cc = 0; cc < BMAX; ++cc) sm.sinv[r][cc] = F.row[cc];
  sm.vw[0][r] = r < R.n ? F.Vh : 0.0;
  sm.vw[1][r] = r < R.n ? F.Ve : 0.0;
  __syncthreads();
  TInv X{0.0, 0.0};
#pragma unroll
  for (int cc = 0; cc < BMAX; ++cc) {
    if (cc < R.B.size) {
      const double s = sm.sinv[R.B.start + cc][R.B.lr];
      X.Xh += s * sm.vw[0][R.B.start + cc];
      X.Xe += s * sm.vw[1][R.B.start + cc];
    }
  }
  return X;
}

// mu = (I - g A(slot))^-T x by iterative refinement with P^-T; wt: the row's
// weight atol + rtol |z|.  fail: a non-finite cotangent
template <int NT, int BMAX>
__device__ __forceinline__ double adj_solve(AdjSmem<NT, BMAX>& sm, const RowState& R, const ColT& T, int& rs,
                                            int& lb, const Inv<BMAX>& F, const TInv& X, double g, int slot,
                                            double x, double wt, int maxit, int& iters, bool& fail) {
  double mu = pt_solve(sm, R, rs, F, X, g, x);
  double d = mu;      // the latest update (the first is mu itself)
  double prev = 0.0;  // its predecessor's weighted norm
  for (int it = 0;; ++it) {
    ++iters;
    double ex[2] = {(d * wt) * (d * wt), (mu * wt) * (mu * wt)};
    const double q = at_apply(sm, R, T, rs, lb, slot, mu, ex, nullptr);
    const double dn = sqrt(ex[0]), mn = sqrt(ex[1]);
    if (!(mn > 0.0)) {  // a zero cotangent, or a non-finite one
      if (!isfinite(mn)) fail = true;
      break;
    }
    if (!isfinite(dn)) { fail = true; break; }
    if (it >= 1) {
      if (dn <= ADJ_RTOL * mn) break;   // the last update is below the tolerance
      if (it >= 2) {
        // remaining error ~ dn th / (1 - th) with the contraction rate th
        const double th = dn / prev;
        if (th < 1.0 && dn * th <= ADJ_RTOL * mn * (1.0 - th)) break;
      }
    }
    // the iteration cap of the tangent-linear kernel's stage tangents (which
    // stop there too): the transposed iteration has the same contraction rate
    if (it >= maxit) break;
    prev = dn;
    d = pt_solve(sm, R, rs, F, X, g, x - mu + g * q);
    mu += d;
  }
  return mu;
}

// per-thread partial cotangents of the scalars of one time, by term: the five
// descriptor terms, then this row's weights in h' and eta'
struct ThetaAcc {
  double a[7];
};

// adds omega^T (dA/dscalar) z to TA; om = (u_h . omega, u_e . omega); z must
// be in zth[tb] (ordered by a barrier since it was written)
template <int NT, int BMAX>
__device__ __forceinline__ void theta_gather(const AdjSmem<NT, BMAX>& sm, const RowState& R, int tb, double omega,
                                             double z, const double* om, ThetaAcc& TA) {
  const double* zt = sm.zth[tb];
#pragma unroll
  for (int j = 0; j < 5; ++j) TA.a[j] += R.w.cs[j] * omega * zt[R.w.col[j]];
  TA.a[5] += om[0] * R.w.whs * z;
  TA.a[6] += om[1] * R.w.wes * z;
}

// theta_gather with its own sum for om
template <int NT, int BMAX>
__device__ __forceinline__ void theta_add(AdjSmem<NT, BMAX>& sm, const RowState& R, int& rs, int& tb, double omega,
                                          double z, ThetaAcc& TA) {
  const int r = threadIdx.x;
  sm.zth[tb][r] = r < R.n ? z : 0.0;
  double v[2] = {R.w.uh * omega, R.w.ue * omega};
  wg_sum(sm.f, rs, v);
  theta_gather(sm, R, tb, omega, z, v, TA);
  tb ^= 1;
}

// the scalar cotangents of one time (a workgroup sum over the rows), pulled
// back through jac[ji] by the input lanes: field cotangents into accf,
// table-value cotangents into the tables at ln a
template <int NT, int BMAX>
__device__ __forceinline__ void theta_reduce(AdjSmem<NT, BMAX>& sm, const RowState& R, int& wb, ThetaAcc& TA, int ji,
                                             double lna, const AbcmbCosmo& c, const AbcmbCotangent& C,
                                             double& accf) {
  double v[SC_COUNT];
#pragma unroll
  for (int s = 0; s < SC_COUNT; ++s) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < 5; ++j) a += R.w.si[j] == s ? TA.a[j] : 0.0;
    a += R.w.whi == s ? TA.a[5] : 0.0;
    a += R.w.wei == s ? TA.a[6] : 0.0;
    v[s] = a;
  }
  block_sum<NT, SC_COUNT>(v, sm.wred[wb]);
  wb ^= 1;
#pragma unroll
  for (int j = 0; j < 7; ++j) TA.a[j] = 0.0;
  const int d = threadIdx.x;
  if (d < NBG) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < SC_COUNT; ++q) s += v[q] * sm.jac[ji][d][q];
    if (d < NBG_FIELDS) accf += s;
    else scatter_value(c, C, d, lna, s);
  }
}

template <int NT, int BMAX>
__device__ __forceinline__ void adj_item(const int item, AdjSmem<NT, BMAX>& sm, const AbcmbCosmo* __restrict__ cosmos,
                                         const AbcmbCotangent* __restrict__ cots,
                                         const AbcmbPertItem* __restrict__ items, const AbcmbOdeSettings& st,
                                         const Bg<0>* __restrict__ bgsave, int n_save_max, int tab_max,
                                         const Seeds& S, double* __restrict__ wgs, int max_ck,
                                         int32_t* __restrict__ status_out, int32_t* __restrict__ steps_out,
                                         int32_t* __restrict__ iters_out) {
  const int r = threadIdx.x;
  double* rec_t = wgs;
  double* rec_h = rec_t + st.max_steps;
  double* cky = rec_h + st.max_steps;
  double* ckf = cky + (size_t)max_ck * NT;
  double* fin_y = ckf + (size_t)max_ck * NT;
  double* fin_f = fin_y + NT;
  double* zb = fin_f + NT;  // [SEG][NSTORE][NT]

  // 1. the forward integration, recorded
  if (r == 0) { sm.sh[0] = 0; sm.sh[1] = ABCMB_K_BAD_INPUT; }
  __syncthreads();
  ode_item<NT, BMAX, 0, AdjRecord>(item, sm.f, cosmos, nullptr, items, st, bgsave, n_save_max, status_out,
                                   steps_out, nullptr, nullptr,
                                   AdjRecord{rec_t, rec_h, cky, ckf, fin_y, fin_f, sm.sh, NT, 0});
  __syncthreads();
  const int nacc = sm.sh[0];
  if (sm.sh[1] != ABCMB_K_OK) {
    if (r == 0) iters_out[item] = 0;
    return;
  }

  const AbcmbPertItem I = items[item];
  const AbcmbCosmo& c = cosmos[I.cosmo];
  if (c.xe_n > tab_max || c.tm_n > tab_max) {  // the seed arrays are shorter than its tables
    if (r == 0) { status_out[item] = ABCMB_K_BAD_INPUT; iters_out[item] = 0; }
    return;
  }
  const AbcmbCotangent& C = cots[I.cosmo];
  const Cosmo<0> P{c, nullptr};
  const double k = I.k;
  const Layout L = make_layout(st);
  const int n = L.n;
  RowState R;
  row_desc_call(L, k, r, R.w);
  R.B = row_block(L, r);
  R.n = n;
#pragma unroll
  for (int j = 0; j < 5; ++j) R.lc[j] = R.w.col[j] - R.B.start;
  ColT T;
  const bool cols_ok = col_terms(L, k, r, T);
  const int maxb = max(L.iu - 2, max(L.lu + 1, L.nq > 0 ? L.ln + 1 : 1));
  auto one_wave = [](int lo, int hi) { return (lo >> 6) == (hi >> 6); };
  bool wl = one_wave(2, L.iu - 1) && one_wave(L.iu, L.inc - 1);
  for (int i = 0; i < L.nq; ++i) wl = wl && one_wave(L.inc + i * (L.ln + 1), L.inc + (i + 1) * (L.ln + 1) - 1);

  const double rtol = I.rtol, atol = I.atol;
  const double inv_n = 1.0 / n;
  const int maxit = st.max_newton + TANGENT_EXTRA;
  const int n_save = c.n_save;
  const double* __restrict__ lna_save = c.lna_save;
  const Bg<0>* __restrict__ bgs = bgsave + (size_t)I.cosmo * n_save_max;
  const size_t fs = (size_t)n_save * c.nk;   // field stride of the table cotangent
  constexpr int LDZ = Smem<NT, BMAX, 0>::LDZ;

  int rs = 0, lb = 0, tb = 0, wb = 0;
  int iters = 0;
  bool fail = !cols_ok;
  double accf = 0.0;   // input lanes d < NBG_FIELDS: cotangent of cosmology field BG_FIELD[d]
  Inv<BMAX> FI;
  ThetaAcc TA;
#pragma unroll
  for (int j = 0; j < 7; ++j) TA.a[j] = 0.0;

  // the cotangent of the saved rows at saved point `is` (state row ys of this
  // thread); the input lanes add the fields' background part
  auto save_adj = [&](int is, double ys) -> double {
    __syncthreads();  // earlier readers of zl are done
    sm.f.zl[0][r] = r < n ? ys : 0.0;
    __syncthreads();
    const double* lp = C.pt + (size_t)is * c.nk + I.ik;
    double lrow = 0.0;
    if (r < n && out_slot(L, r) >= 0) lrow = fields_row_adj<LDZ>(c, S, L, k, bgs[is], sm.f.zl, r, lp, fs);
    if (r < NBG) {
      const double s = fields_bg_adj<LDZ>(c, S, L, k, lna_save[is], sm.f.zl, r, lp, fs);
      if (r < NBG_FIELDS) accf += s;
      else scatter_value(c, C, r, lna_save[is], s);
    }
    return lrow;
  };

  // 2. the reverse sweep
  double ly = 0.0, lf = 0.0;   // cotangents of the state and of f at the end of the step
  int isave = n_save - 1;
  const int nseg = (nacc + SEG - 1) / SEG;
  for (int j = nseg - 1; j >= 0; --j) {
    const int s0 = j * SEG, s1 = min(nacc, s0 + SEG);
    // recompute the segment from its checkpoint
    double fnext;
    {
      D<0> y(cky[(size_t)j * NT + r]), f0(ckf[(size_t)j * NT + r]);
      for (int i = s0; i < s1; ++i) {
        const double t = rec_t[i], hh = rec_h[i];
        set_times(sm.f, P, L, k, t, hh, NTIMES, st.jac_offset);
        bool ok = factor(sm.f, R, rs, maxb, wl, hh * KV_G, FI);
        double* zrow = zb + (size_t)(i - s0) * NSTORE * NT;
        zrow[r] = y.v;
        zrow[NT + r] = f0.v;
        D<0> z;
        int it_fwd = 0;
        ok = step_stages(sm.f, R, rs, FI, st, ok, hh, rtol, atol, inv_n, y, f0, z, it_fwd, StoreStage{zrow, NT});
        if (!ok) fail = true;
        f0 = rhs_eval(sm.f, R, rs, 5, z);
        y = z;
      }
      fnext = f0.v;
    }
    for (int i = s1 - 1; i >= s0; --i) {
      const double t = rec_t[i], hh = rec_h[i];
      const double tn = i + 1 < nacc ? rec_t[i + 1] : 0.0;
      const double g = hh * KV_G;
      const double* zrow = zb + (size_t)(i - s0) * NSTORE * NT;
      const double yv = zrow[r], f0v = zrow[NT + r];
      double zs[7];
      zs[0] = yv;
#pragma unroll
      for (int s = 1; s < 7; ++s) zs[s] = zrow[(1 + s) * NT + r];
      const double fnew = i + 1 < s1 ? zrow[NSTORE * NT + NT + r] : fnext;
      set_times_adj(sm, P, L, k, S, t, hh, st.jac_offset);
      if (!factor(sm.f, R, rs, maxb, wl, g, FI)) fail = true;
      const TInv X = transposed_setup(sm, R, FI);
      double lyn = 0.0;
      double lF[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) lF[q] = 0.0;
      // saved points inside (t, tn]
      while (isave >= 0 && lna_save[isave] > t) {
        const double th = (lna_save[isave] - t) / hh;
        const double h00 = 2 * th * th * th - 3 * th * th + 1, h10 = th * th * th - 2 * th * th + th;
        const double h01 = -2 * th * th * th + 3 * th * th, h11 = th * th * th - th * th;
        const double ys = h00 * yv + h10 * hh * f0v + h01 * zs[6] + h11 * hh * fnew;
        const double ls = save_adj(isave, ys);
        lyn += h00 * ls;
        lF[0] += h10 * hh * ls;
        ly += h01 * ls;
        lf += h11 * hh * ls;
        --isave;
      }
      // f' = A(t + h) y'
      {
        sm.zth[tb][r] = r < n ? zs[6] : 0.0;
        double ex[1] = {0.0};
        double om[2];
        const double q = at_apply(sm, R, T, rs, lb, 5, lf, ex, om);
        theta_gather(sm, R, tb, lf, zs[6], om, TA);
        tb ^= 1;
        ly += q;
      }
      // the stages, last first
#pragma unroll
      for (int s = 6; s >= 1; --s) {
        const int slot = STAGE_SLOT[s];
        double x = s == 6 ? ly : 0.0;
        if (s < 6) {
          double ex[1] = {0.0};
          x += at_apply(sm, R, T, rs, lb, slot, lF[s < 6 ? s : 0], ex, nullptr);
        }
        const double wt = r < n ? atol + rtol * fabs(zs[s]) : 0.0;
        const double mu = adj_solve(sm, R, T, rs, lb, FI, X, g, slot, x, wt, maxit, iters, fail);
        theta_add(sm, R, rs, tb, (s < 6 ? lF[s < 6 ? s : 0] : 0.0) + g * mu, zs[s], TA);
        lyn += mu;
#pragma unroll
        for (int q = 0; q < 6; ++q)
          if (q < s) lF[q] += hh * KV_A[s][q] * mu;
        if (s <= 5) theta_reduce(sm, R, wb, TA, slot - 1, t + TIME_C[slot] * hh, c, C, accf);
      }
      ly = lyn;
      lf = lF[0];
    }
  }

  // 3. the saved points at or before the start, f0 = A(t_start) y0, the initial conditions
  const double t0 = I.lna_start;
  const double y0 = nacc > 0 ? cky[r] : fin_y[r];
  while (isave >= 0) {
    ly += save_adj(isave, y0);
    --isave;
  }
  set_times(sm.f, P, L, k, t0, 0.0, 1, st.jac_offset);
  if (r < NBG) jac_lane(c, L, k, S, r, t0, sm.jac[0][r]);
  {
    sm.zth[tb][r] = r < n ? y0 : 0.0;
    __syncthreads();  // the jac lanes' writes
    double ex[1] = {0.0};
    double om[2];
    const double q = at_apply(sm, R, T, rs, lb, 0, lf, ex, om);
    theta_gather(sm, R, tb, lf, y0, om, TA);
    tb ^= 1;
    ly += q;
    theta_reduce(sm, R, wb, TA, 0, t0, c, C, accf);
  }
  const double tiv = fast_interp<0>(t0, ABCMB_TAU_TAB_LO, 0.0, c.tau_tab, nullptr, ABCMB_TAU_TAB_N).v;
  double v[5];
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = r < n ? ic_adj(c, S, L, k, tiv, r, q, ly) : 0.0;
  v[4] = fail ? 1.0 : 0.0;
  wg_sum(sm.f, rs, v);
  if (r == 0) {
    atomicAdd(C.par + F_OM, v[0]);
    atomicAdd(C.par + F_R_NU, v[1]);
    atomicAdd(C.par + F_R_B, v[2]);
    scatter_value(c, C, DIR_TAU, t0, v[3]);
    if (v[4] != 0.0) status_out[item] = ABCMB_K_ADJOINT;
    iters_out[item] = iters;
  }
  if (r < NBG_FIELDS && accf != 0.0) atomicAdd(C.par + BG_FIELD[r], accf);
}

template <int NT, int BMAX>
__global__ void __launch_bounds__(NT) pert_adj_kernel(const AbcmbCosmo* __restrict__ cosmos,
                                                      const AbcmbCotangent* __restrict__ cots,
                                                      const AbcmbPertItem* __restrict__ items, AbcmbOdeSettings st,
                                                      const Bg<0>* __restrict__ bgsave, int n_save_max, int tab_max,
                                                      Seeds S, double* __restrict__ scratch, size_t wg_doubles,
                                                      int max_ck, int32_t* __restrict__ status_out,
                                                      int32_t* __restrict__ steps_out, int32_t* __restrict__ iters_out,
                                                      int n_items, int32_t* __restrict__ queue) {
  __shared__ AdjSmem<NT, BMAX> sm;
  __shared__ int next;
  double* wgs = scratch + (size_t)blockIdx.x * wg_doubles;
  for (;;) {
    __syncthreads();  // every thread has left the previous item's LDS work
    if (threadIdx.x == 0) next = atomicAdd(queue, 1);
    __syncthreads();
    const int item = next;
    if (item >= n_items) break;
    adj_item<NT, BMAX>(item, sm, cosmos, cots, items, st, bgsave, n_save_max, tab_max, S, wgs, max_ck, status_out,
                       steps_out, iters_out);
  }
}

__global__ void seed_fill_kernel(double* __restrict__ unit, double* __restrict__ ones, double* __restrict__ zeros,
                                 int tab_max) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 64) unit[i] = i == UNIT_AT ? 1.0 : 0.0;
  if (i < tab_max) { ones[i] = 1.0; zeros[i] = 0.0; }
}

// launch shape of a layout: NT threads (at least 128: the input lanes), BMAX
int adj_shape(const AbcmbOdeSettings& st, int& nt, int& bmax) {
  int maxb = 2 + (st.l_g + 1) + (st.l_pol + 1);
  if (st.l_ur + 1 > maxb) maxb = st.l_ur + 1;
  if (st.n_q > 0 && st.l_ncdm + 1 > maxb) maxb = st.l_ncdm + 1;
  ABCMB_REQUIRE(maxb <= 32, "the adjoint supports hierarchy blocks of at most 32 rows (%d)", maxb);
  ABCMB_REQUIRE(st.n <= 192, "state size %d exceeds 192", st.n);
  nt = st.n <= 128 ? 128 : 192;
  bmax = maxb <= 26 ? 26 : 32;
  return ABCMB_OK;
}

size_t adj_wg_doubles(const AbcmbOdeSettings& st, int nt) {
  const size_t max_ck = (size_t)(st.max_steps + SEG - 1) / SEG;
  const size_t d = 2 * (size_t)st.max_steps + 2 * max_ck * nt + 2 * (size_t)nt + (size_t)SEG * NSTORE * nt;
  return (d + 31) / 32 * 32;
}

template <int NT, int BMAX>
int adj_grid(int& grid) {
  int dev = 0, per_cu = 0;
  ABCMB_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  ABCMB_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  ABCMB_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pert_adj_kernel<NT, BMAX>, NT, 0));
  grid = (per_cu > 0 ? per_cu : 1) * prop.multiProcessorCount;
  return ABCMB_OK;
}

int adj_grid_of(int nt, int bmax, int& grid) {
  if (nt == 128) return bmax == 26 ? adj_grid<128, 26>(grid) : adj_grid<128, 32>(grid);
  return bmax == 26 ? adj_grid<192, 26>(grid) : adj_grid<192, 32>(grid);
}

// workspace: saved-point background table, queue counter, seeds, per-workgroup slices
struct AdjLayout {
  size_t bg, queue, seeds, slices, total;
  int grid, nt, bmax;
};

int adj_layout(int n_cosmo, int n_save_max, int tab_max, const AbcmbOdeSettings& st, AdjLayout& A) {
  int rc = adj_shape(st, A.nt, A.bmax);
  if (rc != ABCMB_OK) return rc;
  rc = adj_grid_of(A.nt, A.bmax, A.grid);
  if (rc != ABCMB_OK) return rc;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  A.bg = 0;
  A.queue = up(sizeof(Bg<0>) * (size_t)n_cosmo * (size_t)n_save_max);
  A.seeds = A.queue + 256;
  A.slices = A.seeds + up(sizeof(double) * (64 + 2 * (size_t)tab_max));
  A.total = A.slices + sizeof(double) * adj_wg_doubles(st, A.nt) * (size_t)A.grid;
  return ABCMB_OK;
}

template <int NT, int BMAX>
void adj_launch(const AbcmbCosmo* cosmos, const AbcmbCotangent* cots, const AbcmbPertItem* items, int n_items,
                const AbcmbOdeSettings& st, const Bg<0>* bgs, int n_save_max, int tab_max, const Seeds& S,
                double* scratch, int grid, int32_t* queue, int32_t* status, int32_t* steps, int32_t* iters,
                hipStream_t s) {
  const int g = grid < n_items ? grid : n_items;
  const int max_ck = (st.max_steps + SEG - 1) / SEG;
  hipLaunchKernelGGL((pert_adj_kernel<NT, BMAX>), dim3((unsigned)g), dim3(NT), 0, s, cosmos, cots, items, st, bgs,
                     n_save_max, tab_max, S, scratch, adj_wg_doubles(st, NT), max_ck, status, steps, iters, n_items,
                     queue);
}

}  // namespace
}  // namespace abcmb

extern "C" size_t abcmb_pert_adjoint_workspace_size(int n_cosmo, int n_save_max, int tab_max,
                                                    const AbcmbOdeSettings* settings) {
  using namespace abcmb;
  if (!settings || n_cosmo <= 0 || n_save_max <= 0 || tab_max <= 0) return 0;
  AdjLayout A;
  if (adj_layout(n_cosmo, n_save_max, tab_max, *settings, A) != ABCMB_OK) return 0;
  return A.total;
}

extern "C" int abcmb_pert_adjoint(const AbcmbCosmo* cosmos_dev, const AbcmbCotangent* cot_dev, int n_cosmo,
                                  const AbcmbPertItem* items_dev, int n_items, const AbcmbOdeSettings* settings,
                                  int n_save_max, int tab_max, int32_t* status_dev, int32_t* steps_dev,
                                  int32_t* iters_dev, void* workspace, size_t ws_bytes, void* stream) {
  using namespace abcmb;
  ABCMB_REQUIRE(cosmos_dev && cot_dev && items_dev && settings && status_dev && steps_dev && iters_dev,
                "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && n_items > 0 && n_save_max > 0, "empty launch");
  ABCMB_REQUIRE(tab_max >= ABCMB_TAU_TAB_N && tab_max >= ABCMB_COSMO_DOUBLES, "tab_max %d below the tau table",
                tab_max);
  int maxb = 0;
  ABCMB_REQUIRE(!fluid::kActive, "abcmb_pert_adjoint covers the standard species; this build has user fluids "
                "(signature '%s'): use abcmb_pert_tangent", fluid::kSignature);
  int rc = check_settings(*settings, maxb);
  if (rc != ABCMB_OK) return rc;
  AdjLayout A;
  rc = adj_layout(n_cosmo, n_save_max, tab_max, *settings, A);
  if (rc != ABCMB_OK) return rc;
  ABCMB_REQUIRE(workspace && ws_bytes >= A.total, "workspace too small (%zu < %zu)", ws_bytes, A.total);
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  Bg<0>* bgs = (Bg<0>*)(ws + A.bg);
  int32_t* queue = (int32_t*)(ws + A.queue);
  double* seeds = (double*)(ws + A.seeds);
  Seeds S{seeds, seeds + 64, seeds + 64 + tab_max};
  double* scratch = (double*)(ws + A.slices);
  hipLaunchKernelGGL(seed_fill_kernel, dim3((tab_max + 255) / 256), dim3(256), 0, s, seeds, seeds + 64,
                     seeds + 64 + tab_max, tab_max);
  hipLaunchKernelGGL((bg_save_kernel<0>), dim3((n_save_max + 63) / 64, n_cosmo), dim3(64), 0, s, cosmos_dev,
                     nullptr, n_save_max, bgs);
  hipLaunchKernelGGL(zero_counts_kernel, dim3(1), dim3(64), 0, s, queue, 1);
  ABCMB_HIP_CHECK(hipGetLastError());
  // a build with user fluids returned above: its adjoint kernels are never instantiated
  if constexpr (!fluid::kActive) {
  if (A.nt == 128) {
    if (A.bmax == 26)
      adj_launch<128, 26>(cosmos_dev, cot_dev, items_dev, n_items, *settings, bgs, n_save_max, tab_max, S, scratch,
                          A.grid, queue, status_dev, steps_dev, iters_dev, s);
    else
      adj_launch<128, 32>(cosmos_dev, cot_dev, items_dev, n_items, *settings, bgs, n_save_max, tab_max, S, scratch,
                          A.grid, queue, status_dev, steps_dev, iters_dev, s);
  } else {
    if (A.bmax == 26)
      adj_launch<192, 26>(cosmos_dev, cot_dev, items_dev, n_items, *settings, bgs, n_save_max, tab_max, S, scratch,
                          A.grid, queue, status_dev, steps_dev, iters_dev, s);
    else
      adj_launch<192, 32>(cosmos_dev, cot_dev, items_dev, n_items, *settings, bgs, n_save_max, tab_max, S, scratch,
                          A.grid, queue, status_dev, steps_dev, iters_dev, s);
  }
  }
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}
