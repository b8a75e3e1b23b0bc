/*
 * abcmb.h - C ABI of the MI355X perturbation + line-of-sight library
 * (libabcmb_hip.so) and of its host-side background producer
 * (libabcmb_host.so).
 *
 * The reference (mshettyFE/ABCMB) is a JAX package with no FFI; its
 * boundary for this path is a set of Python method calls.  Each entry point
 * below replaces one of them and cites it (paths relative to the reference
 * repository root).  Conventions:
 *   - every pointer named *_dev is device memory owned by the caller;
 *     host structs are read at call time only;
 *   - nothing here allocates device memory: scratch comes in through
 *     `workspace` / `ws_bytes` (query the size first), except for
 *     abcmb_pert_fwd_noalloc, whose scratch is a fixed device array;
 *   - calls are stream-ordered on `stream` (a hipStream_t passed as void*);
 *   - return value 0 = success, otherwise an ABCMB_E* code; the message is
 *     available from abcmb_last_error() (thread-local).
 */
#ifndef ABCMB_H
#define ABCMB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ABCMB_OK 0
#define ABCMB_EINVAL 1      /* bad argument / shape / layout */
#define ABCMB_EHIP 2        /* HIP runtime error */
#define ABCMB_EWORKSPACE 3  /* workspace too small */

/* per-k ODE status words (AbcmbPertItem.status output) */
#define ABCMB_K_OK 0
#define ABCMB_K_MAX_STEPS 1   /* max_steps reached before ln a = 0 */
#define ABCMB_K_NONFINITE 2   /* NaN/Inf in state or error norm */
#define ABCMB_K_STEP_UNDERFLOW 3  /* step size fell below the spacing of ln a */
#define ABCMB_K_BAD_INPUT 4   /* the cosmology has more saved points than n_save_max */
#define ABCMB_K_ADJOINT 5     /* adjoint: a non-finite cotangent */

#define ABCMB_TAU_TAB_N 10000      /* tau(ln a) table on [-33, 0] */
#define ABCMB_TAU_TAB_LO (-33.0)
#define ABCMB_N_PT_FIELDS 21       /* fields of the perturbation table */

/* Perturbation-table field order inside AbcmbCosmo.pt ([field][n_save][nk]). */
enum {
  ABCMB_PT_DELTA_M = 0, ABCMB_PT_THETA_B_PRIME, ABCMB_PT_ETA, ABCMB_PT_H_PRIME,
  ABCMB_PT_ETA_PRIME, ABCMB_PT_ALPHA, ABCMB_PT_ALPHA_PRIME,
  ABCMB_PT_CDM_DELTA, ABCMB_PT_B_DELTA, ABCMB_PT_B_THETA,
  ABCMB_PT_G_DELTA, ABCMB_PT_G_THETA, ABCMB_PT_G_SIGMA, ABCMB_PT_G_G0, ABCMB_PT_G_G2,
  ABCMB_PT_UR_DELTA, ABCMB_PT_UR_THETA, ABCMB_PT_UR_SIGMA,
  ABCMB_PT_NCDM_DELTA, ABCMB_PT_NCDM_THETA, ABCMB_PT_NCDM_SIGMA
};

/*
 * One cosmology: parameters (input + derived, abcmb/main.py:368-612), the
 * background tables the hot path reads, and where its perturbation table
 * goes.  Pointers are device pointers for the HIP entry points and host
 * pointers for the host producer.
 */
#define ABCMB_FLUID_PARAMS 8  /* parameter slots of user-defined fluids */
/* AbcmbCosmo.species_off bits: a standard species the model leaves out
 * (use_LCDM_species=False with a user fluid in its place,
 * abcmb/model_specs.py:89-116) */
#define ABCMB_NO_LAMBDA 1
#define ABCMB_NO_CDM 2

typedef struct AbcmbCosmo {
  double h, omega_b, omega_cdm, omega_Lambda;
  double A_s, n_s, k_pivot, TCMB0, YHe;
  double N_ur, T_ur, N_ncdm, T_ncdm, m_ncdm;
  double R_nu, R_b, om, omega_r, omega_m;
  double tau0;                                   /* conformal time today, Mpc */
  double xe_x0, xe_dx, xe_lna_last, xe_last;     /* x_e table metadata */
  double tm_x0, tm_dx, tm_lna_last, tm_last;     /* T_m table metadata */
  double lna_rec, lna_transfer_start, z_reion;
  /* parameters of the model's user fluids (abcmb/species.py:11-249), in the
   * order the model's generated fluid code reads them; zero when unused */
  double fluid_par[ABCMB_FLUID_PARAMS];
  int32_t has_ncdm;   /* MassiveNeutrino species present */
  int32_t n_q;        /* momentum bins of its hierarchy (3 or 5) */
  int32_t xe_n, tm_n; /* allocated table lengths (lookups span all of them) */
  int32_t n_save, nk; /* saved ln a points, PT k-modes */
  int32_t species_off; /* ABCMB_NO_* bits */
  int32_t pad0;
  const double* tau_tab;   /* ABCMB_TAU_TAB_N */
  const double* xe_tab;    /* xe_n */
  const double* tm_tab;    /* tm_n */
  const double* lna_save;  /* n_save, ascending, last = 0 */
  const double* k;         /* nk, Mpc^-1 */
  double* pt;              /* [ABCMB_N_PT_FIELDS][n_save][nk] */
  double* pt_fluid;        /* [n_fluid][n_save][nk]: the user-fluid rows at the
                              saved points (AbcmbOdeSettings.n_fluid > 0 only) */
} AbcmbCosmo;

/* One (cosmology, k-mode) ODE work item. */
typedef struct AbcmbPertItem {
  double k, lna_start, rtol, atol;
  int32_t cosmo, ik, pad0, pad1;
} AbcmbPertItem;

/*
 * Forward-mode tangents of one cosmology's inputs along ABCMB_TANGENT_DIRS
 * parameter directions (directions a caller does not need carry zeros).
 * par[d][f] is the derivative of the f-th double field of AbcmbCosmo (in
 * declaration order, h .. z_reion) along direction d; the tables are the
 * derivatives of AbcmbCosmo's tables on their own (fixed) abscissae.  The
 * tangent of the perturbation table comes back in pt.
 */
#define ABCMB_TANGENT_DIRS 5
#define ABCMB_COSMO_DOUBLES (31 + ABCMB_FLUID_PARAMS)
typedef struct AbcmbTangent {
  const double* par;      /* [ABCMB_TANGENT_DIRS][ABCMB_COSMO_DOUBLES] */
  const double* tau_tab;  /* [ABCMB_TANGENT_DIRS][ABCMB_TAU_TAB_N] */
  const double* xe_tab;   /* [ABCMB_TANGENT_DIRS][xe_n] */
  const double* tm_tab;   /* [ABCMB_TANGENT_DIRS][tm_n] */
  double* pt;             /* out: [ABCMB_TANGENT_DIRS][ABCMB_N_PT_FIELDS][n_save][nk] */
  double* pt_fluid;       /* out: [ABCMB_TANGENT_DIRS][n_fluid][n_save][nk] (n_fluid > 0) */
} AbcmbTangent;

/*
 * Reverse-mode cotangents of one cosmology (abcmb_pert_adjoint): pt is the
 * input, dL/d(perturbation table); the others are outputs, accumulated into
 * (the caller zeroes them): dL/d(AbcmbCosmo double fields, declaration
 * order) and dL/d(table entries) of the tau(ln a), x_e and T_m tables.
 */
typedef struct AbcmbCotangent {
  const double* pt;       /* in:  [ABCMB_N_PT_FIELDS][n_save][nk] */
  double* par;            /* out: [ABCMB_COSMO_DOUBLES] */
  double* tau_tab;        /* out: [ABCMB_TAU_TAB_N] */
  double* xe_tab;         /* out: [xe_n] */
  double* tm_tab;         /* out: [tm_n] */
} AbcmbCotangent;

/* Integrator settings shared by a launch (abcmb/model_specs.py:57-68,
 * abcmb/perturbations.py:285-312). */
typedef struct AbcmbOdeSettings {
  double pcoeff, icoeff, dcoeff, dt0;
  double safety, factormin, factormax, newton_kappa;
  double jac_offset;                /* chord Jacobian at t_n + jac_offset * h */
  int32_t max_steps, max_newton;
  int32_t l_g, l_pol, l_ur, l_ncdm; /* hierarchy cut-offs */
  int32_t n_q, n;                   /* massive bins (0 if absent), state size */
  int32_t n_fluid;                  /* rows of user-defined fluids, the last rows of
                                       the state; nonzero only with a library built
                                       for that fluid set (abcmb_fluid_signature) */
  int32_t pad0;
} AbcmbOdeSettings;

/* tanh reionization of one cosmology (abcmb/background.py:939-1017): the
 * optical depth it must reach and the shape parameters of the hydrogen and
 * helium steps. */
typedef struct AbcmbReion {
  double tau_reion;
  double Delta_z_reion, z_reion_He, Delta_z_reion_He, exp_reion;
} AbcmbReion;

/* ---------------- host background producer (libabcmb_host.so) ---------- */

/*
 * Fill the background tables and scalars of `c` from its parameters:
 * conformal time table (abcmb/background.py:252-319), the recombination
 * history x_e/T_m of HyRex (abcmb/hyrex/) with the tanh reionization
 * added (abcmb/background.py:939-1017), lna_rec and lna_transfer_start
 * (abcmb/background.py:555-564).  c->tau_tab, xe_tab and tm_tab must point
 * to writable host arrays of ABCMB_TAU_TAB_N, xe_n and tm_n doubles
 * (xe_n >= abcmb_host_recomb_points(), tm_n >= abcmb_host_tm_points(): the
 * reference's allocated lengths, which its lookups' span depends on; x_e and
 * T_m are written from their first entries, the tails padded with the last
 * value, and xe_lna_last / tm_x0 / tm_lna_last say where).  tau_reion > 0
 * selects the tau-input reionization branch.  n_eff: N_eff of the
 * cosmology, read by the two-photon correction fit (abcmb/hyrex/hydrogen.py:
 * 538-575).  rec_switch (in/out, may be NULL): ABCMB_REC_STAGES grid counts,
 * where each recombination stage handed over (HeIII Saha, post-Saha HeII,
 * HeII rate equation, post-Saha H, two-photon EMLA, EMLA with T_m, three-level
 * atom); < 0 on input lets the stage's own rule decide (and reports it),
 * >= 0 forces it (finite differences hold them fixed, like every other
 * discrete grid choice).  c->lna_rec and c->lna_transfer_start are outputs,
 * unless both are negative on input: then they are kept (a caller that holds
 * the grids at another cosmology's passes that cosmology's values) and their
 * scans are skipped.
 */
#define ABCMB_REC_STAGES 7
int abcmb_host_background(AbcmbCosmo* c, double tau_reion, double z_reion_in,
                          double Delta_z_reion, double z_reion_He,
                          double Delta_z_reion_He, double exp_reion, double n_eff,
                          int32_t* rec_switch);
int abcmb_host_recomb_points(void);
int abcmb_host_tm_points(void);
/* abcmb_host_background without the reionization: x_e holds the recombination
 * history alone and z_reion is 0 (abcmb_reionize adds the rest on the GPU). */
int abcmb_host_recombination(AbcmbCosmo* c, double n_eff, int32_t* rec_switch);

/*
 * Background vectors on the LOS grid lna[0..n-1] (abcmb/spectrum.py:640-651):
 * out[0..5][n] = tau, g, dg/dlna, aH, exp(-kappa), aH' * aH.
 */
int abcmb_host_los_background(const AbcmbCosmo* c, const double* lna, int n,
                              double* out);
/* Per-k starting ln a (abcmb/perturbations.py:119-159 and :279). */
int abcmb_host_lna_start(const AbcmbCosmo* c, const double* k, int nk,
                         double R_tc, double R_large, double* out);
/* Background at arbitrary ln a: out[0..6][n] = aH, tau, tau_c, cs2, x_e, T_m, R. */
int abcmb_host_eval(const AbcmbCosmo* c, const double* lna, int n, double* out);
/*
 * Forward-mode derivatives of the producer along ABCMB_TANGENT_DIRS
 * parameter directions, with every grid choice held.  Replaces the
 * derivatives of the reference's Background construction that jax.jacfwd /
 * jax.grad take through abcmb/background.py:420-768 and abcmb/hyrex/
 * (Model(adjoint=...), abcmb/main.py:111): the same arithmetic is carried on
 * dual numbers (the value parts equal abcmb_host_background's results bit for
 * bit).  c: a cosmology filled by abcmb_host_background (or
 * abcmb_host_recombination when reionization == 0) with lna_rec and
 * lna_transfer_start negative (held); its tables are rewritten with the same
 * values.  rec_switch: that call's ABCMB_REC_STAGES stage ends (all >= 0).
 * dpar [ABCMB_TANGENT_DIRS][ABCMB_COSMO_DOUBLES], in/out: on input the
 * derivatives of the parameter fields (h .. omega_m, fluid_par), on output also those of
 * tau0, xe_last, tm_last and z_reion (zero for the held grid metadata).
 * dtau_reion, dn_eff [ABCMB_TANGENT_DIRS]: derivatives of the tau_reion and
 * n_eff arguments.  dtau [dirs][ABCMB_TAU_TAB_N], dxe [dirs][xe_n],
 * dtm [dirs][tm_n]: the tables' derivatives.  z_reion's derivative is the
 * implicit-function derivative of the tau_reion root.  A build for user
 * fluids differentiates their generated host code the same way (the fluid_par
 * fields take part like the others).
 */
int abcmb_host_background_tangent(AbcmbCosmo* c, double tau_reion, const double* dtau_reion,
                                  double z_reion_in, double Delta_z_reion, double z_reion_He,
                                  double Delta_z_reion_He, double exp_reion, double n_eff,
                                  const double* dn_eff, const int32_t* rec_switch, int reionization,
                                  double* dpar, double* dtau, double* dxe, double* dtm);
/* abcmb_host_los_background (out [6][n]) and its derivatives dout
 * [ABCMB_TANGENT_DIRS][6][n], given the derivatives dpar, dtau, dxe, dtm of
 * the cosmology's fields and tables (from abcmb_host_background_tangent). */
int abcmb_host_los_background_tangent(const AbcmbCosmo* c, const double* dpar, const double* dtau,
                                      const double* dxe, const double* dtm, const double* lna, int n,
                                      double* out, double* dout);
/* abcmb_host_eval (out [7][n]) and its derivatives dout [ABCMB_TANGENT_DIRS][7][n]. */
int abcmb_host_eval_tangent(const AbcmbCosmo* c, const double* dpar, const double* dtau, const double* dxe,
                            const double* dtm, const double* lna, int n, double* out, double* dout);
/* Baryon drag redshift z_d (baryon optical depth kappa_d = 1) and the sound
 * horizon r_s(z_d) in Mpc (Background.z_d / Background.rs_d,
 * abcmb/background.py:774-936), for a cosmology filled by
 * abcmb_host_background. */
int abcmb_host_drag(const AbcmbCosmo* c, double* z_d, double* rs_d);
/* Signature of the user-fluid code this build of the host producer was
 * compiled with ("" for the standard species only).  Replaces the
 * reference's registration of Fluid classes (Model(user_species=...),
 * abcmb/model_specs.py:89-116; abcmb/species.py:11-249): a model's fluids
 * are turned into C++ once and the producer built for them. */
const char* abcmb_host_fluid_signature(void);

/* ---------------- GPU library (libabcmb_hip.so) ---------------------------- */

const char* abcmb_last_error(void);
int abcmb_device_init(void);
/* Signature of the user-fluid code this build of the kernels was compiled
 * with ("" for the standard species only); the ODE entry points accept
 * settings.n_fluid > 0 only from a build whose fluid code has that many rows
 * (the model's Fluid classes, abcmb/species.py:11-249, registered by
 * Model(user_species=...), abcmb/model_specs.py:89-116). */
const char* abcmb_fluid_signature(void);

/*
 * Launch-shape settings that the autotuner (abcmb_amd/autotune.py) picks by
 * timing and caches on disk; process-wide, 0 = the built-in default.  No
 * counterpart in the reference (XLA chooses its own launch shapes).
 *   ABCMB_TUNE_PERT_PERSISTENT: 0 = one workgroup per ODE work item; v > 0 =
 *     a persistent grid of v% of the resident workgroups pulling items from a
 *     device work queue (abcmb_pert_fwd / _tangent / _noalloc).
 *   ABCMB_TUNE_LENS_GEMM_COLS: column tile of the lensing GEMMs
 *     (abcmb_lens_mfma): 0 or 32 = 32 cosmologies per tile, 64 = 64.
 */
#define ABCMB_TUNE_PERT_PERSISTENT 0
#define ABCMB_TUNE_LENS_GEMM_COLS 1
#define ABCMB_TUNE_KEYS 8
int abcmb_tune_set(int key, int value);
int abcmb_tune_get(int key);

/*
 * Perturbation ODE for every work item: adaptive ESDIRK (Kvaerno5) with
 * PID control and chord Newton in FP64, one k-mode per workgroup, dense
 * output at the cosmology's n_save points, output-table assembly (in the
 * kernel, or from the saved state rows by a second pass on the same stream).
 * Replaces PerturbationEvolver.full_evolution (abcmb/perturbations.py:75-117)
 * with evolution_one_k (:250-316) and make_output_table (:318-404).
 * cosmos_dev: n_cosmo device-resident AbcmbCosmo; items_dev: n_items
 * distinct (cosmology, k) pairs;
 * status_dev/steps_dev: int32[n_items] outputs (status, accepted+rejected
 * steps); iters_dev (optional): int32[n_items] chord-Newton iterations;
 * state_dev: optional [n_items][n_save][n] raw states (NULL = off).
 * n_save_max: upper bound of the cosmologies' n_save; the workspace
 * (abcmb_pert_workspace_size(n_cosmo, n_save_max, 0) bytes) holds the
 * background at every saved point, computed once per cosmology, and a count
 * of the launch's items per cosmology.
 */
size_t abcmb_pert_workspace_size(int n_cosmo, int n_save_max, int n_dir);
int abcmb_pert_fwd(const AbcmbCosmo* cosmos_dev, int n_cosmo,
                   const AbcmbPertItem* items_dev, int n_items,
                   const AbcmbOdeSettings* settings, int n_save_max,
                   int32_t* status_dev, int32_t* steps_dev, int32_t* iters_dev,
                   double* state_dev, void* workspace, size_t ws_bytes, void* stream);

/*
 * Tangent-linear perturbation ODE: the same integration as abcmb_pert_fwd
 * (same step sequence, chord matrices and iteration counts, same values in
 * cosmos[].pt) carrying, with every state row, its derivatives along
 * ABCMB_TANGENT_DIRS parameter directions.  The step sizes are held fixed
 * (the derivative of the discrete scheme at the chosen steps), the chord
 * iterations act on value and tangents together, the initial conditions,
 * background scalars and output-table fields are differentiated exactly.
 * Replaces jax.jacfwd through PerturbationEvolver.full_evolution
 * (abcmb/perturbations.py:75-117, :250-404) for the parameters that reach
 * the perturbation equations.  tangents_dev: n_cosmo device-resident
 * AbcmbTangent (inputs and the tangent-table output pointer); workspace:
 * abcmb_pert_workspace_size(n_cosmo, n_save_max, ABCMB_TANGENT_DIRS) bytes.
 */
int abcmb_pert_tangent(const AbcmbCosmo* cosmos_dev, const AbcmbTangent* tangents_dev, int n_cosmo,
                       const AbcmbPertItem* items_dev, int n_items,
                       const AbcmbOdeSettings* settings, int n_save_max,
                       int32_t* status_dev, int32_t* steps_dev, int32_t* iters_dev,
                       void* workspace, size_t ws_bytes, void* stream);

/*
 * abcmb_pert_fwd without a workspace argument and without any allocation, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No counterpart in
 * the reference.  Its scratch (what abcmb_pert_fwd's workspace holds) is a
 * fixed device array of ABCMB_NOALLOC_SCRATCH_BYTES per GPU, kept until the
 * process ends; a launch needing more (abcmb_pert_workspace_size(n_cosmo,
 * n_save_max, 0) > ABCMB_NOALLOC_SCRATCH_BYTES, e.g. above 139 cosmologies of
 * 500 saved points) is rejected with ABCMB_EINVAL.  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (different streams,
 * host threads) race; calls ordered on one stream do not.  The arguments and
 * results are those of abcmb_pert_fwd.
 */
#define ABCMB_NOALLOC_SCRATCH_BYTES (8 << 20)
/*
 * Discrete adjoint of the perturbation ODE: the transpose of
 * abcmb_pert_tangent's linear map, for one cotangent per cosmology.  Replaces
 * reverse-mode differentiation through PerturbationEvolver.full_evolution with
 * Model(adjoint=RecursiveCheckpointAdjoint) (abcmb/main.py:111,
 * abcmb/perturbations.py:75-117, :250-404).  Each (cosmology, k) item is
 * integrated again as abcmb_pert_fwd does (same step sequence), recording
 * every accepted step and a checkpoint every 64 accepted steps; the accepted
 * steps are then swept backwards segment by segment, each segment recomputed
 * from its checkpoint, with transposed stage solves (the chord matrix as the
 * preconditioner of an iterative refinement to 1e-6), the transposed cubic
 * Hermite dense output injecting the table cotangents at the saved points,
 * and the cotangents of the background scalars and of the initial conditions
 * pulled back to the cosmology fields and table entries they are computed
 * from.  The step sizes are held fixed, as in abcmb_pert_tangent.
 * cot_dev: n_cosmo device-resident AbcmbCotangent.  tab_max: at least
 * ABCMB_TAU_TAB_N and every cosmology's xe_n and tm_n.  iters_dev: transposed
 * refinement iterations per item.  Item status ABCMB_K_ADJOINT: a cotangent
 * became non-finite (what was accumulated before is kept).
 * Supports state sizes up to 192 with hierarchy blocks up to 32 rows.
 */
size_t abcmb_pert_adjoint_workspace_size(int n_cosmo, int n_save_max, int tab_max,
                                         const AbcmbOdeSettings* settings);
int abcmb_pert_adjoint(const AbcmbCosmo* cosmos_dev, const AbcmbCotangent* cot_dev, int n_cosmo,
                       const AbcmbPertItem* items_dev, int n_items,
                       const AbcmbOdeSettings* settings, int n_save_max, int tab_max,
                       int32_t* status_dev, int32_t* steps_dev, int32_t* iters_dev,
                       void* workspace, size_t ws_bytes, void* stream);

int abcmb_pert_fwd_noalloc(const AbcmbCosmo* cosmos_dev, int n_cosmo,
                           const AbcmbPertItem* items_dev, int n_items,
                           const AbcmbOdeSettings* settings, int n_save_max,
                           int32_t* status_dev, int32_t* steps_dev, int32_t* iters_dev,
                           double* state_dev, void* stream);

/*
 * The entry points below work on a batch of n_cosmo cosmologies described by
 * the device-resident AbcmbCosmo array (their perturbation tables, k-grids,
 * save grids and parameters); per-cosmology buffers are contiguous blocks of
 * the stated shape, one after another.
 */

/*
 * Cubic-spline interpolation (not-a-knot, in log10 k) of the 10 LOS fields
 * of each perturbation table from its k grid to the transfer grid, fused with
 * the source functions S_T0, S_T1, S_T2, S_E and the LOS quadrature weight
 * w_i / aH_i (abcmb/spectrum.py:664-696, :755-772).
 * bgv_dev: [B][6][n_save-1] from abcmb_host_los_background; kt_dev: [nkt];
 * src_dev out: [B][4][n_save-1][nkt]; scales4: sw, isw, dop, pol (host).
 */
size_t abcmb_sources_workspace_size(int n_cosmo, int nk_max, int n_save);
int abcmb_sources(const AbcmbCosmo* cosmos_dev, int n_cosmo, int nk_max, int n_save,
                  const double* bgv_dev, const double* kt_dev, int nkt,
                  const double* scales4, double* src_dev,
                  void* workspace, size_t ws_bytes, void* stream);

/*
 * Line-of-sight transfer Delta_l(k) for each sampled multipole and the C_l
 * k-quadrature (abcmb/spectrum.py:616-797, Bessel lookups :48-109).
 * bes_dev: tables [3][n_lt][nx] (phi0, phi1, phi2), bes_rng_dev: [3][n_lt][2]
 * x ranges, bes_l_dev: int32[n_lt] multipoles; ell_idx_dev: int32[n_ell]
 * rows of the Bessel tables to use.  precision: 0 = FP64, 1 = FP32, 2 = source
 * and Bessel operands rounded to bfloat16 with FP32 accumulation.
 * cl_dev out: [B][3][n_ell] (TT, TE, EE); delta_dev (optional) out:
 * [B][4][n_ell][nkt].
 */
size_t abcmb_los_workspace_size(int n_cosmo, int n_ell, int nkt);
int abcmb_los_cl(const AbcmbCosmo* cosmos_dev, int n_cosmo, const double* src_dev,
                 int n_tau, const double* bgv_dev, const double* kt_dev, int nkt,
                 const double* bes_dev, const double* bes_rng_dev,
                 const int32_t* bes_l_dev, int n_lt, int nx,
                 const int32_t* ell_idx_dev, int n_ell, int precision,
                 double* cl_dev, double* delta_dev,
                 void* workspace, size_t ws_bytes, void* stream);

/*
 * Tangent-linear line-of-sight C_l: the sampled-l TT, TE, EE of abcmb_los_cl
 * and their derivatives along n_dir (= ABCMB_TANGENT_DIRS) directions
 * of the sources and of the conformal distance chi_i = tau0 - tau_i at the
 * LOS nodes, plus the derivatives with respect to ln A_s and n_s.  The
 * Bessel lookups are differentiated exactly (slope of the interpolated
 * segment, derivative of the asymptotic form).  Replaces jax.jacfwd through
 * SpectrumSolver.Cl_one_ell (abcmb/spectrum.py:616-797).
 * dsrc_dev: [B][n_dir][4][n_tau][nkt] derivatives of src_dev; dchi_dev:
 * [B][n_dir][n_tau] derivatives of chi; cl_dev out: [B][1 + n_dir + 2][3][n_ell]
 * (value, the n_dir directions, ln A_s, n_s; each TT, TE, EE).  precision:
 * 0 FP64; 1 FP32 (Bessel lookups and the contraction over tau in FP32,
 * the k-quadrature in FP64); 2 the same with the contraction's operands
 * rounded to bfloat16 (abcmb_los_cl's variants).
 */
size_t abcmb_los_tangent_workspace_size(int n_cosmo, int n_dir, int n_ell, int nkt);
int abcmb_los_tangent(const AbcmbCosmo* cosmos_dev, int n_cosmo, const double* src_dev,
                      const double* dsrc_dev, const double* dchi_dev, int n_dir, int n_tau,
                      const double* bgv_dev, const double* kt_dev, int nkt,
                      const double* bes_dev, const double* bes_rng_dev,
                      const int32_t* bes_l_dev, int n_lt, int nx,
                      const int32_t* ell_idx_dev, int n_ell, int precision, double* cl_dev,
                      void* workspace, size_t ws_bytes, void* stream);

/*
 * Reverse mode of abcmb_sources + abcmb_los_cl (FP64) and of the C_l
 * k-quadrature: from the cotangents cl_bar [B][3][n_ell] of the sampled TT,
 * TE, EE, the cotangents of the ten LOS fields of every perturbation table
 * (added into pt_bar_dev[c], [ABCMB_N_PT_FIELDS][n_save][nk] device pointers;
 * the caller zeroes them), of the LOS background vectors bgv_bar
 * [B][6][n_save-1] (overwritten; same rows as bgv_dev) and scal_bar [B][3] =
 * (A_s, n_s, tau0) (overwritten).  Replaces reverse-mode differentiation
 * through SpectrumSolver.Cl_one_ell (abcmb/spectrum.py:616-797).  The
 * forward sources and transfer functions are recomputed inside; the Bessel
 * lookups are differentiated as abcmb_los_tangent does.  Every cosmology
 * needs nk >= 4.
 */
size_t abcmb_los_adjoint_workspace_size(int n_cosmo, int nk_max, int n_save, int n_ell, int nkt);
int abcmb_los_adjoint(const AbcmbCosmo* cosmos_dev, int n_cosmo, int nk_max, int n_save,
                      const double* bgv_dev, const double* kt_dev, int nkt, const double* scales4,
                      const double* bes_dev, const double* bes_rng_dev, const int32_t* bes_l_dev, int n_lt,
                      int nx, const int32_t* ell_idx_dev, int n_ell, const double* cl_bar_dev,
                      double* const* pt_bar_dev, double* bgv_bar_dev, double* scal_bar_dev,
                      void* workspace, size_t ws_bytes, void* stream);

/* Dense linear map out[m][r] = sum_c op[r][c] in[m][c] (the l-spline of
 * abcmb/spectrum.py:598-601 applied as a precomputed operator). */
int abcmb_apply_operator(const double* op_dev, int rows, int cols,
                         const double* in_dev, int n_vec, double* out_dev,
                         void* stream);

/* Limber lensing potential C_l^phiphi (abcmb/spectrum.py:349-435).
 * lens_dev: [B][4][n_lna] = ln a axis (lna_rec .. 0), aH and tau on it, and
 * chi(lna_rec) at [3][0]; ells_dev: [n_l]; clpp_dev out: [B][n_l].  Of each
 * cosmology's perturbation table only field ABCMB_PT_DELTA_M (the first, [0])
 * is read, so pt may point to a table holding that field alone. */
int abcmb_lensing_clpp(const AbcmbCosmo* cosmos_dev, int n_cosmo,
                       const double* lens_dev, int n_lna,
                       const double* ells_dev, int n_l, double* clpp_dev,
                       void* stream);

/* Correlation-function lensing of TT/TE/EE (+BB) (abcmb/spectrum.py:437-572,
 * Wigner-d recursions abcmb/ABCMBTools.py:16-185).  cl_in: [B][3][n_l]
 * unlensed on l = 2..n_l+1; clpp: [B][n_l]; mu, w: Gauss-Legendre nodes with
 * mu = 1 (weight 0) appended; out: [B][4][n_l] lensed TT, TE, EE, BB. */
size_t abcmb_lens_workspace_size(int n_cosmo, int n_mu, int n_l);
int abcmb_lens(int n_cosmo, const double* cl_in_dev, const double* clpp_dev, int n_l,
               const double* mu_dev, const double* w_dev, int n_mu,
               double* out_dev, void* workspace, size_t ws_bytes, void* stream);

/*
 * Correlation-function lensing with its two contractions over shared operands
 * on the matrix cores: the Cgl pass (abcmb/spectrum.py:498-503) and the
 * back-transform (:564-570) as batched GEMMs across the n_cosmo cosmologies
 * of the call, against reduced Wigner d-matrices built once by
 * abcmb_lens_dmat (cosmology independent, [6][n_mu][n_l] plus the recursion
 * tables).  precision 0: FP64 operands and MFMA (the results of abcmb_lens to
 * rounding); 1: FP32 operands, FP32 MFMA accumulation; 2: BF16 operands, FP32
 * accumulation.  The correlation pass between them stays FP64 on the vector
 * units.  Arguments and outputs are those of abcmb_lens; the d-matrices must
 * be built for the same mu nodes, n_l and precision.
 */
/*
 * Reverse mode of abcmb_lens (vector or matrix-core path, FP64): from the
 * cotangents lensed_bar [B][4][n_l] of the lensed TT, TE, EE, BB, the
 * cotangents of the unlensed spectra cl_bar [B][3][n_l] and of C_l^phiphi
 * clpp_bar [B][n_l] (both overwritten).  Replaces reverse-mode
 * differentiation through SpectrumSolver.lensed_Cls (abcmb/spectrum.py:437-572).
 * The back-transform and the Cgl pass are transposed as contractions over the
 * nodes; the correlation pass is linear in the unlensed spectra and
 * differentiated in sigma^2(mu) and Cgl2(mu) with dual numbers.
 */
size_t abcmb_lens_adjoint_workspace_size(int n_cosmo, int n_mu, int n_l);
int abcmb_lens_adjoint(int n_cosmo, const double* cl_in_dev, const double* clpp_dev, int n_l,
                       const double* mu_dev, const double* w_dev, int n_mu, const double* lensed_bar_dev,
                       double* cl_bar_dev, double* clpp_bar_dev, void* workspace, size_t ws_bytes,
                       void* stream);

/*
 * Reverse mode of abcmb_lensing_clpp with respect to the perturbation table:
 * clpp_bar [B][n_l] into the ABCMB_PT_DELTA_M field of the table cotangents
 * pt_bar_dev[c] ([ABCMB_N_PT_FIELDS][n_save][nk], device pointers, added to:
 * the caller zeroes them).  Replaces the table part of reverse-mode
 * differentiation through SpectrumSolver.Cl_phiphi (abcmb/spectrum.py:349-435);
 * the background inputs (lens_dev, the cosmology fields) are held fixed.
 */
int abcmb_lensing_clpp_adjoint(const AbcmbCosmo* cosmos_dev, int n_cosmo, const double* lens_dev,
                               int n_lna, const double* ells_dev, int n_l, const double* clpp_bar_dev,
                               double* const* pt_bar_dev, void* stream);

size_t abcmb_lens_dmat_size(int n_mu, int n_l, int precision);
int abcmb_lens_dmat(const double* mu_dev, int n_mu, int n_l, int precision, void* dmat_dev,
                    size_t dmat_bytes, void* stream);
size_t abcmb_lens_mfma_workspace_size(int n_cosmo, int n_mu, int n_l, int precision);
int abcmb_lens_mfma(int n_cosmo, const double* cl_in_dev, const double* clpp_dev, int n_l,
                    const double* mu_dev, const double* w_dev, int n_mu, const void* dmat_dev,
                    int precision, double* out_dev, void* workspace, size_t ws_bytes, void* stream);

/*
 * Reionization on the GPU (Background reionization model and its optical-depth
 * root, abcmb/background.py:939-1017; the reference solves z_reion with
 * optimistix Newton, :1008-1017).  For each cosmology whose x_e table
 * (cosmos_dev[i].xe_tab) holds the recombination history alone
 * (abcmb_host_recombination), solve tau_reion(z_reion) = reion_dev[i].tau_reion
 * on the 2000-point trapezoid over ln a in [-5, 0], and write the table with
 * the hydrogen and helium reionization added (xe_out_dev [n_cosmo][xe_n]: the
 * entries up to xe_lna_last (at most n_rec), the padding copying the last of
 * them) and z_reion
 * (z_out_dev [n_cosmo]).  With tangents_dev (the cosmologies' AbcmbTangent
 * inputs, x_e tables of the recombination history), also the implicit-function
 * derivatives dz_reion/dtheta = (dtau_target/dtheta - dtau/dtheta|_z) /
 * (dtau/dz) (dz_out_dev [n_cosmo][ABCMB_TANGENT_DIRS]; dtarget_dev
 * [n_cosmo][ABCMB_TANGENT_DIRS] holds dtau_target along each direction) and the
 * tangents of the finished table (dxe_out_dev [n_cosmo][ABCMB_TANGENT_DIRS][xe_n]).
 * Every cosmology's xe_n equals the first's (the row stride of the outputs)
 * and is at least n_rec.  status_dev [n_cosmo]: 0, or ABCMB_EINVAL if no root
 * in z in [0, 50] or the sizes break that rule.
 */
int abcmb_reionize(const AbcmbCosmo* cosmos_dev, const AbcmbTangent* tangents_dev, int n_cosmo,
                   const AbcmbReion* reion_dev, const double* dtarget_dev, int n_rec,
                   double* xe_out_dev, double* dxe_out_dev, double* z_out_dev, double* dz_out_dev,
                   int32_t* status_dev, void* stream);

/*
 * The six LOS background vectors on the GPU (what abcmb_host_los_background
 * returns; abcmb/spectrum.py:640-651 with kappa by abcmb/background.py:682-742):
 * out_dev [n_cosmo][6][n] = tau, g, dg/dlna, aH, exp(-kappa), aH' * aH on each
 * cosmology's grid lna_dev [n_cosmo][n], kappa(lna) = int_lna^0 dl / (tau_c aH)
 * by 4-point Gauss-Legendre on the 20000 panels of [-10, 0].  With
 * tangents_dev, also their derivatives dout_dev
 * [n_cosmo][ABCMB_TANGENT_DIRS][6][n].  The x_e table may be the output of
 * abcmb_reionize (a cosmology copied with xe_tab, xe_last and z_reion set).
 */
int abcmb_los_background(const AbcmbCosmo* cosmos_dev, const AbcmbTangent* tangents_dev, int n_cosmo,
                         const double* lna_dev, int n, double* out_dev, double* dout_dev, void* stream);

/* Linear matter power spectrum at z = 0 (abcmb/spectrum.py:268-302):
 * pk_dev out [B][nq] on the query grid kq_dev [nq]. */
int abcmb_pk_lin(const AbcmbCosmo* cosmos_dev, int n_cosmo, const double* kq_dev, int nq,
                 double* pk_dev, void* stream);

/* Linear power spectrum at redshift z of the total matter (kind
 * ABCMB_PK_M: SpectrumSolver.Pk_lin, abcmb/spectrum.py:268-302) or of CDM +
 * baryons (ABCMB_PK_CB: SpectrumSolver.Pk_cb, abcmb/spectrum.py:304-347,
 * normalised by omega_m as the reference does): delta interpolated linearly in
 * ln a over the saved points, then linearly in k, clamped at both ends. */
#define ABCMB_PK_M 0
#define ABCMB_PK_CB 1
int abcmb_pk(const AbcmbCosmo* cosmos_dev, int n_cosmo, const double* kq_dev, int nq,
             double z, int kind, double* pk_dev, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* ABCMB_H */
