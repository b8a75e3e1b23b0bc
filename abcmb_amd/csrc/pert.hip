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
#include "pert_core.h"

namespace abcmb {
namespace {

// The value-only kernel with 32-row blocks fits three waves per SIMD; the
// tangent kernel and 64-row blocks need the whole register file of one wave.
// PERSIST = false: one workgroup per item (the hardware dispatches them in
// item order, longest first).  PERSIST = true: a grid of resident workgroups
// takes items from a device work queue (an atomic counter over the item list,
// zeroed before the launch) until it is empty; every workgroup leaves when the
// counter passes n_items, so the grid drains without any wait.
template <int NT, int BMAX, int N, bool PERSIST>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu((N == 0 && BMAX <= 32) ? 3 : 1)))
pert_fwd_kernel(const AbcmbCosmo* __restrict__ cosmos, const AbcmbTangent* __restrict__ tangents,
                const AbcmbPertItem* __restrict__ items, AbcmbOdeSettings st, const Bg<N>* __restrict__ bgsave,
                int n_save_max, int32_t* __restrict__ status_out, int32_t* __restrict__ steps_out,
                int32_t* __restrict__ iters_out, double* __restrict__ state_out, int n_items,
                int32_t* __restrict__ queue) {
  __shared__ Smem<NT, BMAX, N> sm;
  if constexpr (!PERSIST) {
    ode_item<NT, BMAX, N>(blockIdx.x, sm, cosmos, tangents, items, st, bgsave, n_save_max, status_out, steps_out,
                          iters_out, state_out);
  } else {
    __shared__ int next;
    for (;;) {
      __syncthreads();  // every thread has read the previous item and left its LDS work
      if (threadIdx.x == 0) next = atomicAdd(queue, 1);
      __syncthreads();
      const int item = next;
      if (item >= n_items) break;
      ode_item<NT, BMAX, N>(item, sm, cosmos, tangents, items, st, bgsave, n_save_max, status_out, steps_out,
                            iters_out, state_out);
    }
  }
}

template <int NT, int BMAX, int N>
int launch(const AbcmbCosmo* cosmos, const AbcmbTangent* tangents, int n_cosmo, const AbcmbPertItem* items,
           int n_items, const AbcmbOdeSettings& st, const Bg<N>* bgsave, int n_save_max, int32_t* cnt,
           int32_t* queue, int32_t* status, int32_t* steps, int32_t* iters, double* state, hipStream_t stream) {
  const int sched = tune_get(ABCMB_TUNE_PERT_PERSISTENT);
  if (sched > 0) {
    // resident workgroups: the occupancy of this instance times the CUs, times
    // the tuned fraction (percent) of that
    int dev = 0, per_cu = 0;
    ABCMB_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    ABCMB_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    ABCMB_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pert_fwd_kernel<NT, BMAX, N, true>, NT, 0));
    long long g = (long long)(per_cu > 0 ? per_cu : 1) * prop.multiProcessorCount * sched / 100;
    if (g < 1) g = 1;
    if (g > n_items) g = n_items;
    hipLaunchKernelGGL(zero_counts_kernel, dim3(1), dim3(64), 0, stream, queue, 1);
    hipLaunchKernelGGL((pert_fwd_kernel<NT, BMAX, N, true>), dim3((unsigned)g), dim3(NT), 0, stream, cosmos,
                       tangents, items, st, bgsave, n_save_max, status, steps, iters, state, n_items, queue);
  } else {
    hipLaunchKernelGGL((pert_fwd_kernel<NT, BMAX, N, false>), dim3(n_items), dim3(NT), 0, stream, cosmos, tangents,
                       items, st, bgsave, n_save_max, status, steps, iters, state, n_items, queue);
  }
  ABCMB_HIP_CHECK(hipGetLastError());
  if (12 + 3 * st.n_q <= ABCMB_N_PT_FIELDS) {  // the kernel stored rows: turn them into fields
    hipLaunchKernelGGL(zero_counts_kernel, dim3((n_cosmo + 255) / 256), dim3(256), 0, stream, cnt, n_cosmo);
    hipLaunchKernelGGL(count_items_kernel, dim3((n_items + 255) / 256), dim3(256), 0, stream, items, n_items, status,
                       cnt);
    hipLaunchKernelGGL((fields_cosmo_kernel<N>), dim3((unsigned)n_cosmo * n_save_max), dim3(64), 0, stream, cosmos,
                       tangents, st, bgsave, n_save_max, cnt);
    const long long nthr = (long long)n_items * n_save_max;
    hipLaunchKernelGGL((fields_kernel<N>), dim3((unsigned)((nthr + 63) / 64)), dim3(64), 0, stream, cosmos, tangents,
                       items, n_items, st, bgsave, n_save_max, status, cnt);
    ABCMB_HIP_CHECK(hipGetLastError());
  }
  return ABCMB_OK;
}

// the workspace: the saved-point background table, then at the next 256-byte
// boundary the work-queue counter (256 bytes), then one item count per cosmology
size_t bg_table_bytes(size_t per, int n_cosmo, int n_save_max) {
  const size_t b = per * (size_t)(n_cosmo > 0 ? n_cosmo : 0) * (size_t)(n_save_max > 0 ? n_save_max : 0);
  return (b + 255) / 256 * 256;
}
constexpr size_t QUEUE_BYTES = 256;

template <int N>
int dispatch(const AbcmbCosmo* cosmos, const AbcmbTangent* tangents, int n_cosmo, const AbcmbPertItem* items,
             int n_items, const AbcmbOdeSettings& st, int maxb, int n_save_max, void* workspace, int32_t* status,
             int32_t* steps, int32_t* iters, double* state, hipStream_t s) {
  Bg<N>* bgs = (Bg<N>*)workspace;
  int32_t* queue = (int32_t*)((char*)workspace + bg_table_bytes(sizeof(Bg<N>), n_cosmo, n_save_max));
  int32_t* cnt = (int32_t*)((char*)queue + QUEUE_BYTES);
  hipLaunchKernelGGL((bg_save_kernel<N>), dim3((n_save_max + 63) / 64, n_cosmo), dim3(64), 0, s, cosmos, tangents,
                     n_save_max, bgs);
  ABCMB_HIP_CHECK(hipGetLastError());
#define ABCMB_LAUNCH(NT_, BM_) \
  return launch<NT_, BM_, N>(cosmos, tangents, n_cosmo, items, n_items, st, bgs, n_save_max, cnt, queue, status, \
                             steps, iters, state, s)
  // 26 rows: the photon block of the default cut-offs (l_max_g 12, l_max_pol_g 10)
  if (maxb <= 26) {
    if (st.n <= 64) ABCMB_LAUNCH(64, 26);
    if (st.n <= 128) ABCMB_LAUNCH(128, 26);
    ABCMB_LAUNCH(192, 26);
  }
  if (maxb <= 32) {
    if (st.n <= 64) ABCMB_LAUNCH(64, 32);
    if (st.n <= 128) ABCMB_LAUNCH(128, 32);
    ABCMB_LAUNCH(192, 32);
  }
  if (st.n <= 64) ABCMB_LAUNCH(64, 64);
  if (st.n <= 128) ABCMB_LAUNCH(128, 64);
  ABCMB_LAUNCH(192, 64);
#undef ABCMB_LAUNCH
}

}  // namespace
}  // namespace abcmb

extern "C" const char* abcmb_fluid_signature(void) { return abcmb::fluid::kSignature; }

extern "C" size_t abcmb_pert_workspace_size(int n_cosmo, int n_save_max, int n_dir) {
  using namespace abcmb;
  const size_t per = n_dir > 0 ? sizeof(Bg<ABCMB_TANGENT_DIRS>) : sizeof(Bg<0>);
  return bg_table_bytes(per, n_cosmo, n_save_max) + QUEUE_BYTES + sizeof(int32_t) * (size_t)(n_cosmo > 0 ? n_cosmo : 0);
}

extern "C" int abcmb_pert_fwd(const AbcmbCosmo* cosmos_dev, int n_cosmo, const AbcmbPertItem* items_dev,
                              int n_items, const AbcmbOdeSettings* settings, int n_save_max, int32_t* status_dev,
                              int32_t* steps_dev, int32_t* iters_dev, double* state_dev, void* workspace,
                              size_t ws_bytes, void* stream) {
  using namespace abcmb;
  ABCMB_REQUIRE(cosmos_dev && items_dev && settings && status_dev && steps_dev, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && n_items > 0 && n_save_max > 0, "empty launch");
  ABCMB_REQUIRE(workspace && ws_bytes >= abcmb_pert_workspace_size(n_cosmo, n_save_max, 0), "workspace too small");
  int maxb = 0;
  int rc = check_settings(*settings, maxb);
  if (rc != ABCMB_OK) return rc;
  return dispatch<0>(cosmos_dev, nullptr, n_cosmo, items_dev, n_items, *settings, maxb, n_save_max, workspace,
                     status_dev, steps_dev, iters_dev, state_dev, (hipStream_t)stream);
}

extern "C" int abcmb_pert_tangent(const AbcmbCosmo* cosmos_dev, const AbcmbTangent* tangents_dev, int n_cosmo,
                                  const AbcmbPertItem* items_dev, int n_items, const AbcmbOdeSettings* settings,
                                  int n_save_max, int32_t* status_dev, int32_t* steps_dev, int32_t* iters_dev,
                                  void* workspace, size_t ws_bytes, void* stream) {
  using namespace abcmb;
  ABCMB_REQUIRE(cosmos_dev && tangents_dev && items_dev && settings && status_dev && steps_dev, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && n_items > 0 && n_save_max > 0, "empty launch");
  ABCMB_REQUIRE(workspace && ws_bytes >= abcmb_pert_workspace_size(n_cosmo, n_save_max, ABCMB_TANGENT_DIRS),
                "workspace too small");
  int maxb = 0;
  int rc = check_settings(*settings, maxb);
  if (rc != ABCMB_OK) return rc;
  return dispatch<ABCMB_TANGENT_DIRS>(cosmos_dev, tangents_dev, n_cosmo, items_dev, n_items, *settings, maxb,
                                      n_save_max, workspace, status_dev, steps_dev, iters_dev, nullptr,
                                      (hipStream_t)stream);
}

// ---------------------------------------------------------------- reionization on the device
// tanh reionization (abcmb/background.py:939-1017) of a cosmology whose x_e
// table holds the recombination history alone: the reionization redshift is
// solved on the device for the target optical depth (safeguarded Newton on the
// 2000-point trapezoid over ln a in [-5, 0] that the host producer uses), the
// hydrogen and helium reionization fractions are added to the table, and with
// tangents the derivative of z_reion follows from the implicit-function
// theorem, dz/dtheta = (dtau_target/dtheta - dtau/dtheta|_z) / (dtau/dz).
namespace abcmb {
namespace {

constexpr int REION_GRID = 2000;
constexpr int REION_THREADS = 256;

template <int N>
__device__ __forceinline__ D<N> xe_reion(double lna, const D<N>& zre, const AbcmbReion& R, const D<N>& YHe) {
  const D<N> fHe = YHe * dinv(4.0 * (1.0 - YHe));
  const double z = exp(-lna) - 1.0;
  const double y = pow(1.0 + z, R.exp_reion);
  const D<N> yre = dpow(1.0 + zre, R.exp_reion);
  const D<N> dy = R.exp_reion * dpow(1.0 + zre, R.exp_reion - 1.0) * R.Delta_z_reion;
  const D<N> xH = 0.5 * (1.0 + fHe) * (1.0 + dtanh((yre - y) * dinv(dy)));
  const D<N> xHe = 0.5 * fHe * (1.0 + tanh((R.z_reion_He - z) / R.Delta_z_reion_He));
  return xH + xHe;
}

// optical depth of the reionization fractions and its z derivative (value only,
// or with the parameter tangents at fixed z when N > 0): a workgroup sum
template <int N>
__device__ void reion_tau(const Cosmo<N>& P, const AbcmbReion& R, double zre, D<N>& tau, double& dtau_dz,
                          double* red) {
  constexpr double K_NH = 3.0 * H0_OVER_H * H0_OVER_H / 8 / PI / GN / MH;
  constexpr double K_G = SIGMA_T * C_CM / C_MPC;
  const D<N> YHe = P(F_YHE);
  const D<N> fHe = YHe * dinv(4.0 * (1.0 - YHe));
  const double yre = pow(1.0 + zre, R.exp_reion);
  const double dy = R.exp_reion * pow(1.0 + zre, R.exp_reion - 1.0) * R.Delta_z_reion;
  // d(yre)/dz and d(dy)/dz for the z derivative of the tanh argument
  const double dyre = R.exp_reion * pow(1.0 + zre, R.exp_reion - 1.0);
  const double ddy = R.exp_reion * (R.exp_reion - 1.0) * pow(1.0 + zre, R.exp_reion - 2.0) * R.Delta_z_reion;
  D<N> s(0.0);
  double sz = 0.0;
  for (int i = threadIdx.x; i < REION_GRID; i += blockDim.x) {
    const double l = -5.0 + 5.0 * i / (REION_GRID - 1);
    const double w = (i == 0 || i == REION_GRID - 1 ? 0.5 : 1.0) * (5.0 / (REION_GRID - 1));
    Bg<N> b;
    bg_rho(P, l, b);
    const double a = b.a, z = 1.0 / a - 1.0;
    const double y = pow(1.0 + z, R.exp_reion);
    const D<N> nh = (1 - YHe) * P(F_OMEGA_B) * (K_NH / (a * a * a));
    const double u = (yre - y) / dy;
    const double th = tanh(u);
    const D<N> xH = 0.5 * (1.0 + fHe) * (1.0 + th);
    const D<N> xHe = 0.5 * fHe * (1.0 + tanh((R.z_reion_He - z) / R.Delta_z_reion_He));
    const D<N> g = (a * K_G) * nh * dinv(b.aH);
    s += w * (g * (xH + xHe));
    const double du = (dyre * dy - (yre - y) * ddy) / (dy * dy);
    sz += w * g.v * 0.5 * (1.0 + fHe.v) * (1.0 - th * th) * du;
  }
  // workgroup sums of s (1 + N values) and sz
  double v[2 + N];
  v[0] = s.v;
  v[1] = sz;
  if constexpr (N > 0) { ABCMB_DLOOP(N) v[2 + i_] = s.d[i_]; }
  for (int q = 0; q < 2 + N; ++q) {
    double x = v[q];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if ((threadIdx.x & 63) == 0) red[(threadIdx.x >> 6) * (2 + N) + q] = x;
  }
  __syncthreads();
  for (int q = 0; q < 2 + N; ++q) {
    double x = 0.0;
    for (int w8 = 0; w8 < REION_THREADS / 64; ++w8) x += red[w8 * (2 + N) + q];
    v[q] = x;
  }
  __syncthreads();
  tau = D<N>(v[0]);
  if constexpr (N > 0) { ABCMB_DLOOP(N) tau.d[i_] = v[2 + i_]; }
  dtau_dz = v[1];
}

template <int N>
__global__ void __launch_bounds__(REION_THREADS) reion_kernel(const AbcmbCosmo* __restrict__ cosmos,
                                                             const AbcmbTangent* __restrict__ tangents,
                                                             const AbcmbReion* __restrict__ reion,
                                                             const double* __restrict__ dtarget, int n_rec,
                                                             double* __restrict__ xe_out, double* __restrict__ dxe_out,
                                                             double* __restrict__ z_out, double* __restrict__ dz_out,
                                                             int32_t* __restrict__ status) {
  __shared__ double red[(REION_THREADS / 64) * (2 + ABCMB_TANGENT_DIRS)];
  const int ci = blockIdx.x;
  const AbcmbCosmo& c = cosmos[ci];
  const AbcmbReion R = reion[ci];
  const AbcmbTangent* T = nullptr;
  if constexpr (N > 0) T = tangents + ci;
  // every cosmology's table has the stride of the first, and the recombination
  // points fit in it (uniform across the workgroup: no barrier is skipped by some threads only)
  if (c.xe_n != cosmos[0].xe_n || n_rec > c.xe_n) {
    if (threadIdx.x == 0) status[ci] = ABCMB_EINVAL;
    return;
  }
  const Cosmo<0> P0{c, nullptr};
  // bracket by bisection on [0, 50], then Newton kept inside the bracket
  double lo = 0.0, hi = 50.0, z = 0.0, f = 0.0, fz = 0.0;
  D<0> t0;
  reion_tau<0>(P0, R, lo, t0, fz, red);
  const double flo = t0.v - R.tau_reion;
  bool ok = flo < 0.0;
  for (int it = 0; it < 8; ++it) {
    const double mid = 0.5 * (lo + hi);
    reion_tau<0>(P0, R, mid, t0, fz, red);
    if (t0.v - R.tau_reion < 0.0) lo = mid; else hi = mid;
  }
  z = 0.5 * (lo + hi);
  int it = 0;
  for (; it < 100; ++it) {
    reion_tau<0>(P0, R, z, t0, fz, red);
    f = t0.v - R.tau_reion;
    if (f < 0.0) lo = z; else hi = z;
    double zn = fz > 0.0 ? z - f / fz : 0.5 * (lo + hi);
    if (!(zn > lo && zn < hi)) zn = 0.5 * (lo + hi);
    const double dz = fabs(zn - z);
    z = zn;
    if (dz <= 1e-13 * fmax(1.0, z) || hi - lo <= 1e-13 * fmax(1.0, z)) break;
  }
  ok = ok && it < 100 && isfinite(z);
  // implicit-function derivatives at the root
  D<N> zd(z);
  if constexpr (N > 0) {
    const Cosmo<N> P{c, T};
    D<N> tau;
    double dtdz;
    reion_tau<N>(P, R, z, tau, dtdz, red);
    ABCMB_DLOOP(N) zd.d[i_] = (dtarget[(size_t)ci * N + i_] - tau.d[i_]) / dtdz;
    if (threadIdx.x == 0) { ABCMB_DLOOP(N) dz_out[(size_t)ci * N + i_] = zd.d[i_]; }
  }
  if (threadIdx.x == 0) {
    z_out[ci] = z;
    status[ci] = ok ? ABCMB_OK : ABCMB_EINVAL;
  }
  // the table: reionization added on the recombination grid's valid points
  // (through xe_lna_last, at most n_rec), the padding copies the last of them
  const int n = c.xe_n;
  const Cosmo<N> P{c, T};
  const D<N> YHe = P(F_YHE);
  double* xo = xe_out + (size_t)ci * n;
  const int nv = max(1, min(n_rec, (int)llround((c.xe_lna_last - c.xe_x0) / c.xe_dx) + 1));
  for (int j = threadIdx.x; j < nv; j += blockDim.x) {
    const double l = c.xe_x0 + j * c.xe_dx;
    const D<N> x = tab_at<N>(c.xe_tab, P.dxe(), n, j) + xe_reion<N>(l, zd, R, YHe);
    xo[j] = x.v;
    if constexpr (N > 0) { ABCMB_DLOOP(N) dxe_out[((size_t)ci * N + i_) * n + j] = x.d[i_]; }
  }
  __syncthreads();
  for (int j = nv + threadIdx.x; j < n; j += blockDim.x) {
    xo[j] = xo[nv - 1];
    if constexpr (N > 0) { ABCMB_DLOOP(N) dxe_out[((size_t)ci * N + i_) * n + j] = dxe_out[((size_t)ci * N + i_) * n + nv - 1]; }
  }
}

}  // namespace
}  // namespace abcmb

extern "C" int abcmb_reionize(const AbcmbCosmo* cosmos_dev, const AbcmbTangent* tangents_dev, int n_cosmo,
                              const AbcmbReion* reion_dev, const double* dtarget_dev, int n_rec, double* xe_out_dev,
                              double* dxe_out_dev, double* z_out_dev, double* dz_out_dev, int32_t* status_dev,
                              void* stream) {
  using namespace abcmb;
  ABCMB_REQUIRE(cosmos_dev && reion_dev && xe_out_dev && z_out_dev && status_dev, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && n_rec > 1, "bad sizes");
  const bool tan = tangents_dev != nullptr;
  ABCMB_REQUIRE(!tan || (dtarget_dev && dxe_out_dev && dz_out_dev), "tangents need dtarget, dxe_out and dz_out");
  hipStream_t s = (hipStream_t)stream;
  if (tan)
    hipLaunchKernelGGL(reion_kernel<ABCMB_TANGENT_DIRS>, dim3(n_cosmo), dim3(REION_THREADS), 0, s, cosmos_dev,
                       tangents_dev, reion_dev, dtarget_dev, n_rec, xe_out_dev, dxe_out_dev, z_out_dev, dz_out_dev,
                       status_dev);
  else
    hipLaunchKernelGGL(reion_kernel<0>, dim3(n_cosmo), dim3(REION_THREADS), 0, s, cosmos_dev, nullptr, reion_dev,
                       nullptr, n_rec, xe_out_dev, nullptr, z_out_dev, nullptr, status_dev);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}

// ---------------------------------------------------------------- LOS background on the device
// The six line-of-sight background vectors (abcmb/spectrum.py:640-651) of a
// cosmology on its LOS grid: tau, g, dg/dlna, aH, exp(-kappa) and aH' aH, with
// the optical depth kappa(lna) = int_lna^0 dl / (tau_c aH)
// (abcmb/background.py:682-742) by 4-point Gauss-Legendre over the 20000
// equal panels of [-10, 0] that the host producer uses.  One workgroup per
// cosmology: each thread sums a run of 40 consecutive panels, a suffix scan over
// the runs gives kappa at every run boundary, and each LOS point adds the
// panels between it and the next boundary.  With tangents, every quantity
// carries its derivatives along the parameter directions.
namespace abcmb {
namespace {

constexpr int KAPPA_NODES = 20001;
constexpr double KAPPA_LO = -10.0;
constexpr int LOSBG_THREADS = 512;
constexpr int KAPPA_RUN = (KAPPA_NODES - 1 + LOSBG_THREADS - 1) / LOSBG_THREADS;
__constant__ double GL4_X[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
__constant__ double GL4_W[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};

__device__ __forceinline__ double kappa_node(int i) {
  return i == KAPPA_NODES - 1 ? 0.0 : KAPPA_LO + (0.0 - KAPPA_LO) * i / (KAPPA_NODES - 1);
}

// the optical-depth rate dkappa/d(-lna) = a n_e sigma_T c / aH = 1 / (tau_c aH)
template <int N>
__device__ __forceinline__ D<N> kappa_rate(const Cosmo<N>& P, double l) {
  constexpr double K_NH = 3.0 * H0_OVER_H * H0_OVER_H / 8 / PI / GN / MH;
  constexpr double K_G = SIGMA_T * C_CM / C_MPC;
  Bg<N> b;
  bg_rho(P, l, b);
  const D<N> nh = (1 - P(F_YHE)) * P(F_OMEGA_B) * (K_NH / (b.a * b.a * b.a));
  return (b.a * K_G) * nh * bg_xe(P, l) * dinv(b.aH);
}

template <int N>
__device__ __forceinline__ D<N> kappa_gl4(const Cosmo<N>& P, double lo, double hi) {
  const double m = 0.5 * (lo + hi), hw = 0.5 * (hi - lo);
  D<N> s(0.0);
#pragma unroll
  for (int q = 0; q < 4; ++q) s += GL4_W[q] * kappa_rate<N>(P, m + hw * GL4_X[q]);
  return s * hw;
}

template <int N>
__global__ void __launch_bounds__(LOSBG_THREADS) losbg_kernel(const AbcmbCosmo* __restrict__ cosmos,
                                                              const AbcmbTangent* __restrict__ tangents,
                                                              const double* __restrict__ lna, int n,
                                                              double* __restrict__ out, double* __restrict__ dout) {
  constexpr int NV = 1 + N;
  constexpr double K_NH = 3.0 * H0_OVER_H * H0_OVER_H / 8 / PI / GN / MH;
  constexpr double K_TC = C_MPC / SIGMA_T / C_CM;
  constexpr double K_RHO = 3.0 * H0_OVER_H * H0_OVER_H / 8.0 / PI / GN;
  constexpr double K_AHP = 4.0 * PI * GN / 3.0 / (C_MPC * C_MPC);
  __shared__ double runs[LOSBG_THREADS * NV];
  __shared__ double total[NV];
  const int ci = blockIdx.x;
  const AbcmbCosmo& c = cosmos[ci];
  const AbcmbTangent* T = nullptr;
  if constexpr (N > 0) T = tangents + ci;
  const Cosmo<N> P{c, T};
  // this thread's run of panels
  {
    const int p0 = min(threadIdx.x * KAPPA_RUN, KAPPA_NODES - 1), p1 = min(p0 + KAPPA_RUN, KAPPA_NODES - 1);
    D<N> s(0.0);
    for (int p = p0; p < p1; ++p) s += kappa_gl4<N>(P, kappa_node(p), kappa_node(p + 1));
    runs[threadIdx.x * NV] = s.v;
    if constexpr (N > 0) { ABCMB_DLOOP(N) runs[threadIdx.x * NV + 1 + i_] = s.d[i_]; }
  }
  __syncthreads();
  // exclusive suffix sums over the runs, one component per thread
  if (threadIdx.x < NV) {
    double acc = 0.0;
    for (int r = LOSBG_THREADS - 1; r >= 0; --r) {
      const double v = runs[r * NV + threadIdx.x];
      runs[r * NV + threadIdx.x] = acc;
      acc += v;
    }
    total[threadIdx.x] = acc;
  }
  __syncthreads();
  const double h = kappa_node(1) - kappa_node(0);
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const double l = lna[(size_t)ci * n + j];
    D<N> kap(0.0);
    if (l <= KAPPA_LO) {
      kap.v = total[0];
      if constexpr (N > 0) { ABCMB_DLOOP(N) kap.d[i_] = total[1 + i_]; }
    } else if (l < 0.0) {
      const int i = min((int)((l - KAPPA_LO) / h), KAPPA_NODES - 2);
      // kappa at node i + 1: the suffix of its run plus the panels to that run's end
      const int r = (i + 1) / KAPPA_RUN;
      kap.v = runs[r * NV];
      if constexpr (N > 0) { ABCMB_DLOOP(N) kap.d[i_] = runs[r * NV + 1 + i_]; }
      const int pend = min((r + 1) * KAPPA_RUN, KAPPA_NODES - 1);
      for (int p = i + 1; p < pend; ++p) kap += kappa_gl4<N>(P, kappa_node(p), kappa_node(p + 1));
      kap += kappa_gl4<N>(P, l, kappa_node(i + 1));
    }
    const D<N> em = l < KAPPA_LO ? D<N>(0.0) : dexp(-kap);
    Bg<N> b;
    bg_rho(P, l, b);
    const double ia = 1.0 / b.a;
    const D<N> xe = bg_xe(P, l);
    const D<N> nh = (1 - P(F_YHE)) * P(F_OMEGA_B) * (K_NH * ia * ia * ia);
    const D<N> tc = dinv(nh * xe) * (ia * K_TC);
    // slope of the x_e interpolant (zero off the table and where the index is clipped)
    D<N> dxe(0.0);
    if (l >= c.xe_x0 && l < c.xe_lna_last) {
      const double hi = c.xe_x0 + c.xe_n * c.xe_dx;
      const double fi = (l - c.xe_x0) / (hi - c.xe_x0) * (c.xe_n - 1);
      if (fi >= 1e-6 && fi <= c.xe_n - 1.0 - 1e-6) {
        const int il = (int)floor(fi), iu = min(il + 1, c.xe_n - 1);
        const double sc = (c.xe_n - 1) / (hi - c.xe_x0);
        dxe = (tab_at<N>(c.xe_tab, P.dxe(), c.xe_n, iu) - tab_at<N>(c.xe_tab, P.dxe(), c.xe_n, il)) * sc;
      }
    }
    const D<N> g = em * dinv(tc);
    const D<N> gp = l < KAPPA_LO ? D<N>(0.0) : g * (dinv(tc * b.aH) - (2.0 - dxe * dinv(xe)));
    const D<N> rde = (c.species_off & ABCMB_NO_LAMBDA) ? D<N>(0.0) : P(F_OMEGA_LAMBDA) * K_RHO;
    D<N> rp3 = -2.0 * rde + b.rc + b.rb + 2.0 * (b.rg + b.ru) + b.rn + 3.0 * b.pn;
    if constexpr (fluid::kActive) {
      D<N> rf, pf, rm;
      fluid::bg<N>(P, l, b.a, rf, pf, rm);
      rp3 += rf + 3.0 * pf;
    }
    const D<N> ahp = -(K_AHP * b.a * b.a) * rp3;
    const D<N> v[6] = {fast_interp<N>(l, ABCMB_TAU_TAB_LO, 0.0, c.tau_tab, P.dtau(), ABCMB_TAU_TAB_N), g, gp, b.aH,
                       em, ahp};
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      out[((size_t)ci * 6 + f) * n + j] = v[f].v;
      if constexpr (N > 0) { ABCMB_DLOOP(N) dout[(((size_t)ci * N + i_) * 6 + f) * n + j] = v[f].d[i_]; }
    }
  }
}

}  // namespace
}  // namespace abcmb

extern "C" int abcmb_los_background(const AbcmbCosmo* cosmos_dev, const AbcmbTangent* tangents_dev, int n_cosmo,
                                    const double* lna_dev, int n, double* out_dev, double* dout_dev, void* stream) {
  using namespace abcmb;
  ABCMB_REQUIRE(cosmos_dev && lna_dev && out_dev, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && n > 0, "bad sizes");
  ABCMB_REQUIRE(!tangents_dev || dout_dev, "tangents need dout");
  hipStream_t s = (hipStream_t)stream;
  if (tangents_dev)
    hipLaunchKernelGGL(losbg_kernel<ABCMB_TANGENT_DIRS>, dim3(n_cosmo), dim3(LOSBG_THREADS), 0, s, cosmos_dev,
                       tangents_dev, lna_dev, n, out_dev, dout_dev);
  else
    hipLaunchKernelGGL(losbg_kernel<0>, dim3(n_cosmo), dim3(LOSBG_THREADS), 0, s, cosmos_dev, nullptr, lna_dev, n,
                       out_dev, nullptr);
  ABCMB_HIP_CHECK(hipGetLastError());
  return ABCMB_OK;
}

// ---------------------------------------------------------------- launch without allocation
// A fixed scratch array per device for callers that cannot allocate (graph
// capture, tight serving loops): it holds what abcmb_pert_fwd's workspace
// holds (the saved-point background table and the per-cosmology item counts),
// both fully rewritten by every call before they are read.
namespace abcmb {
static __device__ __align__(256) unsigned char scratch[ABCMB_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
void* g_scratch_ptr[MAX_DEVICES] = {nullptr};
}  // namespace

// The device address of `scratch` on the current device (the array's name is
// a host-side handle, not a device address).  Looked up once per device.
int scratch_address(void** out) {
  int dev = 0;
  ABCMB_HIP_CHECK(hipGetDevice(&dev));
  ABCMB_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "device index %d out of range", dev);
  void* p = __atomic_load_n(&g_scratch_ptr[dev], __ATOMIC_ACQUIRE);
  if (!p) {
    ABCMB_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ABCMB_REQUIRE(p != nullptr, "no device address for the scratch array");
    __atomic_store_n(&g_scratch_ptr[dev], p, __ATOMIC_RELEASE);
  }
  *out = p;
  return ABCMB_OK;
}
}  // namespace abcmb

extern "C" int abcmb_pert_fwd_noalloc(const AbcmbCosmo* cosmos_dev, int n_cosmo, const AbcmbPertItem* items_dev,
                                      int n_items, const AbcmbOdeSettings* settings, int n_save_max,
                                      int32_t* status_dev, int32_t* steps_dev, int32_t* iters_dev,
                                      double* state_dev, void* stream) {
  using namespace abcmb;
  ABCMB_REQUIRE(cosmos_dev && items_dev && settings && status_dev && steps_dev, "null argument");
  ABCMB_REQUIRE(n_cosmo > 0 && n_items > 0 && n_save_max > 0, "empty launch");
  const size_t need = abcmb_pert_workspace_size(n_cosmo, n_save_max, 0);
  ABCMB_REQUIRE(need <= (size_t)ABCMB_NOALLOC_SCRATCH_BYTES,
                "launch needs %zu bytes of scratch (n_cosmo=%d, n_save_max=%d); the fixed scratch holds %d",
                need, n_cosmo, n_save_max, (int)ABCMB_NOALLOC_SCRATCH_BYTES);
  int maxb = 0;
  int rc = check_settings(*settings, maxb);
  if (rc != ABCMB_OK) return rc;
  void* ws = nullptr;
  rc = scratch_address(&ws);
  if (rc != ABCMB_OK) return rc;
  return dispatch<0>(cosmos_dev, nullptr, n_cosmo, items_dev, n_items, *settings, maxb, n_save_max, ws, status_dev,
                     steps_dev, iters_dev, state_dev, (hipStream_t)stream);
}
