// User-fluid hook of the ODE kernels (included by pert_core.h inside
// namespace abcmb::<anonymous>) for a build with the standard species only:
// no rows, no background terms.  abcmb_amd/fluidgen.py writes the same
// interface for a model's Fluid classes (abcmb/species.py:11-249).
#pragma once

namespace fluid {
constexpr bool kActive = false;  // any user fluid (background or rows)
constexpr int kRows = 0;         // rows at the end of the state
constexpr int kSlots = 0;        // background scalar slots of those rows
constexpr const char* kSignature = "";

// summed density, pressure and matter density of the user fluids at ln a (eV cm^-3)
template <int N, class Cos>
__device__ __forceinline__ void bg(const Cos&, double, double, D<N>& rho, D<N>& pres, D<N>& rho_m) {
  rho = D<N>(0.0); pres = D<N>(0.0); rho_m = D<N>(0.0);
}
// the slot values at ln a for mode k (aH, and the h' / eta' weight bases kh, ke)
template <int N, class Cos>
__device__ __forceinline__ void slots(const Cos&, double, double, double, const D<N>&, const D<N>&, const D<N>&,
                                      D<N>*) {}
// terms of user row j: add(local column, static coefficient, slot or -1), the
// row's static h' / eta' coefficients and the slots of its h' / eta' weights
template <class Add>
__device__ __forceinline__ void row(int, double, Add, double&, double&, int&, int&) {}
// initial condition of user row j at conformal time tau
template <int N, class Cos>
__device__ __forceinline__ D<N> ic(int, const Cos&, double, const D<N>&) { return D<N>(0.0); }
// the user rows' terms of the metric and matter sums at a saved point
template <int N, class Cos, class Rows>
__device__ __forceinline__ void sums(const Cos&, double, double, const Rows&, D<N>&, D<N>&, D<N>&, D<N>&, D<N>&) {}
}  // namespace fluid
