// User-fluid hook of the host background producer for a build with the
// standard species only.  abcmb_amd/fluidgen.py writes the same interface
// (a template over the scalar type) for a model's Fluid classes (abcmb/species.py:11-249).
#pragma once

namespace fluid_host {
constexpr bool kActive = false;
constexpr const char* kSignature = "";
// summed density and pressure of the user fluids at ln a (a = e^lna), eV cm^-3;
// C: AbcmbCosmo or the producer's cosmology over dual numbers, S its scalar
template <class C, class S>
inline void bg(const C&, double, double, S& rho, S& pres) {
  rho = S(0.0);
  pres = S(0.0);
}
}  // namespace fluid_host
