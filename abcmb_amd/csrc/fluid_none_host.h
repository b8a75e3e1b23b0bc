// User-fluid hook of the host background producer for a build with the
// standard species only.  abcmb_amd/fluidgen.py writes the same interface
// (plain double C++) for a model's Fluid classes (abcmb/species.py:11-249).
#pragma once

namespace fluid_host {
constexpr bool kActive = false;
constexpr const char* kSignature = "";
// summed density and pressure of the user fluids at ln a (a = e^lna), eV cm^-3
inline void bg(const AbcmbCosmo&, double, double, double& rho, double& pres) { rho = 0.0; pres = 0.0; }
}  // namespace fluid_host
