// HyRex recombination (abcmb/hyrex/*.py) for the host producer (hyrex.cpp).
#pragma once
#include <cstdint>

#include "../../include/abcmb.h"
#include "host_dual.h"

namespace hyrex {

// values with their derivatives along the tangent directions
using Tan = abcmb_ad::Dual<ABCMB_TANGENT_DIRS>;

constexpr double Z0 = 8000.0;    // the grid starts at ln a = -ln(1 + Z0)   (hyrex.py:115-139)
constexpr double DLNA = 5.0e-4;  // and steps by DLNA
constexpr int N_GRID = 17975;    // jnp.arange(-ln 8001, 0, 5e-4).size: the inputs' grid
// allocated lengths of the reference's concatenated arrays (array_with_padding.py):
// x_e: HeIII Saha on z >= 3601 (1597) + post-Saha HeII (4000) + HeII ODE (4096)
//      + post-Saha H (800) + two-photon EMLA (2048) + EMLA with T_m (8192) + TLA (8192)
// T_m: EMLA with T_m (8192) + TLA (8192)
constexpr int XE_ALLOC = 1597 + 4000 + 4096 + 800 + 2048 + 8192 + 8192;
constexpr int TM_ALLOC = 8192 + 8192;
constexpr int N_STAGES = 7;      // the stage ends recorded in / forced by the switch array

struct History {
  int n_xe;       // x_e points written (from the grid start)
  int n_tm;       // T_m points written (from tm_x0)
  double tm_x0;   // ln a of the first T_m point
};

// The history on the grid for helium mass fraction YHe, omega_b, omega_cdm and
// N_eff, from T, n_H and H sampled on the N_GRID grid points (RecombInputs).
// xe [xe_alloc], tm [tm_alloc] are filled, their tails padded with the last
// values.  sw [N_STAGES]: per stage, the grid count after it (< 0 on input:
// the stage's own stopping rule decides; >= 0: forced, for derivatives with
// the grids held).  S: double, or Tan for values with their derivatives
// (then every stage end must be forced).
template <class S>
int recombination(S YHe, S omega_b, S omega_cdm, S neff, const S* T, const S* nh, const S* H, int n_grid,
                  S* xe, int xe_alloc, S* tm, int tm_alloc, int32_t* sw, History& out);

}  // namespace hyrex
