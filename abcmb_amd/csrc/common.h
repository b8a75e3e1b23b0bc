// Shared helpers of the HIP library: error reporting and physical constants.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include <string>

#include "abcmb.h"

namespace abcmb {

void set_error(const char* fmt, ...);
// launch-shape settings chosen by the autotuner (abcmb_tune_set); 0 when unset
int tune_get(int key);

#define ABCMB_HIP_CHECK(expr)                                                   \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      ::abcmb::set_error("%s failed: %s", #expr, hipGetErrorString(_e));        \
      return ABCMB_EHIP;                                                        \
    }                                                                           \
  } while (0)

#define ABCMB_REQUIRE(cond, ...)                                                \
  do {                                                                          \
    if (!(cond)) {                                                              \
      ::abcmb::set_error(__VA_ARGS__);                                          \
      return ABCMB_EINVAL;                                                      \
    }                                                                           \
  } while (0)

namespace cst {
constexpr double PI = 3.14159265358979323846;
constexpr double C_CM = 29979245800.0;
constexpr double C_MPC = 9.71561e-15;
constexpr double H0_OVER_H = 3.24078e-18;
constexpr double HBAR = 6.582119569509e-16;
constexpr double GN = 1.18980205e-40;
constexpr double MP = 938271999.0;
constexpr double ME = 510998.9461;
constexpr double MH = MP + ME + 13.5982860719383;
constexpr double MHE = 3.72839e9;
constexpr double SIGMA_T = 6.652458734e-25;
}  // namespace cst

}  // namespace abcmb
