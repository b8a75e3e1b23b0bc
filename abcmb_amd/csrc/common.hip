// Error state and device initialisation of libabcmb_hip.so.
#include <atomic>
#include <cstdarg>
#include <cstdio>

#include "common.h"

namespace abcmb {

static thread_local char g_err[512] = {0};

int scratch_address(void** out);  // pert.hip

// Launch-shape settings: one value per key, read by every launch of the
// process; writes are atomic, so tuning from one thread while another
// launches is safe (a launch sees the old or the new value).
static std::atomic<int> g_tune[ABCMB_TUNE_KEYS];

int tune_get(int key) { return key >= 0 && key < ABCMB_TUNE_KEYS ? g_tune[key].load(std::memory_order_relaxed) : 0; }

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace abcmb

extern "C" const char* abcmb_last_error(void) { return abcmb::g_err; }

extern "C" int abcmb_device_init(void) {
  int dev = 0;
  ABCMB_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  ABCMB_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) {
    abcmb::set_error("device %d is %s; this library is built for gfx950 only", dev, prop.gcnArchName);
    return ABCMB_EINVAL;
  }
  void* scratch = nullptr;  // the allocation-free entry point's per-device scratch address
  return abcmb::scratch_address(&scratch);
}

extern "C" int abcmb_tune_set(int key, int value) {
  ABCMB_REQUIRE(key >= 0 && key < ABCMB_TUNE_KEYS, "unknown tuning key %d", key);
  ABCMB_REQUIRE(value >= 0, "tuning value must be >= 0");
  if (key == ABCMB_TUNE_PERT_PERSISTENT) ABCMB_REQUIRE(value <= 400, "persistent grid above 400%% of residency");
  if (key == ABCMB_TUNE_LENS_GEMM_COLS) ABCMB_REQUIRE(value == 0 || value == 32 || value == 64, "column tile 32 or 64");
  abcmb::g_tune[key].store(value, std::memory_order_relaxed);
  return ABCMB_OK;
}

extern "C" int abcmb_tune_get(int key) { return abcmb::tune_get(key); }
