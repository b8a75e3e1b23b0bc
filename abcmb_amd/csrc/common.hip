// Error state and device initialisation of libabcmb_hip.so.
#include <cstdarg>
#include <cstdio>

#include "common.h"

namespace abcmb {

static thread_local char g_err[512] = {0};

int scratch_address(void** out);  // pert.hip

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
