// Thread-local error message and library-level entry points of the C ABI.
#include "common.h"

#include <hip/hip_runtime.h>

namespace scxr {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

}  // namespace scxr

extern "C" {

int scxr_version(void) { return SCXR_VERSION; }

const char* scxr_last_error(void) { return scxr::g_last_error.c_str(); }

int scxr_device_info(int32_t* cu_count, int32_t* warp_size, int64_t* lds_bytes,
                     char* arch_name, int32_t arch_len) {
  int dev = 0;
  SCXR_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  SCXR_HIP(hipGetDeviceProperties(&p, dev));
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (warp_size) *warp_size = p.warpSize;
  if (lds_bytes) *lds_bytes = static_cast<int64_t>(p.sharedMemPerBlock);
  if (arch_name && arch_len > 0) snprintf(arch_name, arch_len, "%s", p.gcnArchName);
  return SCXR_OK;
}

}  // extern "C"
