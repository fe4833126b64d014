// Library-wide error state and version for the jfx C ABI.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

namespace jfx {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace jfx

extern "C" {

int jfx_version(void) { return 100; }

const char* jfx_last_error(void) { return jfx::g_err; }

int jfx_device_sync(void) {
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    jfx::set_error("hipDeviceSynchronize: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

}  // extern "C"
