// Library-level entry points: ABI version and per-thread error messages.
#include <stdarg.h>
#include <string.h>

#include "common.h"
#include "hrdaya_hip.h"

namespace hp {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", what, hipGetErrorString(e));
    return -(int)e;
  }
  return 0;
}

}  // namespace hp

extern "C" int hp_abi_version(void) { return HP_ABI_VERSION; }

extern "C" int hp_device_alloc(void** ptr, int64_t bytes, int contiguous) {
  if (!ptr || bytes <= 0) {
    hp::set_error("hp_device_alloc: bad arguments");
    return -1;
  }
  const hipError_t e =
      hipExtMallocWithFlags(ptr, (size_t)bytes, contiguous ? hipDeviceMallocContiguous
                                                           : hipDeviceMallocDefault);
  if (e != hipSuccess) {
    *ptr = nullptr;
    hp::set_error("hp_device_alloc: %lld bytes: %s", (long long)bytes, hipGetErrorString(e));
    return -(int)e;
  }
  return 0;
}

extern "C" int hp_device_free(void* ptr) {
  if (!ptr) return 0;
  const hipError_t e = hipFree(ptr);
  if (e != hipSuccess) {
    hp::set_error("hp_device_free: %s", hipGetErrorString(e));
    return -(int)e;
  }
  return 0;
}

extern "C" int hp_last_error(char* buf, size_t size) {
  if (!buf || size == 0) return -1;
  strncpy(buf, hp::g_err, size - 1);
  buf[size - 1] = 0;
  return 0;
}
