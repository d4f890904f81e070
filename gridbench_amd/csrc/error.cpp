// Status codes and the thread-local "last error" message.
#include <cstdarg>
#include <cstdio>

#include "gbx_internal.h"

static thread_local std::string g_last_error;

int gbx_fail(int status, const char *fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return status;
}

void gbx_clear_error() { g_last_error.clear(); }

extern "C" const char *gbx_last_error(void) { return g_last_error.c_str(); }

extern "C" const char *gbx_strerror(int status)
{
  switch (status) {
  case GBX_OK: return "success";
  case GBX_EINVAL: return "invalid argument";
  case GBX_ENULL: return "null pointer";
  case GBX_EALIAS: return "output aliases input";
  case GBX_ETABLE: return "invalid stencil table";
  case GBX_ENOMEM: return "out of memory";
  case GBX_EHIP: return "HIP runtime error";
  case GBX_ENODEV: return "no usable GPU";
  case GBX_EUNSUPPORTED: return "unsupported";
  case GBX_EMISMATCH: return "mismatched objects";
  case GBX_ENOCONV: return "solver did not converge";
  default: return "unknown status";
  }
}

extern "C" int gbx_version(void) { return 100; }

extern "C" int gbx_device_count(void)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int gbx_require_device()
{
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return gbx_fail(GBX_ENODEV, "no HIP device available (%s)", hipGetErrorString(e));
  }
  return GBX_OK;
}
