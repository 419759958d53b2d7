// Error reporting and version for the C ABI (include/torchmorph_hip.h).
#include <cstdarg>

#include "tm_common.h"

namespace tmh {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* get_error() { return g_err; }

}  // namespace tmh

extern "C" int tm_last_error(char* buf, size_t len) {
  const char* e = tmh::get_error();
  size_t n = strlen(e);
  if (buf && len) {
    size_t k = n < len - 1 ? n : len - 1;
    memcpy(buf, e, k);
    buf[k] = 0;
  }
  return (int)n;
}

extern "C" const char* tm_version(void) { return "torchmorph_hip 0.1 gfx950"; }
