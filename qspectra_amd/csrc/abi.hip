// Library-wide C-ABI helpers: version and per-thread error messages.
#include "qsx_common.h"

#include <cstdarg>
#include <cstring>

namespace qsx {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace qsx

extern "C" int32_t qsx_version(void) { return 100; }  // 0.1.0

extern "C" int32_t qsx_last_error(char* buf, size_t len) {
  if (buf == nullptr || len == 0) return QSX_EINVAL;
  strncpy(buf, qsx::g_err, len - 1);
  buf[len - 1] = '\0';
  return QSX_OK;
}
