// Library-level entry points: version, status text, per-thread error text.
#include <stdarg.h>

#include "reno_common.h"

namespace reno {

static thread_local char t_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_err, sizeof(t_err), fmt, ap);
  va_end(ap);
}

}  // namespace reno

extern "C" int reno_version(void) { return 10000; }

extern "C" const char* reno_last_error(void) { return reno::t_err; }

extern "C" const char* reno_status_string(int status) {
  switch (status) {
    case RENO_OK: return "ok";
    case RENO_EINVAL: return "invalid argument";
    case RENO_EHIP: return "HIP runtime error";
    case RENO_ENOSPACE: return "workspace too small";
    case RENO_ESOLVER: return "rocSOLVER error";
    case RENO_EQN: return "invalid quantum number";
    default: return "unknown status";
  }
}
