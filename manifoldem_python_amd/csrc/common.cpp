// Error reporting and version of libmemi.
#include "common.h"

#include <cstring>

namespace memi {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace memi

extern "C" {

int memi_version(void) { return 10000; }

const char* memi_last_error(void) { return memi::g_err; }

}  // extern "C"
