// Host-side plumbing of libbpl_amd: thread-local error strings and version.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "bpl_common.h"

namespace bpl {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int check_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return BPL_OK;
  set_error("%s: %s", what, hipGetErrorString(e));
  return BPL_E_HIP;
}

}  // namespace bpl

extern "C" int bpl_version(void) { return 1; }
extern "C" const char *bpl_last_error(void) { return bpl::g_err; }
