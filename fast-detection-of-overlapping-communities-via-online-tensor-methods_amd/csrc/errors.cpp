// Thread-local error message storage behind mmsb_last_error(), and the version string.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "mmsb_stpm.h"

namespace mmsb {

static thread_local char g_last_error[1024] = "";

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

void clear_error() { g_last_error[0] = '\0'; }

}  // namespace mmsb

extern "C" const char* mmsb_last_error(void) { return mmsb::g_last_error; }
extern "C" const char* mmsb_version(void) { return "mmsb-amd 0.1 (gfx950)"; }
