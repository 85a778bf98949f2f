// C-ABI housekeeping: error state and version.
#include <string>

#include "internal.h"
#include "mflops/mf.h"

namespace mf {
static thread_local std::string g_last_error;
int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}
}  // namespace mf

extern "C" const char *mf_last_error(void) { return mf::g_last_error.c_str(); }
extern "C" int mf_version(void) { return 100; }
