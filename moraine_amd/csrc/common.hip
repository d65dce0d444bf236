// Error state and ABI version of libmoraine_hip.
#include "common.hpp"

namespace mr {
static thread_local char g_err[512] = "";
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
}  // namespace mr

extern "C" {
const char* mr_last_error_string(void) { return mr::g_err; }
int mr_abi_version(void) { return 1; }
}
