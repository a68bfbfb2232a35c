// Error reporting and version entry points of the C ABI.
#include "pyl_common.hpp"

namespace pyl {

static thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void taylor_matrix(int n, double dt, double* out) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      out[i * n + j] = j <= i ? powi(dt, i - j) / fact_d(i - j) : 0.0;
}

}  // namespace pyl

extern "C" const char* pyl_last_error(void) { return pyl::g_last_error.c_str(); }

extern "C" int pyl_abi_version(void) { return 107; }
