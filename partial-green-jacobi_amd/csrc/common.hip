// Error reporting and version for libpgj.
#include <cstdarg>
#include <cstdio>
#include <string>

#include "pgj_internal.h"

namespace pgj {
namespace {
thread_local std::string g_error;
}

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
}

void clear_error() { g_error.clear(); }

}  // namespace pgj

extern "C" int pgj_version(void) { return 1; }

extern "C" const char* pgj_last_error(void) { return pgj::g_error.c_str(); }
