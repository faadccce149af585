// tt_common.cpp -- error reporting shared by every entry point.
#include <cstdarg>
#include <cstdio>

#include "tritok_bem.h"
#include "tt_internal.h"

namespace {
thread_local char g_err[1024] = "";
}

void tt_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" const char *tt_last_error(void) { return g_err; }
