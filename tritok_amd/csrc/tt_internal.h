// tt_internal.h -- declarations shared by the library's translation units.
#pragma once
#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <stdexcept>

// Records a printf-style message for tt_last_error() on the calling thread.
void tt_set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
