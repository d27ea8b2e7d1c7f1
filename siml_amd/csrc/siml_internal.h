// siml_internal.h -- error plumbing shared by the translation units of
// libsiml_amd.so.
#pragma once

#include <cstdarg>
#include <cstdio>

#include "../../include/siml_amd.h"

namespace siml {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

}  // namespace siml
