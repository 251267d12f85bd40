// ns_internal.h -- shared declarations of the nsphere_amd native library.
#pragma once

#include <cstdint>

#include "nsphere_gpu.h"

namespace ns {

// Thread-local last-error text; returns code for "return set_error(...)".
int set_error(int code, const char *fmt, ...);
void clear_error();

} // namespace ns
