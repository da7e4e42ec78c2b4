// Device workspace of the allocation-free operator entry point.
#pragma once

#include <cstddef>

namespace half {

// Size of the per-device workspace array (bytes).
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;

// Device address of the current device's workspace array, looked up once per
// device (hipGetSymbolAddress with that device current) and kept; no
// allocation.  Throws kInternal if the lookup fails.
void *noalloc_scratch();

}  // namespace half
