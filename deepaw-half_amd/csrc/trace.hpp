// Launch-boundary timeline for measurement: when enabled, an event is
// recorded on the launching stream before every kernel of the operator and
// FFT paths; the time between consecutive marks is that kernel's duration.
// Disabled (the default) it costs one branch per launch.
#pragma once

#include <hip/hip_runtime.h>

namespace half {

void trace_mark(const char *name, hipStream_t s);  // before a launch
void trace_close(hipStream_t s);                   // after the last launch of a sequence
bool trace_enabled();

}  // namespace half
