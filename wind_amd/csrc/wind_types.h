// Types shared by the host library and the device kernels.
#pragma once
#include <stdint.h>

#include "../../include/wind.h"

namespace wind {

// Per-attempt trace of an integration (wind_trace in wind.h, device
// pointers).  capacity == 0: no trace.  With h set, attempt k of system s
// takes the step h[s*capacity + k] and its acceptance from accept[...],
// and the system stops after count[s] attempts (replay), each attempt
// starting from y_start[(s*capacity + k)*N] when that is set; y1, y2 and flag,
// when set, receive each attempt's full-step and two-half-step states and
// the kernel's own error test (1 = its controller would reject).
struct TraceArgs {
    int32_t capacity = 0;
    const float *h = nullptr;
    const int32_t *accept = nullptr;
    const int32_t *count = nullptr;
    const float *y_start = nullptr;  // replay: each attempt's starting state (optional)
    float *y1 = nullptr;
    float *y2 = nullptr;
    int32_t *flag = nullptr;
};

// Kernel arguments common to every integrator kernel.  Pointers are device
// pointers.  Systems are stored as in the reference ABI: equations[s*N + i],
// constants[s*const_stride + k].
struct IntegrateArgs {
    float tnow, tend, h0;          // h0 = (tend - tnow) / n_integration_steps
    float ABS, REL;
    int controller;                // wind_controller
    const float *constants;
    float *equations;
    int32_t nsys;
    int32_t const_stride;
    int64_t max_attempts;          // per system
    int32_t *steps_out;            // optional [nsys]: 3 * attempts
    int32_t *status_out;           // optional [nsys]
    const int32_t *index;          // optional [nsys]: permutation of system ids
    unsigned long long *counters;  // [0] work-queue head, [1] total sub-steps
    TraceArgs trace;               // per-attempt trace / replay (kernels that support it)
    int32_t pivot = 0;             // fast mode: partial pivoting in the LU solves (launchers pick the kernel)
    // optional, with index (lane kernels): record q (lane_record_words(N,
    // NCONST) floats) holds the state, the constants and the id (as int
    // bits) of system index[q], packed in queue order, so that lanes read the
    // systems they take contiguously instead of gathering them
    const float *records = nullptr;
};

// Floats per packed record: state, constants and the system id, rounded up
// to whole 32-byte sectors.
constexpr int lane_record_words(int N, int NC) { return (N + NC + 1 + 7) / 8 * 8; }

}  // namespace wind
