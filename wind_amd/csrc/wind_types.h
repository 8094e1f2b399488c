// Types shared by the host library and the device kernels.
#pragma once
#include <stdint.h>

#include "../../include/wind.h"

namespace wind {

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
};

}  // namespace wind
