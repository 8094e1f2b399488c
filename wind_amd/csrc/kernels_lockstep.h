// Lockstep SIE attempt kernel (the reference's host-driven integrator,
// cuda/SIE_host/harness_kernel.cu:175-246 solveSystem and the error test of
// cuda/SIE_host/vector_kernels.cu:45-70): one launch per attempt over all
// systems, one system per lane.  Writes the two-half-step state and raises
// *flag when any system fails the test (strict '>' comparisons; the
// relative test requires y1 > ABS and y2 > ABS, without absolute values).
#pragma once
#include "kernels_small.h"

namespace wind {

template <class Net>
__global__ void __launch_bounds__(256)
    lockstep_attempt(float tnow, float hh, const float *__restrict__ C, int stride,
                     const float *__restrict__ Y, float *__restrict__ Y2, int nsys, float ABS,
                     float REL, int *flag, typename Net::Params p) {
    constexpr int N = Net::N;
    constexpr int NC = Net::NCONST;
    const float tb = tnow + hh;
    const float hs1 = tb - tnow;          // solveSystem(tnow, tnow + h, 1)
    const float hs2 = (tb - tnow) / 2.0f; // solveSystem(tnow, tnow + h, 2)
    bool fail = false;
    for (int sid = blockIdx.x * blockDim.x + threadIdx.x; sid < nsys; sid += gridDim.x * blockDim.x) {
        float y0[N], y1[N], y2[N], c[NC];
#pragma unroll
        for (int i = 0; i < N; i++) y0[i] = Y[(size_t)sid * N + i];
#pragma unroll
        for (int k = 0; k < NC; k++) c[k] = C[(size_t)sid * stride + k];
        substep_pair<Net, WIND_SIE, MathExact>(p, c, y0, hs1, hs2, y1, y2);
        substep<Net, WIND_SIE, MathExact>(p, c, y2, hs2);
#pragma unroll
        for (int i = 0; i < N; i++) {
            const float d = y1[i] - y2[i];
            if (fabsf(d) > ABS) fail = true;
            if (fabs((double)d / ((double)y2[i] + 1e-12)) > (double)REL && y1[i] > ABS && y2[i] > ABS)
                fail = true;
            Y2[(size_t)sid * N + i] = y2[i];
        }
    }
    if (__any(fail) && __lane_id() == 0) atomicOr(flag, 1);
}

}  // namespace wind
