// Fast-mode lane kernels of the compiled-in networks, in a translation unit
// of their own so that it can be compiled with fused multiply-add
// contraction (-ffp-contract=fast): fast mode promises no operation order,
// exact mode (wind_api.hip, contraction off) keeps the reference's.
#include <hip/hip_runtime.h>

#include "kernels_small.h"
#include "wind_internal.h"

using namespace wind;

#ifndef WIND_LANE_FAST_MIN_BLOCKS
#define WIND_LANE_FAST_MIN_BLOCKS 1
#endif

namespace {

template <class Net, int METHOD>
__global__ void __launch_bounds__(256, WIND_LANE_FAST_MIN_BLOCKS)
    lane_fast_kernel(IntegrateArgs a, typename Net::Params p) {
    lane_integrate_body<Net, METHOD, MathFast>(a, p);
}

}  // namespace

// Persistent blocks per CU.  Fewer lanes than the registers allow run
// faster: on the 2^20-system hhe6 workload (SIE, tol 2e-6) a launch took
// 70.3 / 61.6 / 66.3 / 71.8 ms with 1 / 2 / 3 / 4 blocks (waves per SIMD).
constexpr int kLaneFastBlocksPerCU = 2;

template <class Net, int METHOD>
int launch_lane_fast(int num_cu, const typename Net::Params &p, const IntegrateArgs &a, const SolverChoice &sc,
                     hipStream_t st) {
    auto k = lane_fast_kernel<Net, METHOD>;
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0);
    if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    per_cu = per_cu < kLaneFastBlocksPerCU ? per_cu : kLaneFastBlocksPerCU;
    const int grid = scaled_grid(num_cu, per_cu, sc, a.nsys, 256);
    hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, st, a, p);
    e = hipGetLastError();
    if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    return WIND_OK;
}

#define WIND_LANE_FAST(NET)                                                                                   \
    template int launch_lane_fast<NET, WIND_SIE>(int, const NET::Params &, const IntegrateArgs &,             \
                                                 const SolverChoice &, hipStream_t);                          \
    template int launch_lane_fast<NET, WIND_RK2>(int, const NET::Params &, const IntegrateArgs &,             \
                                                 const SolverChoice &, hipStream_t);
WIND_LANE_FAST(NetNRTest)
WIND_LANE_FAST(NetKatz96)
WIND_LANE_FAST(NetHHe6)
