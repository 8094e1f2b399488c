// Fast-mode lane kernels of the compiled-in networks, in a translation unit
// of their own so that it can be compiled with fused multiply-add
// contraction (-ffp-contract=fast): fast mode promises no operation order,
// exact mode (wind_api.hip, contraction off) keeps the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels_small.h"
#include "wind_internal.h"

using namespace wind;

#ifndef WIND_LANE_FAST_MIN_BLOCKS
#define WIND_LANE_FAST_MIN_BLOCKS 1
#endif

namespace {

// M: MathFast, or MathFastPivot for partial pivoting in the LU solves.
template <class Net, int METHOD, class M>
__global__ void __launch_bounds__(256, WIND_LANE_FAST_MIN_BLOCKS)
    lane_fast_kernel(IntegrateArgs a, typename Net::Params p) {
    lane_integrate_body<Net, METHOD, M>(a, p);
}

// The same integration with a per-attempt trace or replay (a.trace); the
// arithmetic of each attempt is lane_fast_kernel's.
template <class Net, int METHOD, class M>
__global__ void __launch_bounds__(256) lane_fast_trace_kernel(IntegrateArgs a, typename Net::Params p) {
    lane_integrate_body<Net, METHOD, M, true>(a, p);
}

// Start-order estimate from the initial state alone: x = (tend - tnow) *
// max_i |f_i(y0)| / (|y0_i| + ABS), the fastest relative rate of change
// over the interval, one right-hand side per system.  Its float bits go to
// xbits[s] (non-negative floats order as their bit patterns) and the
// largest to *xmax.
template <class Net>
__global__ void __launch_bounds__(256) lane_activity(IntegrateArgs a, typename Net::Params p, unsigned *xbits,
                                                      unsigned *xmax) {
    constexpr int N = Net::N;
    constexpr int NC = Net::NCONST;
    const float span = a.tend - a.tnow;
    const float floor_abs = a.ABS > 0.0f ? a.ABS : 1.17549435e-38f;
    unsigned best = 0u;
    for (int sid = blockIdx.x * blockDim.x + threadIdx.x; sid < a.nsys; sid += gridDim.x * blockDim.x) {
        float y[N], c[NC], f[N];
#pragma unroll
        for (int i = 0; i < N; i++) y[i] = a.equations[(size_t)sid * N + i];
#pragma unroll
        for (int k = 0; k < NC; k++) c[k] = a.constants[(size_t)sid * a.const_stride + k];
        Net::template rhs<MathFast>(p, c, y, f);
        float r = 0.0f;
#pragma unroll
        for (int i = 0; i < N; i++) r = fmaxf(r, fabsf(f[i]) / (fabsf(y[i]) + floor_abs));
        float x = r * span;
        if (!(x >= 0.0f)) x = 0.0f;                   // NaN
        if (x > 3.0e38f) x = 3.0e38f;                 // inf
        const unsigned b = __float_as_uint(x);
        xbits[sid] = b;
        best = b > best ? b : best;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned o = __shfl_xor(best, off);
        best = o > best ? o : best;
    }
    if (__lane_id() == 0 && best) atomicMax(xmax, best);
}

// Start-order estimate from a short probe: the first kProbeAttempts
// attempts of each system from its initial state (lane_fast_kernel's
// attempts; the state is then discarded), extrapolated to the end of the
// interval: attempts so far + (tend - t) / h x attempts per accepted one,
// with h the controller's next step (a system that finishes within the probe
// counts its attempts).  Float bits to xbits[s], the largest to *xmax, as
// lane_activity.
constexpr int kProbeAttempts = 16;

template <class Net, int METHOD, class M>
__global__ void __launch_bounds__(256) lane_probe(IntegrateArgs a, typename Net::Params p, unsigned *xbits,
                                                   unsigned *xmax) {
    constexpr int N = Net::N;
    constexpr int NC = Net::NCONST;
    unsigned best = 0u;
    for (int sid = blockIdx.x * blockDim.x + threadIdx.x; sid < a.nsys; sid += gridDim.x * blockDim.x) {
        LaneSys<N> s;
        float c[NC];
        load_state<N>(a.equations + (size_t)sid * N, s.y);
#pragma unroll
        for (int k = 0; k < NC; k++) c[k] = a.constants[(size_t)sid * a.const_stride + k];
        s.tnow = a.tnow;
        s.h = a.h0;
        s.unsolved = 0;
        s.attempts = 0;
        s.status = WIND_STATUS_OK;
        int accepted = 0;
        bool done = false;
        for (int k = 0; k < kProbeAttempts && !done; k++) {
            const float t0 = s.tnow;
            done = attempt<Net, METHOD, M>(p, a, c, s);
            accepted += s.tnow != t0 ? 1 : 0;
        }
        float x = (float)s.attempts;
        if (!done)
            x += (a.tend - s.tnow) / fmaxf(s.h, 1.0e-30f) * ((float)s.attempts / (float)(accepted > 0 ? accepted : 1));
        if (!(x >= 0.0f)) x = 0.0f;                   // NaN
        if (x > 3.0e38f) x = 3.0e38f;                 // inf
        const unsigned b = __float_as_uint(x);
        xbits[sid] = b;
        best = b > best ? b : best;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned o = __shfl_xor(best, off);
        best = o > best ? o : best;
    }
    if (__lane_id() == 0 && best) atomicMax(xmax, best);
}

// Turns the estimates into cost hints whose classes (four per octave, as
// the start-order sort forms them) keep 29 octaves below the largest: an
// estimate e octaves below it gets class 124 - floor(4 e) (at least 9), and
// the hint is the smallest integer of that class.
__global__ void __launch_bounds__(256) lane_activity_hints(const unsigned *xbits, int n, const unsigned *xmax,
                                                            int32_t *hint) {
    const float lmax = __log2f(fmaxf(__uint_as_float(*xmax), 1.0e-30f));
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const float l = __log2f(fmaxf(__uint_as_float(xbits[s]), 1.0e-30f));
        const float below = fminf(fmaxf(4.0f * (lmax - l), 0.0f), 115.0f);
        const int cls = 124 - (int)below;             // 9 .. 124
        const int lg = (cls - 1) >> 2, sub = (cls - 1) & 3;
        hint[s] = (4 + sub) << (lg - 2);
    }
}

}  // namespace

template <class Net>
int launch_lane_activity(int num_cu, const typename Net::Params &p, const IntegrateArgs &a, int32_t *hint,
                         unsigned *scratch, hipStream_t st) {
    // scratch: [0] the largest estimate, then one word per system
    unsigned *xmax = scratch, *xbits = scratch + 64;
    const int grid = std::min(num_cu * 8, (a.nsys + 255) / 256);
    hipLaunchKernelGGL(lane_activity<Net>, dim3(grid), dim3(256), 0, st, a, p, xbits, xmax);
    hipLaunchKernelGGL(lane_activity_hints, dim3(grid), dim3(256), 0, st, xbits, a.nsys, xmax, hint);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? WIND_OK : wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
}

template <class Net, int METHOD>
int launch_lane_probe(int num_cu, const typename Net::Params &p, const IntegrateArgs &a, int32_t *hint,
                      unsigned *scratch, hipStream_t st) {
    // scratch: [0] the largest estimate, then one word per system
    unsigned *xmax = scratch, *xbits = scratch + 64;
    const int grid = std::min(num_cu * 8, (a.nsys + 255) / 256);
    auto k = a.pivot ? lane_probe<Net, METHOD, MathFastPivot> : lane_probe<Net, METHOD, MathFast>;
    hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, st, a, p, xbits, xmax);
    hipLaunchKernelGGL(lane_activity_hints, dim3(grid), dim3(256), 0, st, xbits, a.nsys, xmax, hint);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? WIND_OK : wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
}

// Persistent blocks per CU, by method.  Fewer lanes than the registers
// allow run faster.  On the 2^20-system hhe6 workload at tol 2e-6 (start
// order from the probe) a launch took, with 1 / 2 / 3 / 4 blocks (waves per
// SIMD): SIE 55.0 / 48.9 ms / - / - (70.3 / 61.6 / 66.3 / 71.8 ms with an
// earlier kernel and no start order); RK2 242 / 278 / 320 / 368 ms (RK2's
// sub-step counts run to 22x their mean, and a lane alone on its SIMD
// finishes the longest systems sooner).
constexpr int lane_fast_blocks_per_cu(int method) { return method == WIND_RK2 ? 1 : 2; }

template <class Net, int METHOD>
int launch_lane_fast(int num_cu, const typename Net::Params &p, const IntegrateArgs &a, const SolverChoice &sc,
                     hipStream_t st) {
    auto k = a.pivot ? (a.trace.capacity > 0 ? lane_fast_trace_kernel<Net, METHOD, MathFastPivot>
                                             : lane_fast_kernel<Net, METHOD, MathFastPivot>)
                     : (a.trace.capacity > 0 ? lane_fast_trace_kernel<Net, METHOD, MathFast>
                                             : lane_fast_kernel<Net, METHOD, MathFast>);
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0);
    if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    per_cu = per_cu < lane_fast_blocks_per_cu(METHOD) ? per_cu : lane_fast_blocks_per_cu(METHOD);
    const int grid = scaled_grid(num_cu, per_cu, sc, a.nsys, 256);
    hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, st, a, p);
    e = hipGetLastError();
    if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    return WIND_OK;
}

#define WIND_LANE_FAST(NET)                                                                                   \
    template int launch_lane_activity<NET>(int, const NET::Params &, const IntegrateArgs &, int32_t *,         \
                                           unsigned *, hipStream_t);                                          \
    template int launch_lane_probe<NET, WIND_SIE>(int, const NET::Params &, const IntegrateArgs &, int32_t *, \
                                                  unsigned *, hipStream_t);                                   \
    template int launch_lane_probe<NET, WIND_RK2>(int, const NET::Params &, const IntegrateArgs &, int32_t *, \
                                                  unsigned *, hipStream_t);                                   \
    template int launch_lane_fast<NET, WIND_SIE>(int, const NET::Params &, const IntegrateArgs &,             \
                                                 const SolverChoice &, hipStream_t);                          \
    template int launch_lane_fast<NET, WIND_RK2>(int, const NET::Params &, const IntegrateArgs &,             \
                                                 const SolverChoice &, hipStream_t);
WIND_LANE_FAST(NetNRTest)
WIND_LANE_FAST(NetKatz96)
WIND_LANE_FAST(NetHHe6)
