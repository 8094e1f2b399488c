// Plugin side of plugin_abi.h: launchers for a generated network type and
// the exported descriptor.  A generated plugin source is
//     #include "<network header>"
//     #include "plugin_impl.h"
//     WIND_DEFINE_PLUGIN(NetType, "name", nreactions, flops_rhs, flops_jac)
// and is compiled twice into the plugin: with WIND_PLUGIN_PART=1 (exact
// mode and lockstep kernels, the descriptor; no FMA contraction) and
// WIND_PLUGIN_PART=2 (the fast-mode kernels, with FMA contraction).
#pragma once
#include "kernels_generated.h"
#include "kernels_lockstep.h"
#include "kernels_small.h"
#include "plugin_abi.h"
#include "wind_internal.h"

namespace wind {

template <class K>
inline int plugin_grid(K kernel, int block, int nsys, int num_cu, float share, int grid_cap, int *grid) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess)
        return WIND_ERR_DEVICE;
    SolverChoice sc;
    sc.share = share;
    sc.grid = grid_cap;
    *grid = scaled_grid(num_cu, per_cu, sc, nsys, block);
    return WIND_OK;
}

#ifndef WIND_PLUGIN_PART
#define WIND_PLUGIN_PART 1
#endif

// Generated networks up to this N also run on the lane-per-system kernels
// (the whole system in registers, code unrolled for N; the code generator's
// LANE_MAX_N); larger ones on the one-workgroup-per-system kernel
// (kernels_generated.h), whose code size does not grow with N.  The
// reference-shaped 64-thread block kernel keeps one equation per lane.
constexpr int kGenLaneMaxN = 16;
constexpr int kGenBlock64MaxN = 64;

// Part 2 (fast mode, compiled with FMA contraction): launch and scratch size.
extern "C" __attribute__((visibility("hidden"))) int wind_plugin_launch_fast(const IntegrateArgs *a, int method,
                                                                          int kernel, int num_cu, float share,
                                                                          int grid, float *scratch, size_t sbytes,
                                                                          hipStream_t st);
extern "C" __attribute__((visibility("hidden"))) int wind_plugin_scratch_fast(int nsys, int method, int kernel,
                                                                           int num_cu, float share, int grid,
                                                                           size_t *bytes);

// Which kernel a launch uses: 0 lane, 1 the 64-thread block kernel, 2 the
// one-workgroup-per-system kernel, -1 unsupported.
template <class Net>
constexpr int gen_kernel_choice(int mode, int kernel) {
    if (kernel == WIND_KERNEL_LANE) return Net::N <= kGenLaneMaxN ? 0 : -1;
    if (kernel == WIND_KERNEL_BLOCK) return (mode == WIND_MODE_EXACT && Net::N <= kGenBlock64MaxN) ? 1 : 2;
    return Net::N <= kGenLaneMaxN ? 0 : 2;
}

#if WIND_PLUGIN_PART == 1
template <class Net, int METHOD>
int plugin_launch_method(const IntegrateArgs &a, int mode, int kernel, int num_cu, float share, int grid_cap,
                         float *scratch, size_t sbytes, hipStream_t st) {
    if (mode != WIND_MODE_EXACT)
        return wind_plugin_launch_fast(&a, METHOD, kernel, num_cu, share, grid_cap, scratch, sbytes, st);
    typename Net::Params p{};
    const int choice = gen_kernel_choice<Net>(mode, kernel);
    if (choice == 1) {
        if constexpr (Net::N <= kGenBlock64MaxN) {
            SolverChoice sc;
            sc.share = share;
            sc.grid = grid_cap;
            const int grid = scaled_grid(num_cu, 32, sc, a.nsys);
            hipLaunchKernelGGL((block_integrate<Net, METHOD>), dim3(grid), dim3(64), 0, st, a, p);
            return hipGetLastError() == hipSuccess ? WIND_OK : WIND_ERR_DEVICE;
        }
    } else if (choice == 0) {
        if constexpr (Net::N <= kGenLaneMaxN) {
            auto k = lane_integrate<Net, METHOD, MathExact>;
            int grid = 0;
            const int rc = plugin_grid(k, 256, a.nsys, num_cu, share, grid_cap, &grid);
            if (rc) return rc;
            hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, st, a, p);
            return hipGetLastError() == hipSuccess ? WIND_OK : WIND_ERR_DEVICE;
        }
    } else if (choice == 2) {
        return gen_block_launch<Net, METHOD, false>(a, p, num_cu, share, grid_cap, scratch, sbytes, st);
    }
    return WIND_ERR_UNSUPPORTED;
}

template <class Net>
int plugin_launch(const IntegrateArgs *a, int method, int mode, int kernel, int num_cu, float share, int grid,
                  float *scratch, size_t sbytes, hipStream_t st) {
    if (a->trace.capacity > 0) return WIND_ERR_UNSUPPORTED;  // no traced kernels in plugins
    if (method == WIND_SIE)
        return plugin_launch_method<Net, WIND_SIE>(*a, mode, kernel, num_cu, share, grid, scratch, sbytes, st);
    if (method == WIND_RK2)
        return plugin_launch_method<Net, WIND_RK2>(*a, mode, kernel, num_cu, share, grid, scratch, sbytes, st);
    return WIND_ERR_ARGUMENT;
}

template <class Net>
int plugin_scratch_bytes(int nsys, int method, int mode, int kernel, int num_cu, float share, int grid,
                         size_t *bytes) {
    *bytes = 0;
    if (method != WIND_SIE && method != WIND_RK2) return WIND_ERR_ARGUMENT;
    if (mode != WIND_MODE_EXACT) return wind_plugin_scratch_fast(nsys, method, kernel, num_cu, share, grid, bytes);
    if (gen_kernel_choice<Net>(mode, kernel) != 2) return WIND_OK;
    GenBlockShape s;
    const int rc = method == WIND_SIE ? gen_block_shape<Net, WIND_SIE, false>(nsys, num_cu, share, grid, &s)
                                      : gen_block_shape<Net, WIND_RK2, false>(nsys, num_cu, share, grid, &s);
    *bytes = s.scratch;
    return rc;
}

template <class Net>
int plugin_lockstep(float tnow, float hh, const float *C, int stride, const float *Y, float *Y2, int nsys,
                    float ABS, float REL, int *flag, int grid) {
    if constexpr (Net::N <= kGenLaneMaxN) {
        typename Net::Params p{};
        hipLaunchKernelGGL(lockstep_attempt<Net>, dim3(grid), dim3(256), 0, 0, tnow, hh, C, stride, Y, Y2, nsys,
                           ABS, REL, flag, p);
        return hipGetLastError() == hipSuccess ? WIND_OK : WIND_ERR_DEVICE;
    } else {
        return WIND_ERR_UNSUPPORTED;  // lockstep attempts keep each system in one lane's registers
    }
}
#else
template <class Net, int METHOD>
int plugin_launch_fast_method(const IntegrateArgs &a, int kernel, int num_cu, float share, int grid_cap,
                              float *scratch, size_t sbytes, hipStream_t st) {
    typename Net::Params p{};
    const int choice = gen_kernel_choice<Net>(WIND_MODE_FAST, kernel);
    if (choice == 0) {
        if constexpr (Net::N <= kGenLaneMaxN) {
            auto k = a.pivot ? lane_integrate<Net, METHOD, MathFastPivot> : lane_integrate<Net, METHOD, MathFast>;
            int grid = 0;
            const int rc = plugin_grid(k, 256, a.nsys, num_cu, share, grid_cap, &grid);
            if (rc) return rc;
            hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, st, a, p);
            return hipGetLastError() == hipSuccess ? WIND_OK : WIND_ERR_DEVICE;
        }
    } else if (choice == 2) {
        return a.pivot ? gen_block_launch<Net, METHOD, true, true>(a, p, num_cu, share, grid_cap, scratch, sbytes, st)
                       : gen_block_launch<Net, METHOD, true>(a, p, num_cu, share, grid_cap, scratch, sbytes, st);
    }
    return WIND_ERR_UNSUPPORTED;
}

template <class Net>
int plugin_scratch_fast(int nsys, int method, int kernel, int num_cu, float share, int grid, size_t *bytes) {
    *bytes = 0;
    if (gen_kernel_choice<Net>(WIND_MODE_FAST, kernel) != 2) return WIND_OK;
    GenBlockShape s;
    const int rc = method == WIND_SIE ? gen_block_shape<Net, WIND_SIE, true>(nsys, num_cu, share, grid, &s)
                                      : gen_block_shape<Net, WIND_RK2, true>(nsys, num_cu, share, grid, &s);
    *bytes = s.scratch;
    return rc;
}
#endif

}  // namespace wind

#if WIND_PLUGIN_PART == 1
#define WIND_DEFINE_PLUGIN(NET, NAME, NREACT, FLOPS_RHS, FLOPS_JAC)                                      \
    extern "C" __attribute__((visibility("default"))) const wind_plugin_desc *wind_plugin_get(void) {     \
        static const wind_plugin_desc d = {WIND_PLUGIN_ABI,       NET::N,                                  \
                                           NET::NCONST,           NREACT,                                  \
                                           NAME,                  FLOPS_RHS,                               \
                                           FLOPS_JAC,             wind::plugin_launch<NET>,                \
                                           wind::plugin_lockstep<NET>, wind::plugin_scratch_bytes<NET>};   \
        return &d;                                                                                         \
    }
#else
#define WIND_DEFINE_PLUGIN(NET, NAME, NREACT, FLOPS_RHS, FLOPS_JAC)                                      \
    extern "C" __attribute__((visibility("hidden"))) int wind_plugin_launch_fast(                         \
        const wind::IntegrateArgs *a, int method, int kernel, int num_cu, float share, int grid,             \
        float *scratch, size_t sbytes, hipStream_t st) {                                                   \
        if (method == WIND_SIE)                                                                            \
            return wind::plugin_launch_fast_method<NET, WIND_SIE>(*a, kernel, num_cu, share, grid, scratch,   \
                                                                  sbytes, st);                             \
        if (method == WIND_RK2)                                                                            \
            return wind::plugin_launch_fast_method<NET, WIND_RK2>(*a, kernel, num_cu, share, grid, scratch,   \
                                                                  sbytes, st);                             \
        return WIND_ERR_ARGUMENT;                                                                          \
    }                                                                                                      \
    extern "C" __attribute__((visibility("hidden"))) int wind_plugin_scratch_fast(                        \
        int nsys, int method, int kernel, int num_cu, float share, int grid, size_t *bytes) {              \
        return wind::plugin_scratch_fast<NET>(nsys, method, kernel, num_cu, share, grid, bytes);           \
    }
#endif
