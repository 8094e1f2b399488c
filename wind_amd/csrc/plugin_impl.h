// Plugin side of plugin_abi.h: launchers for a generated network type and
// the exported descriptor.  A generated plugin source is
//     #include "<network header>"
//     #include "plugin_impl.h"
//     WIND_DEFINE_PLUGIN(NetType, "name", nreactions, flops_rhs, flops_jac)
// and is compiled twice into the plugin: with WIND_PLUGIN_PART=1 (exact
// mode, block and lockstep kernels, the descriptor; no FMA contraction) and
// WIND_PLUGIN_PART=2 (the fast-mode lane kernels, with FMA contraction).
#pragma once
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

extern "C" __attribute__((visibility("hidden"))) int wind_plugin_launch_fast(const IntegrateArgs *a, int method,
                                                                          int num_cu, float share, int grid,
                                                                          hipStream_t st);

template <class Net, int METHOD>
int plugin_launch_method(const IntegrateArgs &a, int mode, int kernel, int num_cu, float share, int grid_cap,
                         hipStream_t st) {
    typename Net::Params p{};
    int grid = 0, rc = WIND_OK;
    if (kernel == WIND_KERNEL_BLOCK) {
        if (mode != WIND_MODE_EXACT) return WIND_ERR_UNSUPPORTED;
        SolverChoice sc;
        sc.share = share;
        sc.grid = grid_cap;
        grid = scaled_grid(num_cu, 32, sc, a.nsys);
        hipLaunchKernelGGL((block_integrate<Net, METHOD>), dim3(grid), dim3(64), 0, st, a, p);
    } else if (mode == WIND_MODE_EXACT) {
        auto k = lane_integrate<Net, METHOD, MathExact>;
        rc = plugin_grid(k, 256, a.nsys, num_cu, share, grid_cap, &grid);
        if (rc) return rc;
        hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, st, a, p);
    } else {
        return wind_plugin_launch_fast(&a, METHOD, num_cu, share, grid_cap, st);
    }
    return hipGetLastError() == hipSuccess ? WIND_OK : WIND_ERR_DEVICE;
}

template <class Net, int METHOD>
int plugin_launch_fast_method(const IntegrateArgs &a, int num_cu, float share, int grid_cap, hipStream_t st) {
    typename Net::Params p{};
    int grid = 0;
    auto k = lane_integrate<Net, METHOD, MathFast>;
    const int rc = plugin_grid(k, 256, a.nsys, num_cu, share, grid_cap, &grid);
    if (rc) return rc;
    hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, st, a, p);
    return hipGetLastError() == hipSuccess ? WIND_OK : WIND_ERR_DEVICE;
}

template <class Net>
int plugin_launch(const IntegrateArgs *a, int method, int mode, int kernel, int num_cu, float share, int grid,
                  hipStream_t st) {
    if (a->trace.capacity > 0) return WIND_ERR_UNSUPPORTED;  // no traced kernels in plugins
    if (method == WIND_SIE) return plugin_launch_method<Net, WIND_SIE>(*a, mode, kernel, num_cu, share, grid, st);
    if (method == WIND_RK2) return plugin_launch_method<Net, WIND_RK2>(*a, mode, kernel, num_cu, share, grid, st);
    return WIND_ERR_ARGUMENT;
}

template <class Net>
int plugin_lockstep(float tnow, float hh, const float *C, int stride, const float *Y, float *Y2, int nsys,
                    float ABS, float REL, int *flag, int grid) {
    typename Net::Params p{};
    hipLaunchKernelGGL(lockstep_attempt<Net>, dim3(grid), dim3(256), 0, 0, tnow, hh, C, stride, Y, Y2, nsys, ABS,
                       REL, flag, p);
    return hipGetLastError() == hipSuccess ? WIND_OK : WIND_ERR_DEVICE;
}

}  // namespace wind

#if WIND_PLUGIN_PART == 1
#define WIND_DEFINE_PLUGIN(NET, NAME, NREACT, FLOPS_RHS, FLOPS_JAC)                                      \
    extern "C" __attribute__((visibility("default"))) const wind_plugin_desc *wind_plugin_get(void) {     \
        static const wind_plugin_desc d = {WIND_PLUGIN_ABI,       NET::N,                                  \
                                           NET::NCONST,           NREACT,                                  \
                                           NAME,                  FLOPS_RHS,                               \
                                           FLOPS_JAC,             wind::plugin_launch<NET>,                \
                                           wind::plugin_lockstep<NET>};                                    \
        return &d;                                                                                         \
    }
#else
#undef WIND_DEFINE_PLUGIN
#define WIND_DEFINE_PLUGIN(NET, NAME, NREACT, FLOPS_RHS, FLOPS_JAC)                                      \
    extern "C" __attribute__((visibility("hidden"))) int wind_plugin_launch_fast(                         \
        const wind::IntegrateArgs *a, int method, int num_cu, float share, int grid, hipStream_t st) {       \
        if (method == WIND_SIE)                                                                            \
            return wind::plugin_launch_fast_method<NET, WIND_SIE>(*a, num_cu, share, grid, st);           \
        if (method == WIND_RK2)                                                                            \
            return wind::plugin_launch_fast_method<NET, WIND_RK2>(*a, num_cu, share, grid, st);           \
        return WIND_ERR_ARGUMENT;                                                                          \
    }
#endif
