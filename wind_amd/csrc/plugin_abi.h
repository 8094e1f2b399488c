// Contract between the core library and network plugins written by the
// code generator (wind_amd/codegen.py).  A plugin is a shared library built
// from a generated network header and plugin_impl.h; it exports
// wind_plugin_get(), which returns a static descriptor.  Core and plugins are
// built from the same sources, so IntegrateArgs crosses the boundary as is;
// WIND_PLUGIN_ABI guards against mixing builds.
#pragma once
#include <hip/hip_runtime.h>

#include "wind_types.h"

#define WIND_PLUGIN_ABI 4

struct wind_plugin_desc {
    int abi;          // WIND_PLUGIN_ABI
    int N;            // species
    int nconst;       // constants per system
    int nreactions;
    const char *name;
    long long flops_rhs, flops_jac;  // arithmetic operations per evaluation
    // Enqueues the per-system integrator on `st` (device pointers in *a;
    // a->counters zeroed).  share / grid as in the core's batch scheduling
    // (0 = whole device).  `scratch` holds at least scratch_bytes() of device
    // memory when that is non-zero.  Returns a wind_error code.
    int (*launch)(const wind::IntegrateArgs *a, int method, int mode, int kernel, int num_cu, float share,
                  int grid, float *scratch, size_t scratch_bytes, hipStream_t st);
    // One lockstep SIE attempt over all systems (see kernels_lockstep.h).
    int (*lockstep_attempt)(float tnow, float hh, const float *C, int stride, const float *Y, float *Y2, int nsys,
                            float ABS, float REL, int *flag, int grid);
    // Device scratch the launch with these arguments needs (the matrices of
    // the one-workgroup-per-system kernel when they exceed LDS), or 0.
    int (*scratch_bytes)(int nsys, int method, int mode, int kernel, int num_cu, float share, int grid,
                         size_t *bytes);
};

typedef const wind_plugin_desc *(*wind_plugin_get_fn)(void);
