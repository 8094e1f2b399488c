// Launchers of the dense LDS/MFMA solver (kernels_dense.h).  A translation
// unit of its own, compiled with fused multiply-add contraction: the dense
// solver exists only in fast mode, which promises no operation order.
#include <hip/hip_runtime.h>

#include "kernels_dense.h"
#include "wind_internal.h"

using namespace wind;

template <int METHOD, int PREC>
int launch_dense(int num_cu, const MassActionDevice &d, const IntegrateArgs &a, int ir, size_t lds, int T,
                 const SolverChoice &sc, hipStream_t st) {
    auto kfn = a.trace.capacity > 0 ? ma_dense_integrate<METHOD, PREC, true> : ma_dense_integrate<METHOD, PREC>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    }
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, T, lds);
    if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    const int grid = scaled_grid(num_cu, per_cu, sc, a.nsys);
    hipLaunchKernelGGL(kfn, dim3(grid), dim3(T), lds, st, a, d, ir);
    e = hipGetLastError();
    if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    return WIND_OK;
}

#define WIND_DENSE(M, P)                                                                                     \
    template int launch_dense<M, P>(int, const MassActionDevice &, const IntegrateArgs &, int, size_t, int,    \
                                    const SolverChoice &, hipStream_t);
WIND_DENSE(WIND_SIE, PREC_VALU)
WIND_DENSE(WIND_SIE, PREC_MFMA_F32)
WIND_DENSE(WIND_SIE, PREC_MFMA_BF16)
WIND_DENSE(WIND_SIE, PREC_MFMA_BF16X3)
WIND_DENSE(WIND_RK2, PREC_VALU)
WIND_DENSE(WIND_RK2, PREC_MFMA_F32)
WIND_DENSE(WIND_RK2, PREC_MFMA_BF16)
WIND_DENSE(WIND_RK2, PREC_MFMA_BF16X3)
