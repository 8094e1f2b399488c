// Launchers of the dense LDS/MFMA solver (kernels_dense.h).  A translation
// unit of its own, compiled with fused multiply-add contraction: the dense
// solver exists only in fast mode, which promises no operation order.
#include <hip/hip_runtime.h>

#include "kernels_dense.h"
#include "kernels_dense_reg.h"
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

// ---------------------------------------------------------------------------
// Register-resident solver (kernels_dense_reg.h).

// Workgroups per CU the register solver runs with on the current device for
// network d (dense_reg_wgs_for), or 0 where it does not fit.
static int dense_reg_wgs_dev(const MassActionDevice &d) {
    int dev = 0, max_lds = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess)
        return 0;
    return dense_reg_wgs_for((d.N + 15) / 16, d.R, d.sc_cap, (size_t)max_lds);
}

bool dense_reg_eligible(const MassActionDevice &d, const SolverChoice &sc) {
    if (sc.dense_kernel == 1 || !d.sc_ent) return false;
    if (sc.pivot || sc.ir_sweeps != 0) return false;
    if (sc.solver != WIND_SOLVER_AUTO && sc.solver != WIND_SOLVER_MFMA_F32) return false;
    const int NT = (d.N + 15) / 16;
    if (!dense_reg_tiles(NT)) return false;
    // three (or four) workgroups per CU must fit, else the LDS-resident solver is the better shape
    return dense_reg_wgs_dev(d) != 0;
}

template <int NT, int WGS>
int launch_dense_reg_nt(int num_cu, const MassActionDevice &d, const IntegrateArgs &a, const SolverChoice &sc,
                        hipStream_t st) {
    auto kfn = a.trace.capacity > 0 ? ma_dense_reg_integrate<NT, WGS, true> : ma_dense_reg_integrate<NT, WGS>;
    const size_t lds = dense_reg_lds_bytes(NT, d.R, WGS, d.sc_cap);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    }
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, 256, lds);
    if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    const int grid = scaled_grid(num_cu, per_cu, sc, a.nsys);
    hipLaunchKernelGGL(kfn, dim3(grid), dim3(256), lds, st, a, d);
    e = hipGetLastError();
    if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    return WIND_OK;
}

int launch_dense_reg(int num_cu, const MassActionDevice &d, const IntegrateArgs &a, const SolverChoice &sc,
                     hipStream_t st) {
    switch ((d.N + 15) / 16) {
    case 2:
        return dense_reg_wgs_dev(d) == 8 ? launch_dense_reg_nt<2, 8>(num_cu, d, a, sc, st)
                                         : launch_dense_reg_nt<2, 3>(num_cu, d, a, sc, st);
    case 4:
        return dense_reg_wgs_dev(d) == 6 ? launch_dense_reg_nt<4, 6>(num_cu, d, a, sc, st)
                                         : launch_dense_reg_nt<4, 3>(num_cu, d, a, sc, st);
    case 5: return launch_dense_reg_nt<5, 3>(num_cu, d, a, sc, st);
    case 6: return launch_dense_reg_nt<6, 3>(num_cu, d, a, sc, st);
    case 7: return launch_dense_reg_nt<7, 3>(num_cu, d, a, sc, st);
    case 8: return launch_dense_reg_nt<8, 3>(num_cu, d, a, sc, st);
    case 9: return launch_dense_reg_nt<9, 3>(num_cu, d, a, sc, st);
    case 10:
        return dense_reg_wgs_dev(d) == 4 ? launch_dense_reg_nt<10, 4>(num_cu, d, a, sc, st)
                                         : launch_dense_reg_nt<10, 3>(num_cu, d, a, sc, st);
    default: return wind_fail(WIND_ERR_UNSUPPORTED, "register-resident dense solver: N out of range");
    }
}
