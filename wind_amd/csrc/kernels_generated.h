// One-workgroup-per-system integrator for generated networks of any size
// (wind_amd/codegen.py): the launch shape of the reference's fused kernel
// (cuda/common/kernel.cu:140-310, one block per system, N < 1024 by
// cuda/common/harness.cu:62-69), with the network evaluated tile-parallel
// (thread t evaluates the RHS and Jacobian block of tiles t, t + T, ...; the
// reference's Ntile copies are independent blocks) and the matrices in LDS
// when they fit, else in a per-workgroup slice of a device scratch buffer.
// Exact mode runs the reference's Gauss-Jordan inverse and mat-vec
// (block_gje_exact), fast mode an LU solve (block_lu_fast); the step control
// is block_integrate_systems'.  The generated code loops over tiles, so the
// kernel's size does not grow with N.
#pragma once
#include <type_traits>

#include "kernels_mass_action.h"

namespace wind {

// Network evaluation of a generated network: constants in LDS, f by tiles,
// J by tiles onto zeros, then M = I + nh J as the reference's
// scaleAndInvertJacobians forms it (sie_gold.c:17-36: every element scaled,
// then the diagonal incremented).
template <class Net, class M>
struct GenEval {
    typename Net::Params p;
    float *cc;  // LDS [NCONST]
    __device__ void load(const IntegrateArgs &a, int sid) {
        for (int k = threadIdx.x; k < Net::NCONST; k += blockDim.x) cc[k] = a.constants[(size_t)sid * a.const_stride + k];
    }
    __device__ void eval(const float *yv, MALayout &L, float nh, bool want_jac) {
        constexpr int N = Net::N;
        const int t = threadIdx.x, T = blockDim.x;
        for (int tt = t; tt < Net::NTILE; tt += T) Net::template rhs_tile<M>(p, cc, yv, L.f, tt);
        if (!want_jac) return;
        for (int e = t; e < N * N; e += T) L.M[e] = 0.0f;
        __syncthreads();
        for (int tt = t; tt < Net::NTILE; tt += T) Net::template jac_tile<M>(p, cc, yv, L.M, tt);
        __syncthreads();
        for (int e = t; e < N * N; e += T) L.M[e] = nh * L.M[e];
        __syncthreads();
        for (int i = t; i < N; i += T) L.M[i * (N + 1)] = L.M[i * (N + 1)] + 1.0f;
        __syncthreads();
    }
};

// Floats of LDS before the matrices: y, yw, y1, f, g and the constants.
template <class Net>
__host__ __device__ constexpr int gen_block_base_floats() {
    return 5 * Net::N + ((Net::NCONST + 3) & ~3);
}

template <class Net, int METHOD, bool FAST, bool PIVOT = false>
__global__ void __launch_bounds__(256) gen_block_integrate(IntegrateArgs a, typename Net::Params p, float *gscratch,
                                                           int mat_in_lds) {
    extern __shared__ float lds[];
    constexpr int N = Net::N;
    MALayout L;
    L.y = lds;
    L.yw = L.y + N;
    L.y1 = L.yw + N;
    L.f = L.y1 + N;
    L.g = L.f + N;
    L.w = L.da = L.db = nullptr;
    float *cc = L.g + N;
    constexpr size_t per_block = (size_t)(FAST ? 1 : 2) * N * N;
    float *mats = mat_in_lds ? lds + gen_block_base_floats<Net>() : gscratch + (size_t)blockIdx.x * per_block;
    L.M = mats;
    L.Inv = FAST ? nullptr : mats + (size_t)N * N;
    using Ev = GenEval<Net, typename std::conditional<FAST, MathFast, MathExact>::type>;
    Ev ev{p, cc};
    block_integrate_systems<METHOD, FAST, Ev, PIVOT>(a, N, L, ev);
}

// Launch shape of gen_block_integrate: threads, grid, LDS bytes, and the
// scratch bytes when the matrices do not fit in LDS (0 otherwise).
struct GenBlockShape {
    int T = 0, grid = 0;
    size_t lds = 0, scratch = 0;
    bool in_lds = true;
};

template <class Net, int METHOD, bool FAST, bool PIVOT = false>
inline int gen_block_shape(int nsys, int num_cu, float share, int grid_cap, GenBlockShape *s) {
    constexpr int N = Net::N;
    int dev = 0, max_lds = 0;
    if (hipGetDevice(&dev) != hipSuccess) return WIND_ERR_DEVICE;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return WIND_ERR_DEVICE;
    const size_t base = sizeof(float) * (size_t)gen_block_base_floats<Net>();
    const size_t mats = sizeof(float) * (FAST ? 1 : 2) * (size_t)N * N;
    s->T = block_threads(N, nsys, num_cu);
    s->in_lds = base + mats + 256 <= (size_t)max_lds;
    s->lds = base + (s->in_lds ? mats : 0);
    auto kfn = gen_block_integrate<Net, METHOD, FAST, PIVOT>;
    if (s->lds > 65536 &&
        hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s->lds) != hipSuccess)
        return WIND_ERR_DEVICE;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, s->T, s->lds) != hipSuccess) return WIND_ERR_DEVICE;
    SolverChoice sc;
    sc.share = share;
    sc.grid = grid_cap;
    s->grid = scaled_grid(num_cu, per_cu, sc, nsys);
    s->scratch = s->in_lds ? 0 : mats * (size_t)s->grid;
    return WIND_OK;
}

template <class Net, int METHOD, bool FAST, bool PIVOT = false>
inline int gen_block_launch(const IntegrateArgs &a, const typename Net::Params &p, int num_cu, float share,
                            int grid_cap, float *scratch, size_t scratch_bytes, hipStream_t st) {
    GenBlockShape s;
    int rc = gen_block_shape<Net, METHOD, FAST, PIVOT>(a.nsys, num_cu, share, grid_cap, &s);
    if (rc) return rc;
    if (s.scratch > 0 && (!scratch || scratch_bytes < s.scratch)) return WIND_ERR_ARGUMENT;
    hipLaunchKernelGGL((gen_block_integrate<Net, METHOD, FAST, PIVOT>), dim3(s.grid), dim3(s.T), s.lds, st, a, p,
                       s.in_lds ? nullptr : scratch, s.in_lds ? 1 : 0);
    return hipGetLastError() == hipSuccess ? WIND_OK : WIND_ERR_DEVICE;
}

}  // namespace wind
