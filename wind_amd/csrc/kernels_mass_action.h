// Integrator kernels for table-driven mass-action networks (any N up to
// kMassActionMaxN): one workgroup per system, state and network values in
// LDS, the matrix in LDS when it fits and otherwise in a per-workgroup slice
// of a device scratch buffer.
//
// Network evaluation is a gather: each reaction's rate w_r = ((k_r s) y_a) y_b
// and partials da_r = (k_r s) y_b, db_r = (k_r s) y_a are computed once (one
// thread per reaction), then each species' f_i and each Jacobian non-zero is
// summed from an ordered contribution list built on the host.  The order of
// every list is the order in which a serial loop over reactions (ascending r,
// updating a, b, c, d; columns a then b) touches that element, so each sum
// rounds exactly like the serial loop.
//
// Exact mode runs the reference's Gauss-Jordan inverse
// (cuda/c_baseline/linear_algebra/gje_gold.c:56-144) with the updates of one
// pivot spread over (row, column) pairs - every element still sees the same
// sequence of operations - followed by the transposed mat-vec
// (matrix_ops.c:6-27).  Fast mode factors (I - hJ) by unpivoted LU with
// fused multiply-adds and substitutes.
#pragma once
#include "wind_device.h"
#include "wind_internal.h"

namespace wind {

struct MALayout {
    float *y, *yw, *y1, *f, *g, *w, *da, *db;
    float *M, *Inv;
};

__device__ inline void ma_eval(const MassActionDevice &net, const float *yv, float scale, MALayout &L,
                               float nh, bool want_jac) {
    const int N = net.N, R = net.R, T = blockDim.x, t = threadIdx.x;
    for (int r = t; r < R; r += T) {
        const int a = net.rx[4 * r], b = net.rx[4 * r + 1];
        const float ks = net.k[r] * scale;
        const float ya = yv[a], yb = yv[b];
        L.w[r] = (ks * ya) * yb;
        L.da[r] = ks * yb;
        L.db[r] = ks * ya;
    }
    __syncthreads();
    for (int i = t; i < N; i += T) {
        float acc = 0.0f;
        for (int e = net.f_ptr[i]; e < net.f_ptr[i + 1]; e++) {
            const int ent = net.f_ent[e];
            const float v = L.w[ent >> 1];
            acc = (ent & 1) ? acc - v : acc + v;
        }
        L.f[i] = acc;
    }
    if (!want_jac) return;
    // M = (-h) * J with J zero outside its non-zeros (sie_gold.c:17-36),
    // then the diagonal + 1 in a second pass.
    const float zs = nh * 0.0f;
    for (int e = t; e < N * N; e += T) L.M[e] = zs;
    __syncthreads();
    for (int z = t; z < net.nnz; z += T) {
        float acc = 0.0f;
        for (int e = net.j_ptr[z]; e < net.j_ptr[z + 1]; e++) {
            const int ent = net.j_ent[e];
            const int r = ent >> 2;
            const float v = (ent & 2) ? L.db[r] : L.da[r];
            acc = (ent & 1) ? acc - v : acc + v;
        }
        L.M[net.j_pos[z]] = nh * acc;
    }
    __syncthreads();
    for (int i = t; i < N; i += T) L.M[i * (N + 1)] = L.M[i * (N + 1)] + 1.0f;
    __syncthreads();
}

// Exact Gauss-Jordan inverse and mat-vec; on return g = inverse * f.
// Threads of a one-workgroup-per-system launch: 64 (one wave) for N <= 64
// when the systems fill the device, four waves otherwise, so that a few
// systems (latency-bound) or large ones spread each elimination step over
// more lanes.  Results do not depend on it.
inline int block_threads(int N, int nsys, int num_cu) { return (N > 64 || nsys <= 4 * num_cu) ? 256 : 64; }

__device__ inline void block_gje_exact(int N, MALayout &L) {
    const int T = blockDim.x, t = threadIdx.x;
    // Elimination loops: groups of TX lanes (the power of two >= N, at most
    // 64) walk a row's columns, the groups walk the rows, so that no thread
    // divides by N.  Every element's arithmetic is the reference's whatever
    // thread does it.
    const int lg = N <= 8 ? 3 : N <= 16 ? 4 : N <= 32 ? 5 : 6, TX = 1 << lg;
    const int tx = t & (TX - 1), ty = t >> lg, TY = T >> lg;
    float *M = L.M, *Inv = L.Inv;
    for (int e = t; e < N * N; e += T) Inv[e] = (e % (N + 1) == 0) ? 1.0f : 0.0f;
    __syncthreads();
    for (int r = 0; r < N; r++) {
        const float s = M[r * N + r];
        for (int c = t; c < N; c += T) {
            if (c > r)
                M[r * N + c] = M[r * N + c] / s;
            else
                Inv[r * N + c] = Inv[r * N + c] / s;
        }
        __syncthreads();
        for (int q = r + 1 + ty; q < N; q += TY) {
            const float sq = M[q * N + r];
            for (int c = tx; c < N; c += TX) {
                if (c > r)
                    M[q * N + c] = M[q * N + c] - sq * M[r * N + c];
                else
                    Inv[q * N + c] = Inv[q * N + c] - sq * Inv[r * N + c];
            }
        }
        __syncthreads();
    }
    for (int r = 0; r < N; r++) {
        const int b = N - 1 - r;
        // rows bq < b, processed in the reference's order bq = b-1, b-2, ...;
        // the rows are independent, so any assignment to threads is exact.
        for (int bq = ty; bq < b; bq += TY) {
            const float m = M[bq * N + b];
            for (int c = tx; c < N; c += TX) Inv[bq * N + c] = Inv[bq * N + c] - m * Inv[b * N + c];
        }
        __syncthreads();
    }
    for (int r = t; r < N; r += T) {
        float acc = 0.0f;
        for (int c = 0; c < N; c++) acc = acc + Inv[c * N + r] * L.f[c];
        L.g[r] = acc;
    }
    __syncthreads();
}

// Row of the largest |A(i, k)|, i >= k (the first of equals), searched by
// wave 0: each lane scans rows k + lane, k + lane + 64, ..., then the wave
// reduces (value, row) pairs with shuffles.  Written to *piv.
__device__ inline void pivot_search(const float *M, int N, int k, int *piv) {
    if (threadIdx.x >= 64) return;
    float best = -1.0f;
    int row = k;
    for (int i = k + (int)threadIdx.x; i < N; i += 64) {
        const float v = fabsf(M[k * N + i]);
        if (v > best) {
            best = v;
            row = i;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float ob = __shfl_xor(best, off);
        const int orow = __shfl_xor(row, off);
        if (ob > best || (ob == best && orow < row)) {
            best = ob;
            row = orow;
        }
    }
    if (threadIdx.x == 0) *piv = row;
}

// Fast LU solve of A x = f, A[i][j] = M[j*N + i]; x in L.g.  Unpivoted like
// the reference's Gauss-Jordan, or (PIVOT) with partial pivoting: before
// eliminating column k, rows k and the pivot_search row exchange their
// entries in columns k.. and in the right-hand side.
template <bool PIVOT = false>
__device__ inline void block_lu_fast(int N, MALayout &L) {
    const int T = blockDim.x, t = threadIdx.x;
    float *M = L.M;
    float *b = L.g;
    __shared__ int s_piv;
    for (int i = t; i < N; i += T) b[i] = L.f[i];
    __syncthreads();
    for (int k = 0; k < N; k++) {
        if constexpr (PIVOT) {
            pivot_search(M, N, k, &s_piv);
            __syncthreads();
            const int p = s_piv;
            if (p != k) {
                for (int j = k + t; j <= N; j += T) {
                    float *a = j < N ? M + (size_t)j * N : b;
                    const float v = a[k];
                    a[k] = a[p];
                    a[p] = v;
                }
                __syncthreads();
            }
        }
        const float ip = __builtin_amdgcn_rcpf(M[k * N + k]);
        for (int i = k + 1 + t; i < N; i += T) M[k * N + i] = M[k * N + i] * ip;  // l_i
        __syncthreads();
        const int rows = N - 1 - k;
        // trailing update A(i,j) -= l_i A(k,j) for i, j > k; b_i -= l_i b_k
        for (int idx = t; idx < rows * (rows + 1); idx += T) {
            const int jj = idx / rows, i = k + 1 + idx % rows;
            const float l = M[k * N + i];
            if (jj < rows) {
                const int j = k + 1 + jj;
                M[j * N + i] = fmaf(-l, M[j * N + k], M[j * N + i]);
            } else {
                b[i] = fmaf(-l, b[k], b[i]);
            }
        }
        __syncthreads();
    }
    for (int j = N - 1; j >= 0; j--) {
        if (t == 0) b[j] = b[j] * __builtin_amdgcn_rcpf(M[j * N + j]);
        __syncthreads();
        const float xj = b[j];
        for (int i = t; i < j; i += T) b[i] = fmaf(-M[j * N + i], xj, b[i]);
        __syncthreads();
    }
}

// The per-system loop of the one-workgroup-per-system kernels: each
// workgroup takes systems blockIdx.x, blockIdx.x + gridDim.x, ... and runs
// the step-doubling attempts of common_gold.c:77-148 with the solver of the
// mode.  `ev` evaluates the network: ev.load(a, sid) reads the system's
// constants, ev.eval(yv, L, nh, want_jac) writes f(yv) to L.f and, with
// want_jac, L.M = I + nh J(yv) (column-major), ending with a barrier.
template <int METHOD, bool FAST, class Eval, bool PIVOT = false>
__device__ inline void block_integrate_systems(const IntegrateArgs &a, int N, MALayout &L, Eval &ev) {
    const int t = threadIdx.x, T = blockDim.x;
    __shared__ int s_flag;
    __shared__ double s_err[4];
    unsigned long long steps = 0;
    for (int sid = blockIdx.x; sid < a.nsys; sid += gridDim.x) {
        for (int i = t; i < N; i += T) L.y[i] = a.equations[(size_t)sid * N + i];
        ev.load(a, sid);
        __syncthreads();
        float tnow = a.tnow, h = a.h0;
        int unsolved = 0, attempts = 0, status = WIND_STATUS_OK;
        while (tnow < a.tend) {
            const float hh = fminf(a.tend - tnow, h);
            const float tb = tnow + hh;
            const float hs_full = tb - tnow, hs_half = (tb - tnow) / 2.0f;
            for (int pass = 0; pass < 3; pass++) {
                const float hs = pass == 0 ? hs_full : hs_half;
                if (pass != 2) {
                    for (int i = t; i < N; i += T) L.yw[i] = L.y[i];
                    __syncthreads();
                }
                ev.eval(L.yw, L, -hs, METHOD == WIND_SIE);
                if (METHOD == WIND_SIE) {
                    if (FAST)
                        block_lu_fast<PIVOT>(N, L);
                    else
                        block_gje_exact(N, L);
                    for (int i = t; i < N; i += T)
                        L.yw[i] = FAST ? fmaf(hs, L.g[i], L.yw[i]) : L.yw[i] + hs * L.g[i];
                } else {
                    __syncthreads();
                    for (int i = t; i < N; i += T)
                        L.yw[i] = FAST ? fmaf(hs, L.f[i], L.yw[i]) : L.yw[i] + hs * L.f[i];
                }
                __syncthreads();
                if (pass == 0)
                    for (int i = t; i < N; i += T) L.y1[i] = L.yw[i];
            }
            if (t == 0) s_flag = 0;
            __syncthreads();
            bool fail = false;
            double err = 0.0;
            if (a.controller == WIND_CTRL_REFERENCE) {
                for (int i = t; i < N; i += T)
                    fail = fail || species_fails<METHOD>(L.y1[i], L.yw[i], a.ABS, a.REL);
                if (fail) s_flag = 1;
            } else {
                for (int i = t; i < N; i += T) err = err_max(err, err_term(L.y1[i], L.yw[i], a.ABS, a.REL));
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) err = err_max(err, __shfl_xor(err, off));
                if (__lane_id() == 0) s_err[t >> 6] = err;
            }
            __syncthreads();
            const bool flag = s_flag != 0;
            if (a.controller != WIND_CTRL_REFERENCE) {
                err = s_err[0];
                for (int wv = 1; wv < (T + 63) / 64; wv++) err = err_max(err, s_err[wv]);
            }
            attempts++;
            const bool acc = control_step<METHOD>(a.controller, a.tend, hh, flag, err, tnow, h, unsolved, status);
            if (acc) {
                const bool extrap = METHOD == WIND_RK2 || a.controller != WIND_CTRL_REFERENCE;
                for (int i = t; i < N; i += T) L.y[i] = extrap ? 2.0f * L.yw[i] - L.y1[i] : L.yw[i];
            }
            __syncthreads();
            if (status != WIND_STATUS_OK) break;
            if (attempts >= a.max_attempts && tnow < a.tend) {
                status = WIND_STATUS_EXHAUSTED;
                break;
            }
        }
        __syncthreads();
        for (int i = t; i < N; i += T) a.equations[(size_t)sid * N + i] = L.y[i];
        if (t == 0) {
            if (a.steps_out) a.steps_out[sid] = 3 * attempts;
            if (a.status_out) a.status_out[sid] = status;
            steps += 3ull * (unsigned long long)attempts;
        }
        __syncthreads();
    }
    if (t == 0 && steps) atomicAdd(a.counters + 1, steps);
}

// Network evaluation of the table-driven mass-action networks (one constant
// per system: the rate scale).
struct MassActionEval {
    const MassActionDevice &net;
    float scale;
    __device__ void load(const IntegrateArgs &a, int sid) { scale = a.constants[(size_t)sid * a.const_stride]; }
    __device__ void eval(const float *yv, MALayout &L, float nh, bool want_jac) {
        ma_eval(net, yv, scale, L, nh, want_jac);
    }
};

template <int METHOD, bool FAST, bool PIVOT = false>
__global__ void __launch_bounds__(256) ma_block_integrate(IntegrateArgs a, MassActionDevice net,
                                                          float *gscratch, int mat_in_lds) {
    extern __shared__ float lds[];
    const int N = net.N, R = net.R;
    MALayout L;
    L.y = lds;
    L.yw = L.y + N;
    L.y1 = L.yw + N;
    L.f = L.y1 + N;
    L.g = L.f + N;
    L.w = L.g + N;
    L.da = L.w + R;
    L.db = L.da + R;
    float *mats = mat_in_lds ? (L.db + R) : gscratch + (size_t)blockIdx.x * 2 * N * N;
    L.M = mats;
    L.Inv = mats + N * N;
    MassActionEval ev{net, 0.0f};
    block_integrate_systems<METHOD, FAST, MassActionEval, PIVOT>(a, N, L, ev);
}

}  // namespace wind
