// Fast-mode integrator for table-driven networks with larger N: one
// workgroup per system, the matrix A = I - hJ dense in LDS (row-major, odd
// row stride to spread LDS banks, padded with identity rows to a multiple of
// 16), factored by a blocked right-looking LU without pivoting (the
// reference's solve is unpivoted too, gje_gold.c:56-134).
//
// Per 16-column panel: the panel is factored in place (16 pivot steps), the
// 16-row block of U to its right is solved against the unit lower diagonal
// block, and the trailing matrix is updated A22 -= L21 * U12 in 16x16 tiles
// spread over the workgroup's waves.  The tile product runs on the matrix
// cores:
//   PREC_VALU      fused multiply-adds on the vector ALU (no MFMA), reference point
//   PREC_MFMA_F32  v_mfma_f32_16x16x4f32, four per tile (K = 16)
//   PREC_MFMA_BF16 v_mfma_f32_16x16x16bf16_1k, one per tile: BF16 inputs, FP32 accumulation
//   PREC_MFMA_BF16X3 three BF16 products (hi*hi + hi*lo + lo*hi) of the split operands
// The 16x16 diagonal block of each panel is factored in one wave with
// register broadcasts; the rows below it are solved against U11 in parallel.
// Forward and back substitution go by 16-row blocks.  Optional iterative
// refinement recomputes the residual r = f - (I - hJ)x in double (J x
// gathered row by row) and solves for a correction with the same factors.
// Every step is deterministic: a system's result does not depend on what
// else runs on the device.
#pragma once
#include "wind_device.h"
#include "wind_internal.h"

namespace wind {

enum DensePrec { PREC_VALU = 0, PREC_MFMA_F32 = 1, PREC_MFMA_BF16 = 2, PREC_MFMA_BF16X3 = 3 };

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short bf16x4_t __attribute__((ext_vector_type(4)));

__device__ inline short to_bf16_rne(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ inline float from_bf16(short h) { return __uint_as_float(((uint32_t)(uint16_t)h) << 16); }

// C tile (16x16) of the trailing update: rows r0.., columns c0.., K = the
// 16 panel columns k0..k0+15.  C -= L21 * U12.
template <int PREC>
__device__ inline void tile_update(float *A, int S, int r0, int c0, int k0, int lane) {
    const int lr = lane & 15, lq = lane >> 4;
    f32x4 c;
#pragma unroll
    for (int r = 0; r < 4; r++) c[r] = A[(r0 + 4 * lq + r) * S + c0 + lr];
    if (PREC == PREC_VALU) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = r0 + 4 * lq + r;
            float acc = c[r];
#pragma unroll
            for (int k = 0; k < 16; k++) acc = fmaf(-A[row * S + k0 + k], A[(k0 + k) * S + c0 + lr], acc);
            c[r] = acc;
        }
    } else if (PREC == PREC_MFMA_F32) {
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
            const float av = -A[(r0 + lr) * S + k0 + kk + lq];
            const float bv = A[(k0 + kk + lq) * S + c0 + lr];
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, c, 0, 0, 0);
        }
    } else {
        bf16x4_t ah, bh, al, bl;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float av = -A[(r0 + lr) * S + k0 + 4 * lq + q];
            const float bv = A[(k0 + 4 * lq + q) * S + c0 + lr];
            ah[q] = to_bf16_rne(av);
            bh[q] = to_bf16_rne(bv);
            if (PREC == PREC_MFMA_BF16X3) {
                al[q] = to_bf16_rne(av - from_bf16(ah[q]));
                bl[q] = to_bf16_rne(bv - from_bf16(bh[q]));
            }
        }
        if (PREC == PREC_MFMA_BF16X3) {
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, c, 0, 0, 0);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, c, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) A[(r0 + 4 * lq + r) * S + c0 + lr] = c[r];
}

// In-place LU (unit lower L below the diagonal, U on and above) of the
// NP x NP matrix at A with row stride S.  Called by the whole workgroup.
template <int PREC>
__device__ inline void dense_lu(float *A, int NP, int S) {
    const int t = threadIdx.x, T = blockDim.x;
    const int wave = t >> 6, lane = t & 63, nw = T >> 6;
    for (int k0 = 0; k0 < NP; k0 += 16) {
        const int kend = k0 + 16;
        // Diagonal block: LU of the 16x16 block in lanes 0-15 of wave 0, lane
        // r holding row k0 + r; pivot rows are broadcast with readlane.
        if (t < 16) {
            float a[16];
#pragma unroll
            for (int j = 0; j < 16; j++) a[j] = A[(k0 + t) * S + k0 + j];
#pragma unroll
            for (int kk = 0; kk < 15; kk++) {
                float pr[16];
#pragma unroll
                for (int j = kk; j < 16; j++) pr[j] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a[j]), kk));
                if (t > kk) {
                    const float l = a[kk] / pr[kk];
                    a[kk] = l;
#pragma unroll
                    for (int j = kk + 1; j < 16; j++) a[j] = fmaf(-l, pr[j], a[j]);
                }
            }
#pragma unroll
            for (int j = 0; j < 16; j++) A[(k0 + t) * S + k0 + j] = a[j];
        }
        __syncthreads();
        // Rows below the diagonal block: solve l U11 = a for each row (L21 = A21 U11^-1).
        for (int i = kend + t; i < NP; i += T) {
            float a[16];
#pragma unroll
            for (int j = 0; j < 16; j++) a[j] = A[i * S + k0 + j];
#pragma unroll
            for (int kk = 0; kk < 16; kk++) {
                const float l = a[kk] / A[(k0 + kk) * S + k0 + kk];
                a[kk] = l;
#pragma unroll
                for (int j = kk + 1; j < 16; j++) a[j] = fmaf(-l, A[(k0 + kk) * S + k0 + j], a[j]);
            }
#pragma unroll
            for (int j = 0; j < 16; j++) A[i * S + k0 + j] = a[j];
        }
        __syncthreads();
        if (kend >= NP) break;
        // U12 = L11^{-1} A12, one column per thread.
        for (int j = kend + t; j < NP; j += T) {
            float u[16];
#pragma unroll
            for (int k = 0; k < 16; k++) u[k] = A[(k0 + k) * S + j];
#pragma unroll
            for (int k = 0; k < 16; k++)
#pragma unroll
                for (int r = k + 1; r < 16; r++) u[r] = fmaf(-A[(k0 + r) * S + k0 + k], u[k], u[r]);
#pragma unroll
            for (int k = 1; k < 16; k++) A[(k0 + k) * S + j] = u[k];
        }
        __syncthreads();
        const int nt = (NP - kend) >> 4;
        for (int tile = wave; tile < nt * nt; tile += nw)
            tile_update<PREC>(A, S, kend + ((tile / nt) << 4), kend + ((tile % nt) << 4), k0, lane);
        __syncthreads();
    }
}

// Solves L U x = b by 16-row blocks: the 16x16 diagonal block in lanes 0-15
// of wave 0 (register shuffles at compile-time lanes), then a parallel
// update of the remaining rows.  b and x may alias; both in LDS.
__device__ inline void dense_lu_solve(const float *A, int NP, int S, const float *b, float *x) {
    const int t = threadIdx.x, T = blockDim.x;
    if (x != b)
        for (int i = t; i < NP; i += T) x[i] = b[i];
    __syncthreads();
    for (int k0 = 0; k0 < NP; k0 += 16) {
        if (t < 16) {
            float xi = x[k0 + t];
            float l[16];
#pragma unroll
            for (int kk = 0; kk < 16; kk++) l[kk] = A[(k0 + t) * S + k0 + kk];
#pragma unroll
            for (int kk = 0; kk < 15; kk++) {
                const float xk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xi), kk));
                if (t > kk) xi = fmaf(-l[kk], xk, xi);
            }
            x[k0 + t] = xi;
        }
        __syncthreads();
        for (int i = k0 + 16 + t; i < NP; i += T) {
            float acc = x[i];
#pragma unroll
            for (int kk = 0; kk < 16; kk++) acc = fmaf(-A[i * S + k0 + kk], x[k0 + kk], acc);
            x[i] = acc;
        }
        __syncthreads();
    }
    for (int k0 = NP - 16; k0 >= 0; k0 -= 16) {
        if (t < 16) {
            float xi = x[k0 + t];
            float u[16];
#pragma unroll
            for (int kk = 0; kk < 16; kk++) u[kk] = A[(k0 + t) * S + k0 + kk];
#pragma unroll
            for (int kk = 15; kk >= 0; kk--) {
                if (t == kk) xi = xi / u[kk];
                const float xk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xi), kk));
                if (t < kk) xi = fmaf(-u[kk], xk, xi);
            }
            x[k0 + t] = xi;
        }
        __syncthreads();
        for (int i = t; i < k0; i += T) {
            float acc = x[i];
#pragma unroll
            for (int kk = 0; kk < 16; kk++) acc = fmaf(-A[i * S + k0 + kk], x[k0 + kk], acc);
            x[i] = acc;
        }
        __syncthreads();
    }
}

struct DenseLayout {
    float *y, *yw, *y1, *f, *x, *r, *A, *w, *hda, *hdb;
    int *ra, *rb;
    int NP, S;
};

__device__ inline int dense_np(int N) { return (N + 15) & ~15; }

// f = rhs(yv) and A = I - hJ (dense, padded).  Phase 1: one thread per
// reaction stores w = k s y_a y_b, h*da = h k s y_b, h*db = h k s y_a and the
// reactant indices in LDS.  Phase 2: one thread per species i walks the
// reactions that touch i (the ordered (reaction, sign) list of f_i) and
// accumulates f_i and row i of A; a thread writes only its own row, so the
// result does not depend on scheduling.
__device__ inline void dense_eval(const MassActionDevice &net, const float *yv, float scale, DenseLayout &L,
                                  float hs, bool want_jac) {
    const int N = net.N, R = net.R, T = blockDim.x, t = threadIdx.x, NP = L.NP, S = L.S;
    for (int r = t; r < R; r += T) {
        const int4 x = reinterpret_cast<const int4 *>(net.rx)[r];
        const float ks = net.k[r] * scale;
        const float ya = yv[x.x], yb = yv[x.y];
        L.w[r] = (ks * ya) * yb;
        L.hda[r] = hs * (ks * yb);
        L.hdb[r] = hs * (ks * ya);
        L.ra[r] = x.x;
        L.rb[r] = x.y;
    }
    __syncthreads();
    for (int i = t; i < NP; i += T) {
        float *row = L.A + i * S;
        if (want_jac)
            for (int j = 0; j < NP; j++) row[j] = (i == j) ? 1.0f : 0.0f;
        float acc = 0.0f;
        if (i < N)
            for (int e = net.f_ptr[i]; e < net.f_ptr[i + 1]; e++) {
                const int ent = net.f_ent[e];
                const int r = ent >> 1;
                const bool neg = ent & 1;
                acc = neg ? acc - L.w[r] : acc + L.w[r];
                if (want_jac) {
                    // J[i][a] += sign * da, J[i][b] += sign * db; A = I - hJ
                    const float ha = L.hda[r], hb = L.hdb[r];
                    row[L.ra[r]] = neg ? row[L.ra[r]] + ha : row[L.ra[r]] - ha;
                    row[L.rb[r]] = neg ? row[L.rb[r]] + hb : row[L.rb[r]] - hb;
                }
            }
        L.f[i] = acc;
    }
    __syncthreads();
}

// One refinement sweep: r = f - (x - h J x) in double, J x gathered by row
// from the same lists, then x += (LU)^{-1} r.
__device__ inline void dense_refine(const MassActionDevice &net, const float *yv, float scale, DenseLayout &L,
                                    float hs) {
    const int N = net.N, T = blockDim.x, t = threadIdx.x, NP = L.NP;
    for (int i = t; i < NP; i += T) {
        double jx = 0.0;
        if (i < N)
            for (int e = net.f_ptr[i]; e < net.f_ptr[i + 1]; e++) {
                const int ent = net.f_ent[e];
                const int r = ent >> 1;
                const int a = L.ra[r], b = L.rb[r];
                const double ks = (double)(net.k[r] * scale);
                const double g = ks * ((double)yv[b] * (double)L.x[a] + (double)yv[a] * (double)L.x[b]);
                jx = (ent & 1) ? jx - g : jx + g;
            }
        L.r[i] = (float)((double)L.f[i] - (double)L.x[i] + (double)hs * jx);
    }
    __syncthreads();
    dense_lu_solve(L.A, L.NP, L.S, L.r, L.r);
    for (int i = t; i < NP; i += T) L.x[i] = L.x[i] + L.r[i];
    __syncthreads();
}

template <int METHOD, int PREC>
__global__ void __launch_bounds__(256) ma_dense_integrate(IntegrateArgs a, MassActionDevice net, int ir_sweeps) {
    extern __shared__ float lds[];
    const int N = net.N, t = threadIdx.x, T = blockDim.x;
    DenseLayout L;
    L.NP = dense_np(N);
    L.S = L.NP + 1;
    const int NP = L.NP;
    const int R = net.R;
    L.y = lds;
    L.yw = L.y + NP;
    L.y1 = L.yw + NP;
    L.f = L.y1 + NP;
    L.x = L.f + NP;
    L.r = L.x + NP;
    L.w = L.r + NP;
    L.hda = L.w + R;
    L.hdb = L.hda + R;
    L.ra = reinterpret_cast<int *>(L.hdb + R);
    L.rb = L.ra + R;
    L.A = reinterpret_cast<float *>(L.rb + R);
    __shared__ int s_flag, s_sid;
    __shared__ double s_err[4];
    unsigned long long steps = 0;
    unsigned int *head = reinterpret_cast<unsigned int *>(a.counters);
    for (;;) {
        if (t == 0) s_sid = (int)atomicAdd(head, 1u);
        __syncthreads();
        const int sid = s_sid;
        if (sid >= a.nsys) break;
        for (int i = t; i < N; i += T) L.y[i] = a.equations[(size_t)sid * N + i];
        const float scale = a.constants[(size_t)sid * a.const_stride];
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
                dense_eval(net, L.yw, scale, L, hs, METHOD == WIND_SIE);
                if (METHOD == WIND_SIE) {
                    dense_lu<PREC>(L.A, NP, L.S);
                    dense_lu_solve(L.A, NP, L.S, L.f, L.x);
                    for (int s = 0; s < ir_sweeps; s++) dense_refine(net, L.yw, scale, L, hs);
                    for (int i = t; i < N; i += T) L.yw[i] = fmaf(hs, L.x[i], L.yw[i]);
                } else {
                    for (int i = t; i < N; i += T) L.yw[i] = fmaf(hs, L.f[i], L.yw[i]);
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

// LDS bytes of ma_dense_integrate for a network.
inline size_t dense_lds_bytes(int N, int R) {
    const int NP = (N + 15) & ~15;
    return sizeof(float) * ((size_t)6 * NP + 5 * (size_t)R + (size_t)NP * (NP + 1));
}

}  // namespace wind
