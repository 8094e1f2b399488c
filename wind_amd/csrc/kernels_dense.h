// Fast-mode integrator for table-driven networks with larger N: one
// workgroup per system, the matrix A = I - hJ dense in LDS (row-major, odd
// row stride to spread LDS banks, padded with identity rows to a multiple of
// 16), factored by a blocked right-looking LU without pivoting (the
// reference's solve is unpivoted too, gje_gold.c:56-134).
//
// Per 16-column panel one wave inverts the 16x16 diagonal block in place
// (G = A11^-1, Gauss-Jordan in registers); the panel blocks W = G A12 and
// the trailing update A22 -= A21 W are 16x16 tile products spread over the
// workgroup's waves, and the next diagonal block is updated and inverted
// while the other waves finish the trailing update.  The tile products run
// on the matrix cores:
//   PREC_VALU      fused multiply-adds on the vector ALU (no MFMA), reference point
//   PREC_MFMA_F32  v_mfma_f32_16x16x4f32, four per tile (K = 16)
//   PREC_MFMA_BF16 v_mfma_f32_16x16x16bf16_1k, one per tile: BF16 inputs, FP32 accumulation
//   PREC_MFMA_BF16X3 three BF16 products (hi*hi + hi*lo + lo*hi) of the split operands
// (the panel products always use FP32: BF16 applies to the trailing update).
// The forward sweep of the solve rides along with the factorisation; the
// backward sweep runs in one wave with the right-hand side in registers, one
// 16-row block at a time.  Optional iterative refinement recomputes the residual
// r = f - (I - hJ)x in double (J x gathered row by row) and solves for a
// correction with the same factors.  Every step is deterministic: a
// system's result does not depend on what else runs on the device.
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

__device__ inline float bcast(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// c += X * Y for 16x16 tiles, X(m, k) and Y(k, n) element functors.  Output
// layout (the MFMA accumulator layout): lane l holds rows 4(l/16) + r,
// r = 0..3, of column l % 16.
template <int PREC, class FX, class FY>
__device__ inline f32x4 tile_mma(FX X, FY Y, f32x4 c, int lane) {
    const int lr = lane & 15, g = lane >> 4;
    if (PREC == PREC_VALU) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            float acc = c[r];
#pragma unroll
            for (int k = 0; k < 16; k++) acc = fmaf(X(4 * g + r, k), Y(k, lr), acc);
            c[r] = acc;
        }
    } else if (PREC == PREC_MFMA_F32) {
        float xa[4], yb[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            xa[q] = X(lr, 4 * q + g);
            yb[q] = Y(4 * q + g, lr);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) c = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[q], yb[q], c, 0, 0, 0);
    } else {
        bf16x4_t ah, bh, al, bl;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float av = X(lr, 4 * g + q);
            const float bv = Y(4 * g + q, lr);
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
    return c;
}

// Row kg (lanes 16 kg .. 16 kg + 15) of v, copied to all four 16-lane rows
// with gfx950's lane swaps: the 16-lane swap exchanges row 1 of its first
// operand with row 0 of its second, and row 3 with row 2; the 32-lane swap
// exchanges the upper half of the first with the lower half of the second.
// With both operands v: p[0] holds rows (0, 0, 2, 2), p[1] rows (1, 1, 3, 3),
// and the 32-lane swap of (a, b, c, d) gives (a, b, a, b) and (c, d, c, d).
// kg is a compile-time value once the callers' loops are unrolled.
__device__ inline float row_to_all(float v, int kg) {
    const unsigned x = __float_as_uint(v);
    const auto p = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    const unsigned y = (kg & 1) ? p[1] : p[0];
    const auto q = __builtin_amdgcn_permlane32_swap(y, y, false, false);
    return __uint_as_float((kg & 2) ? q[1] : q[0]);
}

// Lane n of each 16-lane row, broadcast to the whole row (DPP row_newbcast;
// k is a compile-time value once the callers' loops are unrolled).
__device__ inline float row_bcast(float v, int k) {
    const int x = __float_as_int(v);
    int r = x;
    switch (k) {
#define WIND_ROW_BCAST(n) \
    case n: r = __builtin_amdgcn_update_dpp(0, x, 0x150 + n, 0xF, 0xF, false); break;
        WIND_ROW_BCAST(0) WIND_ROW_BCAST(1) WIND_ROW_BCAST(2) WIND_ROW_BCAST(3)
        WIND_ROW_BCAST(4) WIND_ROW_BCAST(5) WIND_ROW_BCAST(6) WIND_ROW_BCAST(7)
        WIND_ROW_BCAST(8) WIND_ROW_BCAST(9) WIND_ROW_BCAST(10) WIND_ROW_BCAST(11)
        WIND_ROW_BCAST(12) WIND_ROW_BCAST(13) WIND_ROW_BCAST(14) WIND_ROW_BCAST(15)
#undef WIND_ROW_BCAST
    default: break;
    }
    return __int_as_float(r);
}

// Replaces the 16x16 diagonal block at (k0, k0) by its inverse (in-place
// Gauss-Jordan without pivoting).  Executed by one whole wave: lane l holds
// row l % 16, columns 4(l / 16) .. 4(l / 16) + 3.  The pivot row lies in
// each lane's own 16-lane row (DPP row broadcast); the pivot column comes
// from another row of lanes (row_to_all).
__device__ inline void diag_block_invert(float *A, int S, int k0, int lane) {
    const int r = lane & 15, g = lane >> 4;
    float a[4];
#pragma unroll
    for (int c = 0; c < 4; c++) a[c] = A[(k0 + r) * S + k0 + 4 * g + c];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int kg = k >> 2, kc = k & 3;
        float pv[4];
#pragma unroll
        for (int c = 0; c < 4; c++) pv[c] = row_bcast(a[c], k);
        const float f = row_to_all(a[kc], kg);
        const float rp = __builtin_amdgcn_rcpf(bcast(a[kc], k + 16 * kg));
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const float prow = pv[c] * rp;
            const bool pc = 4 * g + c == k;
            if (r == k)
                a[c] = pc ? rp : prow;
            else
                a[c] = pc ? -f * rp : fmaf(-f, prow, a[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; c++) A[(k0 + r) * S + k0 + 4 * g + c] = a[c];
}

// Cost of one diagonal-block inversion in trailing-tile updates (balances
// wave 0's share of the trailing update).
constexpr int kDiagTiles = 2;

// Trailing update C -= A21 W for tiles [lo, hi) (row-major over nt x nt
// tiles of the trailing matrix at (kend, kend)).  With FP32 MFMA the tiles
// are walked row by row: the A21 operand of a tile row is loaded once, and
// four tiles go at a time so that their loads and MFMA chains overlap.
template <int PREC>
__device__ inline void trailing_tiles(float *A, int S, int k0, int kend, int nt, int lo, int hi, int lane) {
    const int lr = lane & 15, g = lane >> 4;
    int tile = lo;
    if (PREC == PREC_MFMA_F32) {
        while (tile < hi) {
            const int rb = tile / nt;
            const int seg_end = min(hi, (rb + 1) * nt);
            const int r0 = kend + (rb << 4);
            float xa[4];
#pragma unroll
            for (int q = 0; q < 4; q++) xa[q] = -A[(r0 + lr) * S + k0 + 4 * q + g];
            for (; tile + 3 < seg_end; tile += 4) {
                const int c0 = kend + ((tile - rb * nt) << 4);
                f32x4 d[4];
                float b[4][4];
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        d[u][q] = A[(r0 + 4 * g + q) * S + c0 + 16 * u + lr];
                        b[u][q] = A[(k0 + 4 * q + g) * S + c0 + 16 * u + lr];
                    }
#pragma unroll
                for (int q = 0; q < 4; q++)
#pragma unroll
                    for (int u = 0; u < 4; u++) d[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[q], b[u][q], d[u], 0, 0, 0);
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int q = 0; q < 4; q++) A[(r0 + 4 * g + q) * S + c0 + 16 * u + lr] = d[u][q];
            }
            for (; tile < seg_end; tile++) {
                const int c0 = kend + ((tile - rb * nt) << 4);
                f32x4 d;
                float b[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    d[q] = A[(r0 + 4 * g + q) * S + c0 + lr];
                    b[q] = A[(k0 + 4 * q + g) * S + c0 + lr];
                }
#pragma unroll
                for (int q = 0; q < 4; q++) d = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[q], b[q], d, 0, 0, 0);
#pragma unroll
                for (int q = 0; q < 4; q++) A[(r0 + 4 * g + q) * S + c0 + lr] = d[q];
            }
        }
        return;
    }
    for (; tile < hi; tile++) {
        const int rb = tile / nt;
        const int r0 = kend + (rb << 4), c0 = kend + ((tile - rb * nt) << 4);
        f32x4 c;
#pragma unroll
        for (int q = 0; q < 4; q++) c[q] = A[(r0 + 4 * g + q) * S + c0 + lr];
        c = tile_mma<PREC>([&](int m, int k) { return -A[(r0 + m) * S + k0 + k]; },
                           [&](int k, int n) { return A[(k0 + k) * S + c0 + n]; }, c, lane);
#pragma unroll
        for (int q = 0; q < 4; q++) A[(r0 + 4 * g + q) * S + c0 + lr] = c[q];
    }
}

// c_k = G_k y_k for the 16-row block at k0 (G_k the inverted diagonal
// block), in place in y; executed by one whole wave (lanes with the same
// row compute the same value).  The wave reads the block before writing it.
__device__ inline void block_rhs_apply_g(const float *A, int S, int k0, float *y, int lane) {
    const int m = lane & 15;
    float v = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; j++) v = fmaf(A[(k0 + m) * S + k0 + j], y[k0 + j], v);
    if (lane < 16) y[k0 + m] = v;
}

// y_i -= A21_ik c_k for the rows i in [lo_row, NP), rows spread over the
// threads [t0, t0 + nt) of the workgroup.
__device__ inline void block_rhs_update(const float *A, int S, int k0, int lo_row, int NP, float *y, int t0,
                                        int nthreads) {
    for (int i = lo_row + (int)threadIdx.x - t0; i < NP; i += nthreads) {
        float acc = y[i];
#pragma unroll
        for (int j = 0; j < 16; j++) acc = fmaf(-A[i * S + k0 + j], y[k0 + j], acc);
        y[i] = acc;
    }
}

// In-place blocked LU without pivoting of the NP x NP matrix at A (row
// stride S), called by the whole workgroup, in the block form
//   A = [I 0; A21 G I] [A11 A12; 0 A22 - A21 G A12],  G = A11^-1.
// On return each 16x16 diagonal block holds its G, the blocks above the
// diagonal hold W = G A12 and the blocks below hold A21 (each as updated
// by the earlier panels).  Per panel: W = G A12 in 16x16 tile products,
// then the trailing update A22 -= A21 W; wave 0 updates the next diagonal
// block first and inverts it while the other waves finish the update.
// With a right-hand side y (LDS, NP values) the forward sweep of the solve
// rides along: per panel c_k = G_k y_k replaces y_k and the rows below are
// updated, y_i -= A21_ik c_k; dense_back_solve then finishes the solve.
template <int PREC>
__device__ inline void dense_lu(float *A, int NP, int S, float *y = nullptr) {
    const int t = threadIdx.x, T = blockDim.x;
    const int wave = t >> 6, lane = t & 63, nw = T >> 6;
    const int lr = lane & 15, g = lane >> 4;
    constexpr int PP = PREC == PREC_VALU ? PREC_VALU : PREC_MFMA_F32;  // W products
    if (wave == 0) diag_block_invert(A, S, 0, lane);
    __syncthreads();
    int k0 = 0;
    for (; k0 + 16 < NP; k0 += 16) {
        const int kend = k0 + 16;
        const int nt = (NP - kend) >> 4;
        for (int u = wave; u < nt; u += nw) {
            const int c0 = kend + (u << 4);
            const f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
            const f32x4 c = tile_mma<PP>([&](int m, int k) { return A[(k0 + m) * S + k0 + k]; },
                                         [&](int k, int n) { return A[(k0 + k) * S + c0 + n]; }, z, lane);
            // the wave reads the whole tile into registers before writing it back
#pragma unroll
            for (int q = 0; q < 4; q++) A[(k0 + 4 * g + q) * S + c0 + lr] = c[q];
        }
        if (y && wave == nt % nw) block_rhs_apply_g(A, S, k0, y, lane);
        __syncthreads();
        // Tiles 1.. are split over the waves; wave 0 first updates tile 0 (the
        // next diagonal block) and inverts it, so it takes fewer of the rest.
        const int ntt = nt * nt, rest = ntt - 1;
        const int w0 = nw == 1 ? rest : max(0, (rest + kDiagTiles) / nw - kDiagTiles);
        const int lo = wave == 0 ? 1 : 1 + w0 + (wave - 1) * (rest - w0) / (nw - 1);
        const int hi = wave == 0 ? 1 + w0 : 1 + w0 + wave * (rest - w0) / (nw - 1);
        if (wave == 0) {
            trailing_tiles<PREC>(A, S, k0, kend, nt, 0, 1, lane);
            diag_block_invert(A, S, kend, lane);
        }
        trailing_tiles<PREC>(A, S, k0, kend, nt, lo, hi, lane);
        if (y) {
            if (nw == 1)
                block_rhs_update(A, S, k0, kend, NP, y, 0, 64);
            else if (wave > 0)
                block_rhs_update(A, S, k0, kend, NP, y, 64, T - 64);
        }
        __syncthreads();
    }
    if (y) {
        if (wave == 0) block_rhs_apply_g(A, S, k0, y, lane);
        __syncthreads();
    }
}

// Backward sweep of the block solve in one wave (lane l keeps rows l,
// l + 64, l + 128 in registers; NP <= 192): x_k = c_k once the blocks after
// k are done, then c_i -= W_ik x_k for the rows above.  c and x may alias.
__device__ inline void dense_back_solve(const float *A, int NP, int S, const float *cv, float *x) {
    const int t = threadIdx.x;
    if (t < 64) {
        const int NB = NP >> 4;
        float xr[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int row = 64 * c + t;
            xr[c] = row < NP ? cv[row] : 0.0f;
        }
#pragma unroll
        for (int kb = 11; kb > 0; kb--) {
            if (kb >= NB) continue;
            const int c = kb >> 2, kq = kb & 3, k0 = kb << 4;
            float xk[16];
#pragma unroll
            for (int j = 0; j < 16; j++) xk[j] = bcast(xr[c], 16 * kq + j);
#pragma unroll
            for (int c2 = 0; c2 < 3; c2++) {
                const int row = 64 * c2 + t;
                if (c2 <= c && row < k0) {
                    float acc = xr[c2];
#pragma unroll
                    for (int j = 0; j < 16; j++) acc = fmaf(-A[row * S + k0 + j], xk[j], acc);
                    xr[c2] = acc;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int row = 64 * c + t;
            if (row < NP) x[row] = xr[c];
        }
    }
    __syncthreads();
}

// Solves A x = b with the block factors left by dense_lu (used by the
// refinement sweeps; the first solve folds its forward sweep into dense_lu).
// Forward sweep in one wave, per 16-row block k: c_k = G_k y_k (block values
// gathered by readlane), written back to their lanes and broadcast, then
// y_i -= A21_ik c_k for the rows below; then dense_back_solve.  b and x may
// alias; both in LDS.
__device__ inline void dense_lu_solve(const float *A, int NP, int S, const float *b, float *x) {
    const int t = threadIdx.x;
    if (t < 64) {
        const int m = t & 15, grp = t >> 4;
        const int NB = NP >> 4;
        float xr[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int row = 64 * c + t;
            xr[c] = row < NP ? b[row] : 0.0f;
        }
#pragma unroll
        for (int kb = 0; kb < 12; kb++) {
            if (kb >= NB) break;
            const int c = kb >> 2, kq = kb & 3, k0 = kb << 4;
            float y = 0.0f;
#pragma unroll
            for (int j = 0; j < 16; j++) y = fmaf(A[(k0 + m) * S + k0 + j], bcast(xr[c], 16 * kq + j), y);
            if (grp == kq) xr[c] = y;
            if (kb + 1 < NB) {
                float yk[16];
#pragma unroll
                for (int j = 0; j < 16; j++) yk[j] = bcast(y, j);
#pragma unroll
                for (int c2 = 0; c2 < 3; c2++) {
                    const int row = 64 * c2 + t;
                    if (c2 >= c && row >= k0 + 16 && row < NP) {
                        float acc = xr[c2];
#pragma unroll
                        for (int j = 0; j < 16; j++) acc = fmaf(-A[row * S + k0 + j], yk[j], acc);
                        xr[c2] = acc;
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int row = 64 * c + t;
            if (row < NP) x[row] = xr[c];
        }
    }
    __syncthreads();
    dense_back_solve(A, NP, S, x, x);
}

struct DenseLayout {
    float *y, *yw, *y1, *f, *x, *r, *A, *w, *hda, *hdb, *k;
    int4 *rx;
    int *fptr, *fent;
    int NP, S;
};

__device__ inline int dense_np(int N) { return (N + 15) & ~15; }
// Float offsets of the LDS layout (int4 tables and the matrix 16-byte aligned).
__host__ __device__ inline int dense_rx_offset(int NP, int R) { return (6 * NP + 4 * R + 3) & ~3; }
__host__ __device__ inline int dense_a_offset(int NP, int R) {
    return (dense_rx_offset(NP, R) + 4 * R + NP + 1 + 4 * R + 3) & ~3;
}

// Carves the LDS layout and copies the network tables into it (called by
// the whole workgroup once; ends with a barrier).
__device__ inline void dense_layout(float *lds, const MassActionDevice &net, DenseLayout &L) {
    const int N = net.N, R = net.R, T = blockDim.x, t = threadIdx.x;
    L.NP = dense_np(N);
    L.S = L.NP + 1;
    const int NP = L.NP;
    L.y = lds;
    L.yw = L.y + NP;
    L.y1 = L.yw + NP;
    L.f = L.y1 + NP;
    L.x = L.f + NP;
    L.r = L.x + NP;
    L.w = L.r + NP;
    L.hda = L.w + R;
    L.hdb = L.hda + R;
    L.k = L.hdb + R;
    L.rx = reinterpret_cast<int4 *>(lds + dense_rx_offset(NP, R));
    L.fptr = reinterpret_cast<int *>(L.rx + R);
    L.fent = L.fptr + NP + 1;
    L.A = lds + dense_a_offset(NP, R);
    for (int r = t; r < R; r += T) {
        L.rx[r] = reinterpret_cast<const int4 *>(net.rx)[r];
        L.k[r] = net.k[r];
    }
    for (int i = t; i <= N; i += T) L.fptr[i] = net.f_ptr[i];
    for (int i = N + 1 + t; i <= NP; i += T) L.fptr[i] = net.f_ptr[N];
    for (int e = t; e < 4 * R; e += T) L.fent[e] = net.f_ent[e];
    __syncthreads();
}

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
        const int4 x = L.rx[r];
        const float ks = L.k[r] * scale;
        const float ya = yv[x.x], yb = yv[x.y];
        L.w[r] = (ks * ya) * yb;
        L.hda[r] = hs * (ks * yb);
        L.hdb[r] = hs * (ks * ya);
    }
    if (want_jac) {
        // 8-byte stores: measured at N = 157 (512 threads), the fill takes
        // 0.84 us with them and 2.7 us with 16-byte stores.
        float2 *A2 = reinterpret_cast<float2 *>(L.A);
        const int n2 = (NP * S + 1) >> 1;
        for (int q = t; q < n2; q += T) A2[q] = make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    // The tables and per-reaction values are read-only here and disjoint from
    // the matrix rows: restrict lets the compiler issue their loads ahead of
    // the row updates.
    const int *__restrict__ fptr = L.fptr;
    const int *__restrict__ fent = L.fent;
    const int4 *__restrict__ rxs = L.rx;
    const float *__restrict__ w = L.w;
    const float *__restrict__ hda = L.hda;
    const float *__restrict__ hdb = L.hdb;
    for (int i = t; i < NP; i += T) {
        float *__restrict__ row = L.A + i * S;
        if (want_jac) row[i] = 1.0f;
        float acc = 0.0f;
        if (i < N) {
            const int e1 = fptr[i + 1];
            for (int e = fptr[i]; e < e1; e++) {
                const int ent = fent[e];
                const int r = ent >> 1;
                const bool neg = ent & 1;
                acc = neg ? acc - w[r] : acc + w[r];
                if (want_jac) {
                    // J[i][a] += sign * da, J[i][b] += sign * db; A = I - hJ
                    const int4 x = rxs[r];
                    const float ha = hda[r], hb = hdb[r];
                    row[x.x] = neg ? row[x.x] + ha : row[x.x] - ha;
                    row[x.y] = neg ? row[x.y] + hb : row[x.y] - hb;
                }
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
            for (int e = L.fptr[i]; e < L.fptr[i + 1]; e++) {
                const int ent = L.fent[e];
                const int r = ent >> 1;
                const int a = L.rx[r].x, b = L.rx[r].y;
                const double ks = (double)(L.k[r] * scale);
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

// TRACE: per-attempt trace and replay through a.trace, as in the lane
// kernels (kernels_small.h, attempt()).
template <int METHOD, int PREC, bool TRACE = false>
__global__ void __launch_bounds__(512) ma_dense_integrate(IntegrateArgs a, MassActionDevice net, int ir_sweeps) {
    extern __shared__ float lds[];
    const int N = net.N, t = threadIdx.x, T = blockDim.x;
    DenseLayout L;
    dense_layout(lds, net, L);
    const int NP = L.NP;
    __shared__ int s_flag, s_sid;
    __shared__ double s_err[8];
    unsigned long long steps = 0;
    unsigned int *head = reinterpret_cast<unsigned int *>(a.counters);
    for (;;) {
        if (t == 0) {
            const unsigned q = atomicAdd(head, 1u);
            s_sid = q < (unsigned)a.nsys ? (a.index ? a.index[q] : (int)q) : a.nsys;
        }
        __syncthreads();
        const int sid = s_sid;
        if (sid >= a.nsys) break;
        for (int i = t; i < N; i += T) L.y[i] = a.equations[(size_t)sid * N + i];
        const float scale = a.constants[(size_t)sid * a.const_stride];
        __syncthreads();
        float tnow = a.tnow, h = a.h0;
        int unsolved = 0, attempts = 0, status = WIND_STATUS_OK;
        const bool replay = TRACE && a.trace.h != nullptr;
        const int rcount = replay ? min(a.trace.count[sid], a.trace.capacity) : 0;
        while (replay ? attempts < rcount : tnow < a.tend) {
            const size_t slot = TRACE ? (size_t)sid * a.trace.capacity + attempts : 0;
            if (replay && a.trace.y_start) {
                for (int i = t; i < N; i += T) L.y[i] = a.trace.y_start[slot * N + i];
                __syncthreads();
            }
            const float hh = replay ? a.trace.h[slot] : fminf(a.tend - tnow, h);
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
                    // x = (I - hJ)^{-1} f: forward sweep inside the factorisation
                    // (f is overwritten by c = L^{-1} f), then the backward sweep.
                    for (int i = t; i < NP; i += T) L.r[i] = L.f[i];
                    __syncthreads();
                    dense_lu<PREC>(L.A, NP, L.S, L.r);
                    dense_back_solve(L.A, NP, L.S, L.r, L.x);
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
            const bool extrap = METHOD == WIND_RK2 || a.controller != WIND_CTRL_REFERENCE;
            if constexpr (TRACE) {
                if (attempts < a.trace.capacity) {
                    for (int i = t; i < N; i += T) {
                        if (a.trace.y1) a.trace.y1[slot * N + i] = L.y1[i];
                        if (a.trace.y2) a.trace.y2[slot * N + i] = L.yw[i];
                    }
                    if (t == 0 && a.trace.flag)
                        a.trace.flag[slot] = a.controller == WIND_CTRL_REFERENCE ? (int)flag : !(err <= 1.0);
                }
                if (replay) {
                    attempts++;
                    if (a.trace.accept[slot]) {
                        tnow = tnow + hh;
                        for (int i = t; i < N; i += T) L.y[i] = extrap ? 2.0f * L.yw[i] - L.y1[i] : L.yw[i];
                    }
                    __syncthreads();
                    continue;
                }
            }
            attempts++;
            const bool acc = control_step<METHOD, true>(a.controller, a.tend, hh, flag, err, tnow, h, unsolved, status);
            if (acc) {
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
    return sizeof(float) * ((size_t)dense_a_offset(NP, R) + (((size_t)NP * (NP + 1) + 3) & ~(size_t)3));
}

}  // namespace wind
