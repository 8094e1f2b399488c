// Register-resident dense SIE solver for mass-action networks with
// 16 NT - 15 <= N <= 16 NT (fast mode).  One 256-thread workgroup (four
// waves) owns one system at a time, systems taken from a device queue, and
// the matrix A = I - hJ lives in the waves' registers instead of LDS: wave
// (wi, wj) holds the 16x16 tiles (i, j) with i = wi, j = wj (mod 2), each in
// the MFMA accumulator layout (lane l holds rows 4(l/16) + q, q = 0..3, of
// column l % 16).  Without the matrix in LDS a workgroup needs a third
// (N > 144 and few enough reactions: a quarter) of a CU's LDS and
// registers, so three (four) systems share a CU and each one's serial
// parts (the pivot chain of a diagonal block, the barriers) overlap with
// the others' work.
//
// Per pass (step-doubling sub-step) of an attempt:
// * assembly: the per-reaction rate terms and f as in the LDS solver
//   (kernels_dense.h, dense_eval, same order, so f and every matrix element
//   round identically); the matrix elements with Jacobian terms are summed
//   into a compact LDS buffer (the elements dealt to all 256 threads on the
//   host, each element its terms in sweep order), and the owner lanes of
//   the tiles read them into registers through per-lane offset and mask
//   words (1 on the diagonal, 0 elsewhere where absent);
// * block LU without pivoting, A = [I 0; A21 G I][A11 A12; 0 A22 - A21 G A12],
//   G = A11^-1, per 16-column panel k, with two barriers:
//     B: W_kj = G_k A_kj for the row panel (G_k from LDS, A_kj in the
//        owner's registers; W replaces A_kj and goes to LDS), the column
//        panel A_ik to LDS, and the forward sweep r_i -= A_ik c_k;
//     C: the trailing update A_ij -= A_ik W_kj on the matrix cores
//        (v_mfma_f32_16x16x4f32, operands from LDS, accumulator in place);
//        the owner of the next diagonal tile updates it first, inverts it in
//        registers (Gauss-Jordan on its transpose, which the accumulator
//        layout holds as rows), publishes G_{k+1} and forms c_{k+1} =
//        G_{k+1} r_{k+1};
// * backward sweep x_k = c_k - sum_{j>k} W_kj x_j by columns, one barrier per
//   block, the W tiles still in registers;
// * y += h x.
// Tile matrix-vector products reduce over the 16 lanes of a row with DPP
// shifts in a fixed order; every result is deterministic and independent of
// the grid.  The step control is the LDS solver's.
#pragma once
#include "kernels_dense.h"

namespace wind {

// Workgroups per CU (WGS) the solver is built for: three; four at ten tiles
// per dimension (N > 144), six at four tiles (N <= 64) and eight at two
// (N <= 32) when the network's LDS footprint allows it (dense_reg_wgs_for).
// Built for more than three, the kernel reads the rate constants from
// global memory instead of LDS and the compiler keeps it within 512 / WGS
// VGPRs (at ten tiles spilling some).  Measured: N = 157 four per CU 3%
// faster than three; N = 64 six per CU 6% faster than three (five, the
// register-limited occupancy of the three-per-CU build); N = 32 at eight
// per CU 8% faster than the LDS solver.

// Assembly stream words a thread loads at a time (streams are padded to a
// multiple of it).
constexpr int kDenseRegBatch = 4;

// Row-panel tiles a wave holds at a time in the trailing update: two from
// eight tiles per dimension (N = 157: 2.26 k -> 2.31 k systems/s; N = 64,
// four tiles, 1.5% slower with two), one below.
template <int NT>
constexpr int kDenseRegGW = NT >= 8 ? 2 : 1;

// LDS layout of the register-resident solver (offsets in floats, 16-byte
// aligned): per-system vectors, per-reaction rate terms, reactant indices
// and rate constants, and a work area shared by the compact element buffer
// of the assembly (cap floats) and the factorisation's panel buffers.  The term lists stay in global memory.
struct DenseRegOffsets {
    int y, yw, y1, f, r, w, hda, hdb, k, ab, fptr, work, total;
};

__host__ __device__ inline DenseRegOffsets dense_reg_offsets(int NT, int R, int WGS, int cap) {
    const int NP = 16 * NT;
    auto up = [](int x) { return (x + 3) & ~3; };
    DenseRegOffsets o;
    int p = 0;
    o.y = p;
    p = up(p + NP);
    o.yw = p;
    p = up(p + NP);
    o.y1 = p;
    p = up(p + NP);
    o.f = p;
    p = up(p + NP);
    o.r = p;
    p = up(p + NP);
    o.w = p;
    p = up(p + R);
    o.hda = p;
    p = up(p + R);
    o.hdb = p;
    p = up(p + R);
    o.k = p;
    if (WGS == 3) p = up(p + R);  // rate constants in LDS
    o.ab = p;
    p = up(p + 2 * R);
    o.fptr = p;
    p = up(p + NP + 1);
    o.work = p;
    const int lu = (2 * NT + 1) * 256;
    p = up(p + (cap > lu ? cap : lu));
    o.total = p;
    return o;
}

template <int CTRL>
__device__ inline float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

// Sum over the 16 lanes of each DPP row (lanes 16m .. 16m + 15), as a
// prefix sum: lane 15 of the row ends with the row's total.
__device__ inline float row_sum16(float v) {
    v += dpp_f<0x111>(v);  // row_shr:1
    v += dpp_f<0x112>(v);  // row_shr:2
    v += dpp_f<0x114>(v);  // row_shr:4
    v += dpp_f<0x118>(v);  // row_shr:8
    return v;
}

// Inverts a 16x16 tile held in the accumulator layout (lane (c, g) holds
// T[4g + q][c]) in place.  Read as rows, those registers are T^T in the
// layout diag_block_invert works on (lane (r, g) holds row r, columns
// 4g .. 4g + 3); Gauss-Jordan on T^T gives (T^T)^-1 = (T^-1)^T, which read
// back in the accumulator layout is T^-1.  No pivoting: the leading
// principal minors of T^T are those of T.
__device__ inline void tile_invert_acc(f32x4 &m, int lane) {
    // the lane index made opaque, so that the compiler derives the per-step
    // lane masks here instead of keeping 32 of them live across the kernel
    asm volatile("" : "+v"(lane));
    const int r = lane & 15, g = lane >> 4;
    float a[4] = {m[0], m[1], m[2], m[3]};
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
    m = f32x4{a[0], a[1], a[2], a[3]};
}

// dst[16 i + 4g + q] (q = 0..3, lanes with c = 15 store) = mode: dst - T v,
// or T v; T in the accumulator layout, v[c] the lane's column entry.
template <bool SUBTRACT>
__device__ inline void tile_matvec_store(const f32x4 &m, float v, float *dst, int lane) {
    const int c = lane & 15, g = lane >> 4;
    float p[4];
#pragma unroll
    for (int q = 0; q < 4; q++) p[q] = row_sum16(m[q] * v);
    if (c == 15) {
#pragma unroll
        for (int q = 0; q < 4; q++) dst[4 * g + q] = SUBTRACT ? dst[4 * g + q] - p[q] : p[q];
    }
}

// TRACE: per-attempt trace and replay through a.trace, as in the other
// kernels.
template <int NT, int WGS, bool TRACE = false>
__global__ void __launch_bounds__(256, WGS) ma_dense_reg_integrate(IntegrateArgs a, MassActionDevice net) {
    constexpr int NP = 16 * NT, NL = (NT + 1) / 2;
    extern __shared__ float lds[];
    const int N = net.N, R = net.R, t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wi = wave >> 1, wj = wave & 1;
    const int lr = lane & 15, g = lane >> 4;
    const DenseRegOffsets o = dense_reg_offsets(NT, R, WGS, net.sc_cap);
    float *y = lds + o.y, *yw = lds + o.yw, *y1 = lds + o.y1, *f = lds + o.f, *rv = lds + o.r;
    float *w = lds + o.w, *hda = lds + o.hda, *hdb = lds + o.hdb, *kk = lds + o.k;
    int2 *ab = reinterpret_cast<int2 *>(lds + o.ab);
    int *fptr = reinterpret_cast<int *>(lds + o.fptr);
    const int *__restrict__ fent = net.f_ent;
    const int *__restrict__ scent = net.sc_ent;
    const unsigned *__restrict__ sctile = net.sc_tile;
    float *work = lds + o.work;
    float *Gbuf = work, *Cbuf = work + 256, *Wbuf = work + 256 * (NT + 1);
    // Per-lane bases of this wave's tiles (tile row i = 2 ta + wi, column
    // j = 2 tb + wj): the accesses below add compile-time offsets.
    float *const cA = Cbuf + wi * 256 + lr * 16 + 4 * g;  // A-operand reads of Cbuf tile i: + ta * 512
    float *const cS = Cbuf + wi * 256 + 64 * g + lr;      // stores to Cbuf tile i: + ta * 512 + q * 16
    float *const wB = Wbuf + wj * 256 + lane * 4;          // Wbuf tile j: + tb * 512
    float *const rI = rv + 16 * wi;                        // rows of tile row i: + 32 * ta
    for (int r = t; r < R; r += 256) {
        ab[r] = make_int2(net.rx[4 * r], net.rx[4 * r + 1]);
        if (WGS == 3) kk[r] = net.k[r];
    }
    for (int i = t; i <= N; i += 256) fptr[i] = net.f_ptr[i];
    __shared__ int s_flag, s_sid;
    __shared__ double s_err[4];
    f32x4 M[NL][NL];
    unsigned long long steps = 0;
    unsigned int *head = reinterpret_cast<unsigned int *>(a.counters);
    __syncthreads();

    // f and A = I - hs J at yv into f and the register tiles (ends with a
    // barrier).
    auto assemble = [&](const float *yv, float scale, float hs) __attribute__((always_inline)) {
        constexpr int B = kDenseRegBatch;
        // the tiles' old values are dead (every tile is stored again below,
        // under conditions the compiler cannot prove cover all of them)
#pragma unroll
        for (int ta = 0; ta < NL; ta++)
#pragma unroll
            for (int tb = 0; tb < NL; tb++) M[ta][tb] = __builtin_nondeterministic_value(M[ta][tb]);
        // the first words of the term stream, loaded while the rate terms
        // and f are formed
        int en[B];
        const int *sp = scent + t;
#pragma unroll
        for (int u = 0; u < B; u++) en[u] = sp[u * 256];
        for (int r = t; r < R; r += 256) {
            const int2 x = ab[r];
            const float ks = (WGS == 3 ? kk[r] : net.k[r]) * scale;
            const float ya = yv[x.x], yb = yv[x.y];
            w[r] = (ks * ya) * yb;
            hda[r] = hs * (ks * yb);
            hdb[r] = hs * (ks * ya);
        }
        __syncthreads();
        // f: the row's reaction list in order, its loads issued 8 at a time
        if (t < NP) {
            float acc = 0.0f;
            if (t < N) {
                const int e1 = fptr[t + 1];
                for (int e = fptr[t]; e < e1; e += 8) {
                    int fe[8];
                    float wv[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) fe[u] = e + u < e1 ? fent[e + u] : 0;
#pragma unroll
                    for (int u = 0; u < 8; u++) wv[u] = w[fe[u] >> 1];
#pragma unroll
                    for (int u = 0; u < 8; u++)
                        if (e + u < e1) acc = (fe[u] & 1) ? acc - wv[u] : acc + wv[u];
                }
            }
            f[t] = acc;
        }
        // The matrix's non-zero elements through a compact buffer: each
        // thread walks its term stream (the elements dealt to the 256
        // threads on the host, interleaved so that a batch of words is one
        // coalesced load per word); an element is its terms summed in order
        // in a register (from 1 on the diagonal), stored once at its offset.
        // The owner lane of each tile element then reads it from the buffer
        // through its per-lane offset and mask word (1 on the diagonal, 0
        // elsewhere where the element is absent).
        {
            const int len = net.sc_len;
            float acc = 0.0f;
            bool start = true;
            for (int j = 0;;) {
                float v[B];
#pragma unroll
                for (int u = 0; u < B; u++) v[u] = lds[(unsigned)en[u] >> 16];
#pragma unroll
                for (int u = 0; u < B; u++) {
                    // zero words (padding) load and store nothing that matters
                    const int e = en[u];
                    acc = start ? ((e & 4) ? 1.0f : 0.0f) : acc;
                    acc = (e & 1) ? acc + v[u] : acc - v[u];
                    start = (e & 2) != 0;
                    if (start) work[(e >> 3) & 8191] = acc;
                }
                j += B;
                if (j >= len) break;
#pragma unroll
                for (int u = 0; u < B; u++) en[u] = sp[(j + u) * 256];
            }
        }
        __syncthreads();
        // (an opaque offset per pass, so that the compiler does not keep
        // every tile's words live across passes)
        int tsel = 0;
        asm volatile("" : "+s"(tsel));
        const unsigned *tpass = sctile + tsel;
#pragma unroll
        for (int ta = 0; ta < NL; ta++) {
            const int i = 2 * ta + wi;
            // this lane's word of each tile in the row: offset << 4 | the
            // mask of its q present (a row of tiles at a time, so that few
            // loads are in flight beside the tiles already filled)
            unsigned wrow[NL];
#pragma unroll
            for (int tb = 0; tb < NL; tb++) {
                const int jb = 2 * tb + wj;
                wrow[tb] = (i < NT && jb < NT) ? tpass[(i * NT + jb) * 64 + lane] : 0u;
            }
#pragma unroll
            for (int tb = 0; tb < NL; tb++) {
                const int jb = 2 * tb + wj;
                if (i >= NT || jb >= NT) continue;
                const unsigned wd = wrow[tb];
                unsigned p = wd >> 4;
                f32x4 m;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const bool has = ((wd >> q) & 1u) != 0;
                    const float x = work[p];
                    m[q] = has ? x : (i == jb && 4 * g + q == lr ? 1.0f : 0.0f);
                    p += has ? 1u : 0u;
                }
                M[ta][tb] = m;
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };

    // Diagonal tile (k, k) of its owner: invert, publish G_k, c_k = G_k r_k.
    auto diag_step = [&](f32x4 &m, int k) __attribute__((always_inline)) {
        tile_invert_acc(m, lane);
#pragma unroll
        for (int q = 0; q < 4; q++) Gbuf[(4 * g + q) * 16 + lr] = m[q];
        const float v = rv[16 * k + lr];
        tile_matvec_store<false>(m, v, rv + 16 * k, lane);
    };

    // Block LU without pivoting with the forward sweep on rv (= f on entry),
    // per 16-column panel k (two barriers): B) the row panel W_kj = G_k A_kj
    // (G_k from LDS; W replaces A_kj and goes to LDS), the column panel A_ik
    // (i > k) to LDS and r_i -= A_ik c_k; C) A_ij -= A_ik W_kj on the matrix
    // cores, the owner of the next diagonal tile updating it first, inverting
    // it and forming c_{k+1} = G_{k+1} r_{k+1}.  Leaves c in rv.
    auto factor = [&]() __attribute__((always_inline)) {
        if (t < NP) rv[t] = f[t];
        __syncthreads();
        if (wave == 0) diag_step(M[0][0], 0);
        __syncthreads();
        for (int k = 0; k < NT - 1; k++) {
            if (wi == (k & 1)) {
                const f32x4 gr = *reinterpret_cast<const f32x4 *>(Gbuf + lr * 16 + 4 * g);
#pragma unroll
                for (int ta = 0; ta < NL; ta++) {
                    if (2 * ta + wi != k) continue;
#pragma unroll
                    for (int tb = 0; tb < NL; tb++) {
                        const int j = 2 * tb + wj;
                        if (j <= k || j >= NT) continue;
                        f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                        for (int q = 0; q < 4; q++) z = __builtin_amdgcn_mfma_f32_16x16x4f32(gr[q], M[ta][tb][q], z, 0, 0, 0);
                        M[ta][tb] = z;
                        *reinterpret_cast<f32x4 *>(wB + tb * 512) = z;
                    }
                }
            }
            if (wj == (k & 1)) {
                const float ck = rv[16 * k + lr];
#pragma unroll
                for (int tb = 0; tb < NL; tb++) {
                    if (2 * tb + wj != k) continue;
#pragma unroll
                    for (int ta = 0; ta < NL; ta++) {
                        const int i = 2 * ta + wi;
                        if (i <= k || i >= NT) continue;
                        const f32x4 m = M[ta][tb];
#pragma unroll
                        for (int q = 0; q < 4; q++) cS[ta * 512 + q * 16] = -m[q];
                        tile_matvec_store<true>(m, ck, rI + 32 * ta, lane);
                    }
                }
            }
            __syncthreads();
            const int k1 = k + 1;
            if (wi == (k1 & 1) && wj == (k1 & 1)) {
#pragma unroll
                for (int ta = 0; ta < NL; ta++) {
                    if (2 * ta + wi != k1) continue;
                    const f32x4 xa = *reinterpret_cast<const f32x4 *>(cA + ta * 512);
                    const f32x4 wv = *reinterpret_cast<const f32x4 *>(wB + ta * 512);
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        M[ta][ta] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[q], wv[q], M[ta][ta], 0, 0, 0);
                    diag_step(M[ta][ta], k1);
                }
            }
            // GW row-panel tiles W_kj held at a time, each A_ik read once
            // per group (fewer LDS reads than one W read per tile update)
#pragma unroll
            for (int tb0 = 0; tb0 < NL; tb0 += kDenseRegGW<NT>) {
                constexpr int GW = kDenseRegGW<NT>;
                f32x4 wv[GW];
                bool vv[GW], any = false;
#pragma unroll
                for (int u = 0; u < GW; u++) {
                    const int j = 2 * (tb0 + u) + wj;
                    vv[u] = tb0 + u < NL && j > k && j < NT;
                    any = any || vv[u];
                    wv[u] = vv[u] ? *reinterpret_cast<const f32x4 *>(wB + (tb0 + u) * 512) : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                }
                if (!any) continue;
#pragma unroll
                for (int ta = 0; ta < NL; ta++) {
                    const int i = 2 * ta + wi;
                    if (i <= k || i >= NT) continue;
                    const f32x4 xa = *reinterpret_cast<const f32x4 *>(cA + ta * 512);
#pragma unroll
                    for (int u = 0; u < GW; u++) {
                        const int tb = tb0 + u < NL ? tb0 + u : NL - 1;
                        const int j = 2 * (tb0 + u) + wj;
                        if (!vv[u] || (i == k1 && j == k1)) continue;
#pragma unroll
                        for (int q = 0; q < 4; q++)
                            M[ta][tb] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[q], wv[u][q], M[ta][tb], 0, 0, 0);
                    }
                }
            }
            __syncthreads();
        }
    };

    // Backward sweep by columns: x_k = c_k, then c_i -= W_ik x_k above.
    auto back = [&]() __attribute__((always_inline)) {
        for (int k = NT - 1; k >= 1; k--) {
            if (wj == (k & 1)) {
                const float xk = rv[16 * k + lr];
#pragma unroll
                for (int tb = 0; tb < NL; tb++) {
                    if (2 * tb + wj != k) continue;
#pragma unroll
                    for (int ta = 0; ta < NL; ta++) {
                        const int i = 2 * ta + wi;
                        if (i >= k) continue;
                        tile_matvec_store<true>(M[ta][tb], xk, rI + 32 * ta, lane);
                    }
                }
            }
            __syncthreads();
        }
    };

    for (;;) {
        if (t == 0) {
            const unsigned q = atomicAdd(head, 1u);
            s_sid = q < (unsigned)a.nsys ? (a.index ? a.index[q] : (int)q) : a.nsys;
        }
        __syncthreads();
        const int sid = s_sid;
        if (sid >= a.nsys) break;
        for (int i = t; i < N; i += 256) y[i] = a.equations[(size_t)sid * N + i];
        const float scale = a.constants[(size_t)sid * a.const_stride];
        __syncthreads();
        float tnow = a.tnow, h = a.h0;
        int unsolved = 0, attempts = 0, status = WIND_STATUS_OK;
        const bool replay = TRACE && a.trace.h != nullptr;
        const int rcount = replay ? min(a.trace.count[sid], a.trace.capacity) : 0;
        while (replay ? attempts < rcount : tnow < a.tend) {
            const size_t slot = TRACE ? (size_t)sid * a.trace.capacity + attempts : 0;
            if (replay && a.trace.y_start) {
                for (int i = t; i < N; i += 256) y[i] = a.trace.y_start[slot * N + i];
                __syncthreads();
            }
            const float hh = replay ? a.trace.h[slot] : fminf(a.tend - tnow, h);
            const float tb = tnow + hh;
            const float hs_full = tb - tnow, hs_half = (tb - tnow) / 2.0f;
            for (int pass = 0; pass < 3; pass++) {
                const float hs = pass == 0 ? hs_full : hs_half;
                if (pass != 2) {
                    for (int i = t; i < N; i += 256) yw[i] = y[i];
                    __syncthreads();
                }
                assemble(yw, scale, hs);
                factor();
                back();
                for (int i = t; i < N; i += 256) yw[i] = fmaf(hs, rv[i], yw[i]);
                __syncthreads();
                if (pass == 0)
                    for (int i = t; i < N; i += 256) y1[i] = yw[i];
            }
            if (t == 0) s_flag = 0;
            __syncthreads();
            bool flag = false;
            double err = 0.0;
            if (a.controller == WIND_CTRL_REFERENCE) {
                bool fail = false;
                for (int i = t; i < N; i += 256) fail = fail || species_fails<WIND_SIE>(y1[i], yw[i], a.ABS, a.REL);
                if (fail) s_flag = 1;
            } else {
                for (int i = t; i < N; i += 256) err = err_max(err, err_term(y1[i], yw[i], a.ABS, a.REL));
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) err = err_max(err, __shfl_xor(err, off));
                if (lane == 0) s_err[wave] = err;
            }
            __syncthreads();
            flag = s_flag != 0;
            if (a.controller != WIND_CTRL_REFERENCE) {
                err = s_err[0];
                for (int wv = 1; wv < 4; wv++) err = err_max(err, s_err[wv]);
            }
            const bool extrap = a.controller != WIND_CTRL_REFERENCE;
            if constexpr (TRACE) {
                if (attempts < a.trace.capacity) {
                    for (int i = t; i < N; i += 256) {
                        if (a.trace.y1) a.trace.y1[slot * N + i] = y1[i];
                        if (a.trace.y2) a.trace.y2[slot * N + i] = yw[i];
                    }
                    if (t == 0 && a.trace.flag)
                        a.trace.flag[slot] = a.controller == WIND_CTRL_REFERENCE ? (int)flag : !(err <= 1.0);
                }
                if (replay) {
                    attempts++;
                    if (a.trace.accept[slot]) {
                        tnow = tnow + hh;
                        for (int i = t; i < N; i += 256) y[i] = extrap ? 2.0f * yw[i] - y1[i] : yw[i];
                    }
                    __syncthreads();
                    continue;
                }
            }
            attempts++;
            const bool acc = control_step<WIND_SIE, true>(a.controller, a.tend, hh, flag, err, tnow, h, unsolved, status);
            if (acc)
                for (int i = t; i < N; i += 256) y[i] = extrap ? 2.0f * yw[i] - y1[i] : yw[i];
            __syncthreads();
            if (status != WIND_STATUS_OK) break;
            if (attempts >= a.max_attempts && tnow < a.tend) {
                status = WIND_STATUS_EXHAUSTED;
                break;
            }
        }
        __syncthreads();
        for (int i = t; i < N; i += 256) a.equations[(size_t)sid * N + i] = y[i];
        if (t == 0) {
            if (a.steps_out) a.steps_out[sid] = 3 * attempts;
            if (a.status_out) a.status_out[sid] = status;
            steps += 3ull * (unsigned long long)attempts;
        }
        __syncthreads();
    }
    if (t == 0 && steps) atomicAdd(a.counters + 1, steps);
}

// LDS bytes of ma_dense_reg_integrate (dynamic part) for a network.
inline size_t dense_reg_lds_bytes(int NT, int R, int WGS, int cap) {
    return sizeof(float) * (size_t)dense_reg_offsets(NT, R, WGS, cap).total;
}

// The workgroups per CU to build for: four at ten tiles, six at four and
// eight at two when they fit the CU's LDS (max_lds bytes), else three when
// three fit, else 0 (the solver does not apply).  cap: the network's
// compact element buffer (floats).
inline int dense_reg_wgs_for(int NT, int R, int cap, size_t max_lds) {
    const int want = NT >= 10 ? 4 : NT == 4 ? 6 : NT == 2 ? 8 : 3;
    if (want != 3 && want * (dense_reg_lds_bytes(NT, R, want, cap) + 256) <= max_lds) return want;
    return 3 * (dense_reg_lds_bytes(NT, R, 3, cap) + 256) <= max_lds ? 3 : 0;
}

}  // namespace wind
