// Integrator kernels for compiled-in networks with small N (N <= 8).
//
// lane_integrate:  one system per lane, the whole state (y, y1, y2, the
//   matrix and its inverse) in VGPRs, fully unrolled for the network's N.
//   Persistent: a fixed grid sized to the device; each wave fetches system
//   indices from a device-wide queue (one atomic per wave per refill, lanes
//   ranked by ballot), so a lane that finishes its system immediately takes
//   the next one.  Every lane runs the same attempt code; per-system control
//   (h, tnow, refinement count) lives in the lane.
//
// block_integrate: one workgroup (one wave) per system, one thread per
//   equation, matrices in LDS - the launch shape of the reference's fused
//   kernel (cuda/common/kernel.cu:140-310, launched one block per system at
//   cuda/common/harness.cu:76-87).  Kept as the straightforward baseline.
//
// Both follow the reference c_baseline's per-system algorithm
// (cuda/c_baseline/common/common_gold.c:48-153 with take_step from
// sie_gold.c:38-79 / rk2_gold.c:18-46).  In exact mode every float
// operation matches the c_baseline operation sequence.
#pragma once
#include "wind_device.h"

namespace wind {

// ---------------------------------------------------------------------------
// Exact-mode solve: the reference's explicit Gauss-Jordan inverse of the
// stored matrix followed by its transposed mat-vec (gje_gold.c:56-144,
// matrix_ops.c:6-27).  Operations whose operands are structurally zero in
// the reference (the cleared part of a pivot row, the zero half of the
// identity during the forward sweep, every matrix update of the backward
// sweep) are skipped: for finite values they leave every non-zero bit
// unchanged, so the result is the reference's.  g = inverse * f.
// ---------------------------------------------------------------------------
template <int N>
__device__ inline void gje_solve_exact(float (&Mx)[N * N], const float (&f)[N], float (&g)[N]) {
    float Inv[N * N];
#pragma unroll
    for (int e = 0; e < N * N; e++) Inv[e] = 0.0f;
#pragma unroll
    for (int i = 0; i < N; i++) Inv[i * (N + 1)] = 1.0f;
#pragma unroll
    for (int r = 0; r < N; r++) {
        const float s = Mx[r * N + r];
#pragma unroll
        for (int c = r + 1; c < N; c++) Mx[r * N + c] = Mx[r * N + c] / s;
#pragma unroll
        for (int c = 0; c <= r; c++) Inv[r * N + c] = Inv[r * N + c] / s;
#pragma unroll
        for (int q = r + 1; q < N; q++) {
            const float sq = Mx[q * N + r];
#pragma unroll
            for (int c = r + 1; c < N; c++) Mx[q * N + c] = Mx[q * N + c] - sq * Mx[r * N + c];
#pragma unroll
            for (int c = 0; c <= r; c++) Inv[q * N + c] = Inv[q * N + c] - sq * Inv[r * N + c];
        }
    }
#pragma unroll
    for (int r = 0; r < N; r++) {
        const int b = N - 1 - r;
#pragma unroll
        for (int q = r + 1; q < N; q++) {
            const int bq = N - 1 - q;
            const float sq = Mx[bq * N + b];
#pragma unroll
            for (int c = 0; c < N; c++) Inv[bq * N + c] = Inv[bq * N + c] - sq * Inv[b * N + c];
        }
    }
#pragma unroll
    for (int r = 0; r < N; r++) {
        float acc = 0.0f;
#pragma unroll
        for (int c = 0; c < N; c++) acc = acc + Inv[c * N + r] * f[c];
        g[r] = acc;
    }
}

// Fast-mode solve of A x = f with A[i][j] = Ms[j*N + i] (the stored matrix
// is column-major): LU elimination, fused multiply-adds and one reciprocal
// per pivot.  About N^3/3 FMAs against the inverse's N^3.  Unpivoted like the
// reference's Gauss-Jordan, or (PIVOT) with partial pivoting: at column k the
// row i >= k with the largest |A(i, k)| (the first of equals) is exchanged
// with row k, by selects, since the registers cannot be indexed at run time.
template <int N, bool PIVOT = false>
__device__ inline void lu_solve_fast(float (&Ms)[N * N], const float (&f)[N], float (&x)[N]) {
    float b[N], ip[N];
#pragma unroll
    for (int i = 0; i < N; i++) b[i] = f[i];
#pragma unroll
    for (int k = 0; k < N; k++) {
        if constexpr (PIVOT) {
            float best = fabsf(Ms[k * N + k]);
            int p = k;
#pragma unroll
            for (int i = k + 1; i < N; i++) {
                const float v = fabsf(Ms[k * N + i]);
                if (v > best) {
                    best = v;
                    p = i;
                }
            }
#pragma unroll
            for (int i = k + 1; i < N; i++) {
                const bool sw = p == i;
#pragma unroll
                for (int j = k; j < N; j++) {
                    const float t = Ms[j * N + k], u = Ms[j * N + i];
                    Ms[j * N + k] = sw ? u : t;
                    Ms[j * N + i] = sw ? t : u;
                }
                const float t = b[k], u = b[i];
                b[k] = sw ? u : t;
                b[i] = sw ? t : u;
            }
        }
        ip[k] = __builtin_amdgcn_rcpf(Ms[k * N + k]);
#pragma unroll
        for (int i = k + 1; i < N; i++) {
            const float l = Ms[k * N + i] * ip[k];
#pragma unroll
            for (int j = k + 1; j < N; j++) Ms[j * N + i] = fmaf(-l, Ms[j * N + k], Ms[j * N + i]);
            b[i] = fmaf(-l, b[k], b[i]);
        }
    }
#pragma unroll
    for (int i = N - 1; i >= 0; i--) {
        float acc = b[i];
#pragma unroll
        for (int j = i + 1; j < N; j++) acc = fmaf(-Ms[j * N + i], x[j], acc);
        x[i] = acc * ip[i];
    }
}

// Fast-mode SIE update y + h x with (I - h J) x = f, solved as
// (J - I/h) z = -f, z = h x: the shift touches only the diagonal (no h J
// products) and the update is y + z.  A step that rounds to zero, or whose
// reciprocal overflows, shifts the diagonal to -inf: the pivot reciprocals
// are -0, every multiplier is +-0 and, for finite J and f, z = +-0 and y
// stays as it is (a non-finite f or J gives NaN, as y + 0 (I - 0 J)^-1 f
// does).  The factorisation overwrites J.
template <int N, bool PIVOT = false>
__device__ inline void sie_update_fast(float (&J)[N * N], const float (&f)[N], float hs, const float (&y)[N],
                                       float (&out)[N]) {
    const float c = __builtin_amdgcn_rcpf(hs);
#pragma unroll
    for (int i = 0; i < N; i++) J[i * (N + 1)] = J[i * (N + 1)] - c;
    float nf[N], z[N];
#pragma unroll
    for (int i = 0; i < N; i++) nf[i] = -f[i];
    lu_solve_fast<N, PIVOT>(J, nf, z);
#pragma unroll
    for (int i = 0; i < N; i++) out[i] = y[i] + z[i];
}

// One sub-step of take_step (sie_gold.c:54-77, rk2_gold.c:27-44) on a
// register-resident state.
template <class Net, int METHOD, class M>
__device__ inline void substep(const typename Net::Params &p, const float *c, float (&y)[Net::N],
                               float hs) {
    constexpr int N = Net::N;
    float f[N];
    Net::template rhs<M>(p, c, y, f);
    if (METHOD == WIND_SIE) {
        float A[N * N];
#pragma unroll
        for (int e = 0; e < N * N; e++) A[e] = 0.0f;
        Net::template jac<M>(p, c, y, A);
        if constexpr (M::kFast) {
            float yn[N];
            sie_update_fast<N, M::kPivot>(A, f, hs, y, yn);
#pragma unroll
            for (int i = 0; i < N; i++) y[i] = yn[i];
        } else {
            const float nh = -hs;
#pragma unroll
            for (int e = 0; e < N * N; e++) A[e] = nh * A[e];
#pragma unroll
            for (int i = 0; i < N; i++) A[i * (N + 1)] = A[i * (N + 1)] + 1.0f;
            float g[N];
            gje_solve_exact<N>(A, f, g);
#pragma unroll
            for (int i = 0; i < N; i++) y[i] = M::mad(hs, g[i], y[i]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) y[i] = M::mad(hs, f[i], y[i]);
    }
}

// The full step and the first half step of an attempt start from the same
// state: f(y) and J(y) are evaluated once and both systems
// (I - h J) x = f, h = hs1 and hs2, are solved from them.  Every value is the
// one two separate sub-steps would compute, so results are unchanged.
template <class Net, int METHOD, class M>
__device__ inline void substep_pair(const typename Net::Params &p, const float *c, const float (&y)[Net::N],
                                    float hs1, float hs2, float (&y1)[Net::N], float (&y2)[Net::N]) {
    constexpr int N = Net::N;
    float f[N];
    Net::template rhs<M>(p, c, y, f);
    if (METHOD == WIND_SIE) {
        float J[N * N];
#pragma unroll
        for (int e = 0; e < N * N; e++) J[e] = 0.0f;
        Net::template jac<M>(p, c, y, J);
        if constexpr (M::kFast) {
            float A[N * N];
#pragma unroll
            for (int e = 0; e < N * N; e++) A[e] = J[e];
            sie_update_fast<N, M::kPivot>(A, f, hs1, y, y1);
            sie_update_fast<N, M::kPivot>(J, f, hs2, y, y2);
            return;
        }
#pragma unroll
        for (int pass = 0; pass < 2; pass++) {
            const float hs = pass == 0 ? hs1 : hs2;
            const float nh = -hs;
            float A[N * N];
#pragma unroll
            for (int e = 0; e < N * N; e++) A[e] = nh * J[e];
#pragma unroll
            for (int i = 0; i < N; i++) A[i * (N + 1)] = A[i * (N + 1)] + 1.0f;
            float g[N];
            gje_solve_exact<N>(A, f, g);
#pragma unroll
            for (int i = 0; i < N; i++) {
                if (pass == 0)
                    y1[i] = M::mad(hs, g[i], y[i]);
                else
                    y2[i] = M::mad(hs, g[i], y[i]);
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) {
            y1[i] = M::mad(hs1, f[i], y[i]);
            y2[i] = M::mad(hs2, f[i], y[i]);
        }
    }
}

// Per-lane system state between attempts.
template <int N>
struct LaneSys {
    float y[N];
    float tnow, h;
    int unsolved;
    int attempts;
    int status;
};

// One step-doubling attempt (common_gold.c:77-148) followed by the timestep
// control of a.controller.  Returns true when the system is finished.
// TRACE: the attempt is recorded in a.trace (y1, y2 and the kernel's own
// error test, as the reference prints them under DEBUGBLOCK,
// cuda/common/kernel.cu:214-249, common_gold.c:102-113) and, with a.trace.h
// set, takes its step and acceptance from the given sequence instead of the
// controller (replay).
template <class Net, int METHOD, class M, bool TRACE = false>
__device__ inline bool attempt(const typename Net::Params &p, const IntegrateArgs &a, const float *c,
                               LaneSys<Net::N> &s, int sid = 0) {
    constexpr int N = Net::N;
    size_t slot = 0;
    bool replay = false;
    if constexpr (TRACE) {
        replay = a.trace.h != nullptr;
        if (replay && s.attempts >= min(a.trace.count[sid], a.trace.capacity)) return true;
        slot = (size_t)sid * a.trace.capacity + s.attempts;
        if (replay && a.trace.y_start)
#pragma unroll
            for (int i = 0; i < N; i++) s.y[i] = a.trace.y_start[slot * N + i];
    }
    const float hh = replay ? a.trace.h[slot] : fminf(a.tend - s.tnow, s.h);
    const float tb = s.tnow + hh;
    const float hs1 = tb - s.tnow;
    const float hs2 = (tb - s.tnow) / 2.0f;
    float y1[N], y2[N];
    substep_pair<Net, METHOD, M>(p, c, s.y, hs1, hs2, y1, y2);
    substep<Net, METHOD, M>(p, c, y2, hs2);
    bool flag = false;
    double err = 0.0;
    if (a.controller == WIND_CTRL_REFERENCE) {
#pragma unroll
        for (int i = 0; i < N; i++) flag = flag || species_fails<METHOD>(y1[i], y2[i], a.ABS, a.REL);
    } else if constexpr (M::kFast) {
        err = err_norm_fast<N>(y1, y2, a.ABS, a.REL);
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) err = err_max(err, err_term(y1[i], y2[i], a.ABS, a.REL));
    }
    const bool extrap = METHOD == WIND_RK2 || a.controller != WIND_CTRL_REFERENCE;
    if constexpr (TRACE) {
        if (s.attempts < a.trace.capacity) {
            if (a.trace.y1)
#pragma unroll
                for (int i = 0; i < N; i++) a.trace.y1[slot * N + i] = y1[i];
            if (a.trace.y2)
#pragma unroll
                for (int i = 0; i < N; i++) a.trace.y2[slot * N + i] = y2[i];
            if (a.trace.flag) a.trace.flag[slot] = a.controller == WIND_CTRL_REFERENCE ? (int)flag : !(err <= 1.0);
        }
        if (replay) {
            s.attempts++;
            if (a.trace.accept[slot]) {
                s.tnow = s.tnow + hh;
#pragma unroll
                for (int i = 0; i < N; i++) s.y[i] = extrap ? 2.0f * y2[i] - y1[i] : y2[i];
            }
            return s.attempts >= min(a.trace.count[sid], a.trace.capacity);
        }
    }
    s.attempts++;
    const float hh_used = hh;
    if (control_step<METHOD, M::kFast>(a.controller, a.tend, hh_used, flag, err, s.tnow, s.h, s.unsolved,
                                       s.status)) {
#pragma unroll
        for (int i = 0; i < N; i++) s.y[i] = extrap ? 2.0f * y2[i] - y1[i] : y2[i];
    }
    if (s.status != WIND_STATUS_OK) return true;
    if (!(s.tnow < a.tend)) return true;
    if (s.attempts >= a.max_attempts) {
        s.status = WIND_STATUS_EXHAUSTED;
        return true;
    }
    return false;
}

// A system's state in and out of HBM.  For even N the system's first float
// is 8-byte aligned (equations[s*N] with s*N even), so 8-byte vector
// accesses halve the memory instructions, and each one covers contiguous
// bytes: fewer partially written lines in the system-major layout.
template <int N>
__device__ inline void load_state(const float *ys, float (&y)[N]) {
    if constexpr (N % 2 == 0) {
#pragma unroll
        for (int i = 0; i < N / 2; i++) {
            const float2 v = reinterpret_cast<const float2 *>(ys)[i];
            y[2 * i] = v.x;
            y[2 * i + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) y[i] = ys[i];
    }
}

template <int N>
__device__ inline void store_state(float *ys, const float (&y)[N]) {
    if constexpr (N % 2 == 0) {
#pragma unroll
        for (int i = 0; i < N / 2; i++) reinterpret_cast<float2 *>(ys)[i] = make_float2(y[2 * i], y[2 * i + 1]);
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) ys[i] = y[i];
    }
}

// Ranks the lanes that set `need` and hands each the next queue index.
__device__ inline int queue_fetch(unsigned int *head, bool need) {
    const unsigned long long mask = __ballot(need);
    if (mask == 0ull) return 0;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    unsigned int base = 0;
    if (lane == leader) base = atomicAdd(head, (unsigned int)__popcll(mask));
    base = __shfl(base, leader);
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    return (int)(base + (unsigned int)__popcll(mask & below));
}

__device__ inline void wave_add_steps(unsigned long long *total, unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (__lane_id() == 0 && v) atomicAdd(total, v);
}

// The body of lane_integrate, shared with kernels that set other launch
// bounds or are compiled with other floating-point contraction.
template <class Net, int METHOD, class M, bool TRACE = false>
__device__ inline void lane_integrate_body(const IntegrateArgs &a, const typename Net::Params &p) {
    constexpr int N = Net::N;
    constexpr int NC = Net::NCONST;
    unsigned int *head = reinterpret_cast<unsigned int *>(a.counters);
    LaneSys<N> s;
    float c[NC];
    int sid = 0;
    bool active = false, need = true;
    unsigned long long steps = 0;
    for (;;) {
        const int got = queue_fetch(head, need);
        if (need) {
            need = false;
            active = got < a.nsys;
            if (active) {
                if (a.records) {
                    // the system's packed record, read in 16-byte pieces
                    constexpr int R = lane_record_words(N, NC);
                    float r[R];
                    const float4 *rp = reinterpret_cast<const float4 *>(a.records + (size_t)got * R);
#pragma unroll
                    for (int v = 0; v < R / 4; v++) {
                        const float4 x = rp[v];
                        r[4 * v] = x.x;
                        r[4 * v + 1] = x.y;
                        r[4 * v + 2] = x.z;
                        r[4 * v + 3] = x.w;
                    }
#pragma unroll
                    for (int i = 0; i < N; i++) s.y[i] = r[i];
#pragma unroll
                    for (int k = 0; k < NC; k++) c[k] = r[N + k];
                    sid = __float_as_int(r[N + NC]);
                } else {
                    sid = a.index ? a.index[got] : got;
                    const float *ys = a.equations + (size_t)sid * N;
                    const float *cs = a.constants + (size_t)sid * a.const_stride;
                    load_state<N>(ys, s.y);
#pragma unroll
                    for (int k = 0; k < NC; k++) c[k] = cs[k];
                }
                s.tnow = a.tnow;
                s.h = a.h0;
                s.unsolved = 0;
                s.attempts = 0;
                s.status = WIND_STATUS_OK;
            }
        }
        if (!__any(active)) break;
        if (active && attempt<Net, METHOD, M, TRACE>(p, a, c, s, sid)) {
            float *ys = a.equations + (size_t)sid * N;
            store_state<N>(ys, s.y);
            if (a.steps_out) a.steps_out[sid] = 3 * s.attempts;
            if (a.status_out) a.status_out[sid] = s.status;
            steps += 3ull * (unsigned long long)s.attempts;
            active = false;
            need = true;
        }
    }
    wave_add_steps(a.counters + 1, steps);
}

template <class Net, int METHOD, class M, bool TRACE = false>
__global__ void __launch_bounds__(256) lane_integrate(IntegrateArgs a, typename Net::Params p) {
    lane_integrate_body<Net, METHOD, M, TRACE>(a, p);
}

// ---------------------------------------------------------------------------
// block_integrate: one wave per system, equation i on lane i, matrices in
// LDS.  Lane c owns column c of the stored matrix and of its inverse, so the
// only cross-lane dependence is the pivot column read in the forward sweep.
// Exact arithmetic only (it is the reference-shaped baseline).
// ---------------------------------------------------------------------------
template <class Net, int METHOD>
__global__ void __launch_bounds__(64) block_integrate(IntegrateArgs a, typename Net::Params p) {
    constexpr int N = Net::N;
    constexpr int NC = Net::NCONST;
    static_assert(N <= 64, "block_integrate keeps one system per wave");
    __shared__ float y[N], yw[N], y1[N], f[N], cc[NC], Mx[N * N], Inv[N * N];
    const int t = threadIdx.x;
    unsigned long long steps = 0;
    for (int sid = blockIdx.x; sid < a.nsys; sid += gridDim.x) {
        if (t < N) y[t] = a.equations[(size_t)sid * N + t];
        if (t < NC) cc[t] = a.constants[(size_t)sid * a.const_stride + t];
        __syncthreads();
        float tnow = a.tnow, h = a.h0;
        int unsolved = 0, attempts = 0, status = WIND_STATUS_OK;
        while (tnow < a.tend) {
            const float hh = fminf(a.tend - tnow, h);
            const float tb = tnow + hh;
            float hs_full = tb - tnow, hs_half = (tb - tnow) / 2.0f;
            for (int pass = 0; pass < 3; pass++) {
                const float hs = pass == 0 ? hs_full : hs_half;
                if (pass != 2) {
                    if (t < N) yw[t] = y[t];
                    __syncthreads();
                }
                // f and J at yw (thread 0 evaluates the network, as the
                // reference's thread 0 writes J: device_dydt.cu:63-114)
                for (int e = t; e < N * N; e += 64) Mx[e] = 0.0f;
                __syncthreads();
                if (t == 0) {
                    Net::template rhs<MathExact>(p, cc, yw, f);
                    if (METHOD == WIND_SIE) Net::template jac<MathExact>(p, cc, yw, Mx);
                }
                __syncthreads();
                if (METHOD == WIND_SIE) {
                    const float nh = -hs;
                    for (int e = t; e < N * N; e += 64) {
                        float v = nh * Mx[e];
                        if (e % (N + 1) == 0) v = v + 1.0f;
                        Mx[e] = v;
                        Inv[e] = (e % (N + 1) == 0) ? 1.0f : 0.0f;
                    }
                    __syncthreads();
                    for (int r = 0; r < N; r++) {
                        const float s = Mx[r * N + r];
                        if (t < N) {
                            if (t > r) Mx[r * N + t] = Mx[r * N + t] / s;
                            else Inv[r * N + t] = Inv[r * N + t] / s;
                            for (int q = r + 1; q < N; q++) {
                                const float sq = Mx[q * N + r];
                                if (t > r) Mx[q * N + t] = Mx[q * N + t] - sq * Mx[r * N + t];
                                else Inv[q * N + t] = Inv[q * N + t] - sq * Inv[r * N + t];
                            }
                        }
                        __syncthreads();
                    }
                    if (t < N) {
                        for (int r = 0; r < N; r++) {
                            const int b = N - 1 - r;
                            for (int q = r + 1; q < N; q++) {
                                const int bq = N - 1 - q;
                                Inv[bq * N + t] = Inv[bq * N + t] - Mx[bq * N + b] * Inv[b * N + t];
                            }
                        }
                        float acc = 0.0f;
                        for (int cidx = 0; cidx < N; cidx++) acc = acc + Inv[cidx * N + t] * f[cidx];
                        yw[t] = yw[t] + hs * acc;
                    }
                } else {
                    if (t < N) yw[t] = yw[t] + hs * f[t];
                }
                __syncthreads();
                if (pass == 0 && t < N) y1[t] = yw[t];
            }
            bool fail = false;
            double err = 0.0;
            if (a.controller == WIND_CTRL_REFERENCE) {
                if (t < N) fail = species_fails<METHOD>(y1[t], yw[t], a.ABS, a.REL);
            } else {
                if (t < N) err = err_term(y1[t], yw[t], a.ABS, a.REL);
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) err = err_max(err, __shfl_xor(err, off));
            }
            const bool flag = __any(fail);
            attempts++;
            const bool acc = control_step<METHOD>(a.controller, a.tend, hh, flag, err, tnow, h, unsolved, status);
            if (acc && t < N) {
                const bool extrap = METHOD == WIND_RK2 || a.controller != WIND_CTRL_REFERENCE;
                y[t] = extrap ? 2.0f * yw[t] - y1[t] : yw[t];
            }
            __syncthreads();
            if (status != WIND_STATUS_OK) break;
            if (attempts >= a.max_attempts && tnow < a.tend) {
                status = WIND_STATUS_EXHAUSTED;
                break;
            }
        }
        __syncthreads();
        if (t < N) a.equations[(size_t)sid * N + t] = y[t];
        if (t == 0) {
            if (a.steps_out) a.steps_out[sid] = 3 * attempts;
            if (a.status_out) a.status_out[sid] = status;
            steps += 3ull * (unsigned long long)attempts;
        }
        __syncthreads();
    }
    if (t == 0 && steps) atomicAdd(a.counters + 1, steps);
}

}  // namespace wind
