// Device-side building blocks shared by every integrator kernel: math
// policies, the compiled-in ODE networks, and the per-system step-doubling
// control.  Written for gfx950 (CDNA4, wave64); compiled with
// -ffp-contract=off so that every float expression rounds as written.  The
// fast paths request fused multiply-adds explicitly with fmaf().
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wind_types.h"

namespace wind {

// ---------------------------------------------------------------------------
// Math policies.  Exact: portable float exp/log made of plain IEEE float
// operations (Cody-Waite reduction by ln2 with a degree-7 Taylor polynomial;
// log by exponent split and the atanh series), so a CPU build of the same
// algorithm gives the same bits.  Fast: the hardware transcendental unit.
// ---------------------------------------------------------------------------
struct MathExact {
    static constexpr bool kFast = false;
    static constexpr bool kPivot = false;
    __device__ static inline float ex(float x) {
        if (!(x < 88.0f)) x = 88.0f;
        if (x < -87.0f) x = -87.0f;
        float kf = x * 1.44269504f;
        kf = kf + (kf >= 0.0f ? 0.5f : -0.5f);
        int k = (int)kf;
        float fk = (float)k;
        float r = x - fk * 0.693145752f;
        r = r - fk * 1.42860677e-06f;
        float p = 1.98412698e-04f;
        p = p * r + 1.38888889e-03f;
        p = p * r + 8.33333333e-03f;
        p = p * r + 4.16666667e-02f;
        p = p * r + 1.66666667e-01f;
        p = p * r + 0.5f;
        p = p * r + 1.0f;
        p = p * r + 1.0f;
        return p * __uint_as_float((uint32_t)(k + 127) << 23);
    }
    __device__ static inline float lg(float x) {
        uint32_t u = __float_as_uint(x);
        int e = (int)((u >> 23) & 0xffu) - 127;
        float m = __uint_as_float((u & 0x007fffffu) | 0x3f800000u);
        if (m > 1.41421356f) {
            m = m * 0.5f;
            e = e + 1;
        }
        float s = (m - 1.0f) / (m + 1.0f);
        float s2 = s * s;
        float p = 1.11111111e-01f;
        p = p * s2 + 1.42857143e-01f;
        p = p * s2 + 2.0e-01f;
        p = p * s2 + 3.33333333e-01f;
        p = p * s2;
        float t = s + s * p;
        return (float)e * 0.693147181f + (t + t);
    }
    __device__ static inline float mad(float a, float b, float c) { return a * b + c; }
    __device__ static inline float div(float a, float b) { return a / b; }
};

struct MathFast {
    static constexpr bool kFast = true;
    static constexpr bool kPivot = false;
    __device__ static inline float ex(float x) { return __expf(x); }
    // log2 on the transcendental unit times ln 2; x must be a normal float
    // (the only caller takes the log of a temperature of at least 10 K)
    __device__ static inline float lg(float x) { return __builtin_amdgcn_logf(x) * 0.693147181f; }
    __device__ static inline float mad(float a, float b, float c) { return fmaf(a, b, c); }
    // hardware reciprocal (1 ulp) instead of the IEEE division sequence
    __device__ static inline float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
};

// Fast arithmetic with partial pivoting in the LU solves (WIND_FLAG_PIVOT).
struct MathFastPivot : MathFast {
    static constexpr bool kPivot = true;
};

// ---------------------------------------------------------------------------
// Networks.  Each provides N, NCONST, a Params block passed by value in the
// kernel arguments, rhs(p, c, y, f) and jac(p, c, y, J) with J column-major
// (J[j*N + i] = df_i/dy_j) and pre-zeroed by the caller; jac writes the
// non-zeros.  The float expressions are the networks' definitions: in exact
// mode they must round exactly like the CPU definitions.
// ---------------------------------------------------------------------------

// NR_test, eq. 16.6.1 of Numerical Recipes (python/ode_systems/NR_test.py:106-121).
struct NetNRTest {
    static constexpr int N = 2;
    static constexpr int NCONST = 4;
    struct Params {};
    template <class M>
    __device__ static inline void rhs(const Params &, const float *c, const float *y, float *f) {
        f[0] = c[0] * y[0] + c[1] * y[1];
        f[1] = c[2] * y[0] + c[3] * y[1];
    }
    template <class M>
    __device__ static inline void jac(const Params &, const float *c, const float *, float *J) {
        J[0] = c[0];
        J[1] = c[2];
        J[2] = c[1];
        J[3] = c[3];
    }
};

// Katz96 H/He ionisation network (cuda/ode_system/ode_gold.c:3-93), including
// the double-precision `y[4] * 2.0` term and J[13] = J[12] as written there.
struct NetKatz96 {
    static constexpr int N = 5;
    static constexpr int NCONST = 10;
    struct Params {};
    template <class M>
    __device__ static inline void rhs(const Params &, const float *c, const float *y, float *f) {
        float ne = y[1] + y[3] + y[4] * 2.0;
        f[0] = c[2] * ne * y[1] - (c[0] * ne + c[1]) * y[0];
        f[1] = -c[2] * ne * y[1] + (c[0] * ne + c[1]) * y[0];
        f[2] = (c[7] + c[8]) * ne * y[3] - (c[3] * ne + c[4]) * y[2];
        f[3] = c[9] * ne * y[4] + (c[3] * ne + c[4]) * y[2] - (c[7] + c[8]) * ne * y[3] -
               (c[5] * ne + c[6]) * y[3];
        f[4] = (c[5] * ne + c[6]) * y[3] - c[9] * ne * y[4];
    }
    template <class M>
    __device__ static inline void jac(const Params &, const float *c, const float *y, float *J) {
        float ne = y[1] + y[3] + y[4] * 2.0;
        J[0] = -(c[0] * ne + c[1]);
        J[1] = -J[0];
        J[6] = -c[2] * ne;
        J[5] = -J[6];
        J[12] = -(c[3] * ne + c[4]);
        J[13] = J[12];
        J[19] = c[5] * ne + c[6];
        J[17] = (c[7] + c[8]) * ne;
        J[18] = -J[17] - J[19];
        J[24] = -c[9] * ne;
        J[23] = -J[24];
    }
};

// hhe6: H/He ionisation with an internal-energy equation.
// State y = [x_H0, x_H+, x_He0, x_He+, x_He++, e]: abundances per H nucleus
// and the internal energy per H nucleus in units of k_B * 1e4 K; time in
// years; c[0] = n_H [cm^-3].  Two-body rates k_j = A_j T^b_j exp(-C_j/T)
// (0-2 collisional ionisation of H0, He0, He+; 3 H+ recombination; 4, 5
// radiative and dielectronic He+ recombination; 6 He++ recombination) and
// photoionisation rates Gamma_0..2 of H0, He0, He+.  The energy equation is
// photoheating minus collisional-ionisation, recombination and free-free
// cooling.  T_raw = 1e4 e / (1.5 mu); the rates use T = max(T_raw, 10 K) and
// the cooling is scaled by phi = min(1, T_raw / 10 K), so it vanishes as
// e -> 0.  The Jacobian is the exact derivative of this RHS (one-sided at
// the floor).
struct NetHHe6 {
    static constexpr int N = 6;
    static constexpr int NCONST = 1;
    struct Params {
        float lnA[7], b[7], C[7], Gamma[3];
    };
    static constexpr float YR = 3.15576e7f, TUNIT = 1.0e4f, TFLOOR = 10.0f;
    static constexpr float E0 = 4.2f, E1 = 5.8f, E2 = 11.6f;
    static constexpr float CHI0 = 15.78f, CHI1 = 28.55f, CHI2 = 63.13f;
    static constexpr float LFF = 1.0285e-15f, RECC = 0.75f;

    struct State {
        float xe, mu, T, lnT, invT, nHyr, sqT, phi;
        float Rj[7], gj[7], G[3], dT[6], dphi[6];
    };

    template <class M>
    __device__ static inline void state(const Params &p, const float *c, const float *y, State &s,
                                        bool want_derivs) {
        s.xe = (y[1] + y[3]) + (y[4] + y[4]);
        s.mu = (((y[0] + y[1]) + (y[2] + y[3])) + y[4]) + s.xe;
        float dTde = M::div(TUNIT, 1.5f * s.mu);
        float Traw = y[5] * dTde;
        bool clamped = !(Traw > TFLOOR);
        float T = clamped ? TFLOOR : Traw;
        s.phi = clamped ? Traw * 0.1f : 1.0f;
        s.T = T;
        s.invT = M::div(1.0f, T);
        s.nHyr = c[0] * YR;
        if constexpr (M::kFast) {
            // k_j = 2^(log2(A_j) + b_j log2(T) - C_j log2(e) / T) on the
            // transcendental unit (T >= 10 K is a normal float).  Per attempt
            // this stays as close to the float64 value of the attempt as the
            // exact path's float arithmetic does (tested against the oracle's
            // float64 replay of the same attempts).
            constexpr float LOG2E = 1.44269504f;
            const float l2T = __builtin_amdgcn_logf(T);
            s.lnT = l2T * 0.693147181f;
#pragma unroll
            for (int j = 0; j < 7; j++)
                s.Rj[j] = __builtin_amdgcn_exp2f(fmaf(p.b[j], l2T, p.lnA[j] * LOG2E) - (p.C[j] * LOG2E) * s.invT) *
                          s.nHyr;
            s.sqT = __builtin_amdgcn_sqrtf(T);
        } else {
            s.lnT = M::lg(T);
#pragma unroll
            for (int j = 0; j < 7; j++) s.Rj[j] = M::ex((p.lnA[j] + p.b[j] * s.lnT) - p.C[j] * s.invT) * s.nHyr;
            s.sqT = M::ex(0.5f * s.lnT);
        }
        if (want_derivs)
#pragma unroll
            for (int j = 0; j < 7; j++) s.gj[j] = s.Rj[j] * ((p.b[j] + p.C[j] * s.invT) * s.invT);
#pragma unroll
        for (int j = 0; j < 3; j++) s.G[j] = p.Gamma[j] * YR;
        if (want_derivs) {
            float Tmu = M::div(Traw, s.mu);
            const float dmu[5] = {1.0f, 2.0f, 1.0f, 2.0f, 3.0f};
            float dTraw[6];
#pragma unroll
            for (int k = 0; k < 5; k++) dTraw[k] = -(Tmu * dmu[k]);
            dTraw[5] = dTde;
#pragma unroll
            for (int k = 0; k < 6; k++) {
                s.dT[k] = clamped ? 0.0f : dTraw[k];
                s.dphi[k] = clamped ? dTraw[k] * 0.1f : 0.0f;
            }
        }
    }

    template <class M>
    __device__ static inline void rhs(const Params &p, const float *c, const float *y, float *f) {
        State s;
        state<M>(p, c, y, s, false);
        const float xe = s.xe;
        float Rr = s.Rj[4] + s.Rj[5];
        float w0 = (s.Rj[0] * xe + s.G[0]) * y[0];
        float w1 = (s.Rj[1] * xe + s.G[1]) * y[2];
        float w2 = (s.Rj[2] * xe + s.G[2]) * y[3];
        float w3 = (s.Rj[3] * xe) * y[1];
        float w4 = (Rr * xe) * y[3];
        float w5 = (s.Rj[6] * xe) * y[4];
        f[0] = w3 - w0;
        f[1] = w0 - w3;
        f[2] = w4 - w1;
        f[3] = (w1 + w5) - (w2 + w4);
        f[4] = w2 - w5;
        float heat = ((s.G[0] * E0) * y[0] + (s.G[1] * E1) * y[2]) + (s.G[2] * E2) * y[3];
        float cci = (((s.Rj[0] * xe) * y[0]) * CHI0 + ((s.Rj[1] * xe) * y[2]) * CHI1) +
                    ((s.Rj[2] * xe) * y[3]) * CHI2;
        float T4 = s.T * 1.0e-4f;
        float crec = ((w3 + w4) + w5) * (RECC * T4);
        float ion = (y[1] + y[3]) + 4.0f * y[4];
        float cff = ((LFF * s.nHyr) * s.sqT) * (xe * ion);
        f[5] = heat - ((cci + crec) + cff) * s.phi;
    }

    // d/dy_k of w = (R xe + G) y_s.
    __device__ static inline void dw(const State &s, float Rv, float g, float G, int sidx,
                                     const float *y, float *out) {
        const float dxe[6] = {0.0f, 1.0f, 0.0f, 1.0f, 2.0f, 0.0f};
#pragma unroll
        for (int k = 0; k < 6; k++) {
            float t = (g * s.dT[k]) * s.xe + Rv * dxe[k];
            float v = t * y[sidx];
            if (k == sidx) v = v + (Rv * s.xe + G);
            out[k] = v;
        }
    }

    // Fast-mode Jacobian from its structure (other rounding than jac's exact
    // form): every term w_j = (R_j xe + G_j) y_(s_j) has the derivative
    // a_j dT_k + b_j dxe_k + [k == s_j] (R_j xe + G_j) with a_j = g_j xe y_(s_j)
    // and b_j = R_j y_(s_j), so the species rows are U_i dT_k + V_i dxe_k plus
    // a few entries on their own species, and the energy row is
    // heat_k - ((ET dT_k + EX dxe_k + EI dion_k + s_k) phi + cool dphi_k).
    __device__ static inline void jac_fast(const State &s, const float *y, float *J) {
        const float xe = s.xe;
        const float Rr = s.Rj[4] + s.Rj[5], gr = s.gj[4] + s.gj[5];
        // w0..w5 of rhs: species s_j = 0, 2, 3, 1, 3, 4
        const float a0 = (s.gj[0] * xe) * y[0], b0 = s.Rj[0] * y[0], e0 = s.Rj[0] * xe;
        const float a1 = (s.gj[1] * xe) * y[2], b1 = s.Rj[1] * y[2], e1 = s.Rj[1] * xe;
        const float a2 = (s.gj[2] * xe) * y[3], b2 = s.Rj[2] * y[3], e2 = s.Rj[2] * xe;
        const float a3 = (s.gj[3] * xe) * y[1], b3 = s.Rj[3] * y[1], e3 = s.Rj[3] * xe;
        const float a4 = (gr * xe) * y[3], b4 = Rr * y[3], e4 = Rr * xe;
        const float a5 = (s.gj[6] * xe) * y[4], b5 = s.Rj[6] * y[4], e5 = s.Rj[6] * xe;
        const float dxe[6] = {0.0f, 1.0f, 0.0f, 1.0f, 2.0f, 0.0f};
        const float dion[6] = {0.0f, 1.0f, 0.0f, 1.0f, 4.0f, 0.0f};
        // species rows (f0 = w3 - w0, f1 = -f0, f2 = w4 - w1, f3 = w1 + w5 - w2 - w4, f4 = w2 - w5)
        const float U[5] = {a3 - a0, a0 - a3, a4 - a1, (a1 + a5) - (a2 + a4), a2 - a5};
        const float V[5] = {b3 - b0, b0 - b3, b4 - b1, (b1 + b5) - (b2 + b4), b2 - b5};
#pragma unroll
        for (int k = 0; k < 6; k++)
#pragma unroll
            for (int i = 0; i < 5; i++) J[6 * k + i] = fmaf(U[i], s.dT[k], V[i] * dxe[k]);
        const float f0 = e0 + s.G[0], f1 = e1 + s.G[1], f2 = e2 + s.G[2];
        J[0] -= f0;           // d f0 / d y0
        J[1] += f0;           // d f1 / d y0
        J[6] += e3;           // d f0 / d y1
        J[7] -= e3;           // d f1 / d y1
        J[14] -= f1;          // d f2 / d y2
        J[15] += f1;          // d f3 / d y2
        J[20] += e4;          // d f2 / d y3
        J[21] -= f2 + e4;     // d f3 / d y3
        J[22] += f2;          // d f4 / d y3
        J[27] += e5;          // d f3 / d y4
        J[28] -= e5;          // d f4 / d y4
        // energy row
        const float W = ((e3 * y[1]) + (e4 * y[3])) + (e5 * y[4]);
        const float recT = RECC * (s.T * 1.0e-4f);
        const float ion = (y[1] + y[3]) + 4.0f * y[4];
        const float LN = LFF * s.nHyr;
        const float LNs = LN * s.sqT;
        const float xion = xe * ion;
        const float dsqc = (0.5f * s.sqT) * s.invT;
        const float cci = ((e0 * y[0]) * CHI0 + (e1 * y[2]) * CHI1) + (e2 * y[3]) * CHI2;
        const float cool = (cci + W * recT) + LNs * xion;
        const float ET = ((CHI0 * a0 + CHI1 * a1) + CHI2 * a2) + ((a3 + a4) + a5) * recT +
                         (W * (RECC * 1.0e-4f) + (LN * dsqc) * xion);
        const float EX = ((CHI0 * b0 + CHI1 * b1) + CHI2 * b2) + ((b3 + b4) + b5) * recT + LNs * ion;
        const float EI = LNs * xe;
        const float sp[6] = {CHI0 * e0, recT * e3, CHI1 * e1, CHI2 * e2 + recT * e4, recT * e5, 0.0f};
        const float heat[6] = {s.G[0] * E0, 0.0f, s.G[1] * E1, s.G[2] * E2, 0.0f, 0.0f};
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const float sum = fmaf(ET, s.dT[k], EX * dxe[k] + EI * dion[k]) + sp[k];
            J[6 * k + 5] = heat[k] - fmaf(sum, s.phi, cool * s.dphi[k]);
        }
    }

    template <class M>
    __device__ static inline void jac(const Params &p, const float *c, const float *y, float *J) {
        State s;
        state<M>(p, c, y, s, true);
        if constexpr (M::kFast) {
            jac_fast(s, y, J);
            return;
        }
        const float xe = s.xe;
        float Rr = s.Rj[4] + s.Rj[5];
        float gr = s.gj[4] + s.gj[5];
        float dw0[6], dw1[6], dw2[6], dw3[6], dw4[6], dw5[6], dc0[6], dc1[6], dc2[6];
        dw(s, s.Rj[0], s.gj[0], s.G[0], 0, y, dw0);
        dw(s, s.Rj[1], s.gj[1], s.G[1], 2, y, dw1);
        dw(s, s.Rj[2], s.gj[2], s.G[2], 3, y, dw2);
        dw(s, s.Rj[3], s.gj[3], 0.0f, 1, y, dw3);
        dw(s, Rr, gr, 0.0f, 3, y, dw4);
        dw(s, s.Rj[6], s.gj[6], 0.0f, 4, y, dw5);
        dw(s, s.Rj[0], s.gj[0], 0.0f, 0, y, dc0);
        dw(s, s.Rj[1], s.gj[1], 0.0f, 2, y, dc1);
        dw(s, s.Rj[2], s.gj[2], 0.0f, 3, y, dc2);
        float w3 = (s.Rj[3] * xe) * y[1];
        float w4 = (Rr * xe) * y[3];
        float w5 = (s.Rj[6] * xe) * y[4];
        float W = (w3 + w4) + w5;
        float T4 = s.T * 1.0e-4f;
        float recT = RECC * T4;
        float ion = (y[1] + y[3]) + 4.0f * y[4];
        float LN = LFF * s.nHyr;
        float LNs = LN * s.sqT;
        float xion = xe * ion;
        float dsqc = (0.5f * s.sqT) * s.invT;
        float cci = (((s.Rj[0] * xe) * y[0]) * CHI0 + ((s.Rj[1] * xe) * y[2]) * CHI1) +
                    ((s.Rj[2] * xe) * y[3]) * CHI2;
        float cool = (cci + W * recT) + LNs * xion;
        const float dxe[6] = {0.0f, 1.0f, 0.0f, 1.0f, 2.0f, 0.0f};
        const float dion[6] = {0.0f, 1.0f, 0.0f, 1.0f, 4.0f, 0.0f};
#pragma unroll
        for (int k = 0; k < 6; k++) {
            float *col = J + 6 * k;
            col[0] = dw3[k] - dw0[k];
            col[1] = dw0[k] - dw3[k];
            col[2] = dw4[k] - dw1[k];
            col[3] = (dw1[k] + dw5[k]) - (dw2[k] + dw4[k]);
            col[4] = dw2[k] - dw5[k];
            float heat = 0.0f;
            if (k == 0) heat = s.G[0] * E0;
            if (k == 2) heat = s.G[1] * E1;
            if (k == 3) heat = s.G[2] * E2;
            float dcci = (dc0[k] * CHI0 + dc1[k] * CHI1) + dc2[k] * CHI2;
            float dcrec = ((dw3[k] + dw4[k]) + dw5[k]) * recT + W * (RECC * (s.dT[k] * 1.0e-4f));
            float dcff = (LN * (dsqc * s.dT[k])) * xion + LNs * (dxe[k] * ion + xe * dion[k]);
            col[5] = heat - (((dcci + dcrec) + dcff) * s.phi + cool * s.dphi[k]);
        }
    }
};

// ---------------------------------------------------------------------------
// Error test, restating cuda/c_baseline/common/common_gold.c:8-46: the
// absolute test on |y2 - y1| >= ABS and the relative test in double,
// |(y2 - y1) / (y2 + 1e-12)| >= REL (SIE) or / (2 y2 - y1 + 1e-12) (RK2),
// applied only where |y1| and |y2| both exceed ABS.  No early exit changes
// the result, so the loop may stop at the first failing species.
// ---------------------------------------------------------------------------
template <int METHOD>
__device__ inline bool species_fails(float y1, float y2, float ABS, float REL) {
    float d = y2 - y1;
    if (fabsf(d) >= ABS) return true;
    double den;
    if (METHOD == WIND_RK2)
        den = (double)(2.0f * y2 - y1) + 1e-12;
    else
        den = (double)y2 + 1e-12;
    double rel = fabs((double)d / den);
    return rel >= (double)REL && fabsf(y1) > ABS && fabsf(y2) > ABS;
}

// ---------------------------------------------------------------------------
// Timestep control after one step-doubling attempt of size hh.
//
// REFERENCE restates common_gold.c:121-148: on failure halve (at most
// kRefineCap times in a row, then accept anyway); on acceptance advance tnow
// and retry the whole remaining interval (SIE) or double the step (RK2).
// ADAPTIVE accepts iff the double-precision error norm is <= 1 (a NaN norm
// fails) and rescales the step by clamp(0.9/sqrt(err), 0.2, 4).
// Both set status to STALLED where the next attempt could never advance:
// an accepted step that left tnow unchanged with nothing left to grow it
// (reference), or a step that underflowed to zero.
// Returns true when the attempt is accepted.
// ---------------------------------------------------------------------------
constexpr int kRefineCap = 10;  // common_gold.c:121

// FAST (fast-mode kernels): the step factor in float with a hardware
// reciprocal square root instead of the double sqrt and division; the
// accept test is the same.
template <int METHOD, bool FAST = false>
__device__ inline bool control_step(int controller, float tend, float hh, bool ref_flag, double err,
                                    float &tnow, float &h, int &unsolved, int &status) {
    if (controller == WIND_CTRL_REFERENCE) {
        if (ref_flag && unsolved < kRefineCap) {
            h = hh * 0.5f;
            unsolved++;
            return false;
        }
        const float tprev = tnow;
        tnow = tnow + hh;
        h = (METHOD == WIND_RK2) ? hh * 2.0f : tend - tnow;
        unsolved = 0;
        if (tnow == tprev && (METHOD == WIND_SIE || h == 0.0f)) status = WIND_STATUS_STALLED;
        return true;
    }
    const bool ok = err <= 1.0;
    if constexpr (FAST) {
        // err beyond float range gives rsq(inf) = 0 or rsq(0) = inf; the clamp
        // takes both to the ends of [0.2, 4]
        const float r = 0.9f * __builtin_amdgcn_rsqf((float)err);
        float fac;
        if (ok)
            fac = err > 0.0 ? r : 4.0f;
        else
            fac = (err == err) ? r : 0.2f;
        fac = fminf(4.0f, fmaxf(0.2f, fac));
        h = hh * fac;
    } else {
        double fac;
        if (ok)
            fac = err > 0.0 ? 0.9 / sqrt(err) : 4.0;
        else
            fac = (err == err) ? 0.9 / sqrt(err) : 0.2;
        fac = fmin(4.0, fmax(0.2, fac));
        h = (float)((double)hh * fac);
    }
    if (ok) tnow = tnow + hh;
    if (h == 0.0f && tnow < tend) status = WIND_STATUS_STALLED;
    return ok;
}

// One species' term of the adaptive error norm, |y2 - y1| / (ABS + REL |y2|).
__device__ inline double err_term(float y1, float y2, float ABS, float REL) {
    return fabs((double)y2 - (double)y1) / ((double)ABS + (double)REL * fabs((double)y2));
}

// max() that keeps a NaN once one has been seen (a NaN norm must fail).
__device__ inline double err_max(double a, double b) { return (b > a || b != b || a != a) ? (a != a ? a : b) : a; }

// The adaptive norm for fast mode.  With ABS >= 2^-120 every denominator
// ABS + REL |y2| is a normal float whose hardware reciprocal is finite, so
// the terms are taken in float as |y2 - y1| * rcp(den) (a few ulp from the
// double quotient; a term beyond float range is +inf and fails, as its
// double value would).  A smaller ABS (a uniform kernel argument, so the
// branch does not diverge) takes the double terms, the largest selected by
// cross-multiplication and divided once.  Either way any NaN term,
// including 0/0, makes the norm NaN as err_term and err_max do.
template <int N>
__device__ inline double err_norm_fast(const float (&y1)[N], const float (&y2)[N], float ABS, float REL) {
    bool nan = false;
    if (ABS >= 0x1p-120f) {
        float m = 0.0f;
#pragma unroll
        for (int i = 0; i < N; i++) {
            const float r = fabsf(y2[i] - y1[i]) * __builtin_amdgcn_rcpf(fmaf(REL, fabsf(y2[i]), ABS));
            nan = nan || r != r;
            m = fmaxf(m, r);
        }
        return nan ? __builtin_nan("") : (double)m;
    }
    double nb = 0.0, db = 1.0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const double num = fabs((double)y2[i] - (double)y1[i]);
        const double den = (double)ABS + (double)REL * fabs((double)y2[i]);
        nan = nan || num != num || den != den || (num == 0.0 && den == 0.0);
        if (num * db > nb * den) {
            nb = num;
            db = den;
        }
    }
    return nan ? __builtin_nan("") : nb / db;
}

}  // namespace wind
