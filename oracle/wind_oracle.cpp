// TEST INFRASTRUCTURE ONLY - the oracle (see wind_oracle.h for scope, the
// reference files restated, and what pins each function).
//
// Build: g++ -O2 -ffp-contract=off (plain x86-64 SSE2 code, no FMA), so
// every float operation rounds exactly as the reference C does when built
// the same way.  Operation order below follows the reference line by line;
// a comment cites the line each block restates.
#include "wind_oracle.h"

#include <type_traits>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Portable exp / log used by the hhe6 network in float.  Plain IEEE float
// operations only (no FMA, no libm), so the CPU and a GPU evaluate them
// bit-identically.  exp: Cody-Waite reduction by ln2 (hi/lo split) and a
// degree-7 Taylor polynomial on |r| <= ln2/2; log: exponent split with the
// mantissa in [sqrt(1/2), sqrt(2)) and the atanh series
// log(m) = 2 (s + s^3/3 + ... + s^9/9), s = (m-1)/(m+1).
// ---------------------------------------------------------------------------
union FBits { float f; uint32_t u; };

float wexpf(float x) {
    if (!(x < 88.0f)) x = 88.0f;          // also maps NaN to the clamp
    if (x < -87.0f) x = -87.0f;
    float kf = x * 1.44269504f;
    kf = kf + (kf >= 0.0f ? 0.5f : -0.5f);
    int k = (int)kf;
    float fk = (float)k;
    float r = x - fk * 0.693145752f;
    r = r - fk * 1.42860677e-06f;
    float p = 1.98412698e-04f;             // 1/7!
    p = p * r + 1.38888889e-03f;           // 1/6!
    p = p * r + 8.33333333e-03f;           // 1/5!
    p = p * r + 4.16666667e-02f;           // 1/4!
    p = p * r + 1.66666667e-01f;           // 1/3!
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    FBits s;
    s.u = (uint32_t)(k + 127) << 23;
    return p * s.f;
}

float wlogf(float x) {
    FBits b;
    b.f = x;
    int e = (int)((b.u >> 23) & 0xffu) - 127;
    b.u = (b.u & 0x007fffffu) | 0x3f800000u;
    float m = b.f;
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = 1.11111111e-01f;             // 1/9
    p = p * s2 + 1.42857143e-01f;          // 1/7
    p = p * s2 + 2.0e-01f;                 // 1/5
    p = p * s2 + 3.33333333e-01f;          // 1/3
    p = p * s2;
    float t = s + s * p;
    return (float)e * 0.693147181f + (t + t);
}

template <class R> struct Math;
template <> struct Math<float> {
    static float ex(float x) { return wexpf(x); }
    static float lg(float x) { return wlogf(x); }
};
template <> struct Math<double> {
    static double ex(double x) { return std::exp(x); }
    static double lg(double x) { return std::log(x); }
};

// ---------------------------------------------------------------------------
// Networks.
// ---------------------------------------------------------------------------

// Katz96 H/He ionisation network, restating cuda/ode_system/ode_gold.c:3-48
// (calculate_dydt) and :50-93 (calculate_jacobian), including the double
// literal in `equations[4]*2.0` and J[13] = J[12] as written there.
template <class R>
void katz96_rhs(const R *y, const R *c, R *f) {
    R ne = y[1] + y[3] + y[4] * 2.0;
    f[0] = c[2] * ne * y[1] - (c[0] * ne + c[1]) * y[0];
    f[1] = -c[2] * ne * y[1] + (c[0] * ne + c[1]) * y[0];
    f[2] = (c[7] + c[8]) * ne * y[3] - (c[3] * ne + c[4]) * y[2];
    f[3] = c[9] * ne * y[4] + (c[3] * ne + c[4]) * y[2] - (c[7] + c[8]) * ne * y[3] -
           (c[5] * ne + c[6]) * y[3];
    f[4] = (c[5] * ne + c[6]) * y[3] - c[9] * ne * y[4];
}

template <class R>
void katz96_jac(const R *y, const R *c, R *J) {
    R ne = y[1] + y[3] + y[4] * 2.0;
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

// NR_test (Numerical Recipes eq. 16.6.1), restating the text that
// python/ode_systems/NR_test.py:115-121 (RHS) and :106-113 (Jacobian)
// generate for Ntile = 1.
template <class R>
void nr_test_rhs(const R *y, const R *c, R *f) {
    f[0] = c[0] * y[0] + c[1] * y[1];
    f[1] = c[2] * y[0] + c[3] * y[1];
}

template <class R>
void nr_test_jac(const R *, const R *c, R *J) {
    J[0] = c[0];
    J[1] = c[2];
    J[2] = c[1];
    J[3] = c[3];
}

// hhe6: H/He ionisation plus internal energy (authored for this project,
// not part of the reference).  State y = [x_H0, x_H+, x_He0, x_He+, x_He++, e]
// with abundances per H nucleus and e the internal energy per H nucleus in
// units of k_B * 1e4 K; time in years; per-system constant c[0] = n_H [cm^-3].
// Network parameters p: lnA[7] (cm^3 s^-1), b[7], C[7] (K), Gamma[3] (s^-1).
// Two-body rates k_j = A_j T^b_j exp(-C_j / T):
//   0 H0+e->H+, 1 He0+e->He+, 2 He+ +e->He++  (collisional ionisation)
//   3 H+ +e->H0, 4 He+ +e->He0, 5 He+ +e->He0 (dielectronic), 6 He++ +e->He+
// Photoionisation Gamma_0..2 of H0, He0, He+.
// Energy: photoheating minus collisional-ionisation, recombination and
// free-free cooling.  T_raw = 1e4 e / (1.5 mu); rates use T = max(T_raw, 10 K)
// and the cooling is scaled by phi = min(1, T_raw / 10 K), so it vanishes as
// e -> 0.
namespace hhe6 {
constexpr float YR = 3.15576e7f;
constexpr float TUNIT = 1.0e4f;
constexpr float TFLOOR = 10.0f;
constexpr float E0 = 4.2f, E1 = 5.8f, E2 = 11.6f;        // photoheating per ionisation
constexpr float CHI0 = 15.78f, CHI1 = 28.55f, CHI2 = 63.13f;  // ionisation potentials
constexpr float LFF = 1.0285e-15f;                       // free-free, cm^3 s^-1 K^-1/2
constexpr float RECC = 0.75f;                            // recombination cooling per kT
}  // namespace hhe6

template <class R>
struct HHe6State {
    R xe, mu, T, lnT, invT, nHyr, sqT, phi;
    R Rj[7], gj[7], G[3], dT[6], dphi[6];
    bool clamped;
};

template <class R>
void hhe6_state(const R *y, const R *c, const float *p, HHe6State<R> &s) {
    using namespace hhe6;
    s.xe = (y[1] + y[3]) + (y[4] + y[4]);
    s.mu = (((y[0] + y[1]) + (y[2] + y[3])) + y[4]) + s.xe;
    R dTde = (R)TUNIT / ((R)1.5f * s.mu);
    R Traw = y[5] * dTde;
    s.clamped = !(Traw > (R)TFLOOR);
    R T = s.clamped ? (R)TFLOOR : Traw;
    s.phi = s.clamped ? Traw * (R)0.1f : (R)1.0f;
    s.T = T;
    s.lnT = Math<R>::lg(T);
    s.invT = (R)1.0f / T;
    s.nHyr = c[0] * (R)YR;
    for (int j = 0; j < 7; j++) {
        R lnA = (R)p[j], b = (R)p[7 + j], C = (R)p[14 + j];
        s.Rj[j] = Math<R>::ex((lnA + b * s.lnT) - C * s.invT) * s.nHyr;
        s.gj[j] = s.Rj[j] * ((b + C * s.invT) * s.invT);
    }
    for (int j = 0; j < 3; j++) s.G[j] = (R)p[21 + j] * (R)YR;
    s.sqT = Math<R>::ex((R)0.5f * s.lnT);
    // derivatives of T_raw; the rates see them above the floor, phi below it
    R Tmu = Traw / s.mu;
    const R dmu[5] = {1, 2, 1, 2, 3};
    R dTraw[6];
    for (int k = 0; k < 5; k++) dTraw[k] = -(Tmu * dmu[k]);
    dTraw[5] = dTde;
    for (int k = 0; k < 6; k++) {
        s.dT[k] = s.clamped ? (R)0.0f : dTraw[k];
        s.dphi[k] = s.clamped ? dTraw[k] * (R)0.1f : (R)0.0f;
    }
}

template <class R>
void hhe6_rhs(const R *y, const R *c, const float *p, R *f) {
    using namespace hhe6;
    HHe6State<R> s;
    hhe6_state(y, c, p, s);
    const R xe = s.xe;
    R Rr = s.Rj[4] + s.Rj[5];
    R w0 = (s.Rj[0] * xe + s.G[0]) * y[0];
    R w1 = (s.Rj[1] * xe + s.G[1]) * y[2];
    R w2 = (s.Rj[2] * xe + s.G[2]) * y[3];
    R w3 = (s.Rj[3] * xe) * y[1];
    R w4 = (Rr * xe) * y[3];
    R w5 = (s.Rj[6] * xe) * y[4];
    f[0] = w3 - w0;
    f[1] = w0 - w3;
    f[2] = w4 - w1;
    f[3] = (w1 + w5) - (w2 + w4);
    f[4] = w2 - w5;
    R heat = ((s.G[0] * (R)E0) * y[0] + (s.G[1] * (R)E1) * y[2]) + (s.G[2] * (R)E2) * y[3];
    R cci = ((((s.Rj[0] * xe) * y[0]) * (R)CHI0 + ((s.Rj[1] * xe) * y[2]) * (R)CHI1) +
             ((s.Rj[2] * xe) * y[3]) * (R)CHI2);
    R T4 = s.T * (R)1.0e-4f;
    R crec = ((w3 + w4) + w5) * ((R)RECC * T4);
    R ion = (y[1] + y[3]) + (R)4.0f * y[4];
    R cff = (((R)LFF * s.nHyr) * s.sqT) * (xe * ion);
    f[5] = heat - ((cci + crec) + cff) * s.phi;
}

// Partial derivatives of w = (R xe + G) y_s with respect to y_k.
template <class R>
void hhe6_dw(const HHe6State<R> &s, R Rv, R g, R G, int sidx, const R *y, R *dw) {
    const R dxe[6] = {0, 1, 0, 1, 2, 0};
    for (int k = 0; k < 6; k++) {
        R t = (g * s.dT[k]) * s.xe + Rv * dxe[k];
        R v = t * y[sidx];
        if (k == sidx) v = v + (Rv * s.xe + G);
        dw[k] = v;
    }
}

template <class R>
void hhe6_jac(const R *y, const R *c, const float *p, R *J) {
    using namespace hhe6;
    HHe6State<R> s;
    hhe6_state(y, c, p, s);
    const R xe = s.xe;
    const R zero = (R)0.0f;
    R Rr = s.Rj[4] + s.Rj[5];
    R gr = s.gj[4] + s.gj[5];
    R dw0[6], dw1[6], dw2[6], dw3[6], dw4[6], dw5[6];
    R dc0[6], dc1[6], dc2[6];
    hhe6_dw(s, s.Rj[0], s.gj[0], s.G[0], 0, y, dw0);
    hhe6_dw(s, s.Rj[1], s.gj[1], s.G[1], 2, y, dw1);
    hhe6_dw(s, s.Rj[2], s.gj[2], s.G[2], 3, y, dw2);
    hhe6_dw(s, s.Rj[3], s.gj[3], zero, 1, y, dw3);
    hhe6_dw(s, Rr, gr, zero, 3, y, dw4);
    hhe6_dw(s, s.Rj[6], s.gj[6], zero, 4, y, dw5);
    hhe6_dw(s, s.Rj[0], s.gj[0], zero, 0, y, dc0);
    hhe6_dw(s, s.Rj[1], s.gj[1], zero, 2, y, dc1);
    hhe6_dw(s, s.Rj[2], s.gj[2], zero, 3, y, dc2);
    R w3 = (s.Rj[3] * xe) * y[1];
    R w4 = (Rr * xe) * y[3];
    R w5 = (s.Rj[6] * xe) * y[4];
    R W = (w3 + w4) + w5;
    R T4 = s.T * (R)1.0e-4f;
    R recT = (R)RECC * T4;
    R ion = (y[1] + y[3]) + (R)4.0f * y[4];
    R LN = (R)LFF * s.nHyr;
    R LNs = LN * s.sqT;
    R xion = xe * ion;
    R dsqc = ((R)0.5f * s.sqT) * s.invT;
    R cci = ((((s.Rj[0] * xe) * y[0]) * (R)CHI0 + ((s.Rj[1] * xe) * y[2]) * (R)CHI1) +
             ((s.Rj[2] * xe) * y[3]) * (R)CHI2);
    R cool = (cci + W * recT) + LNs * xion;
    const R dxe[6] = {0, 1, 0, 1, 2, 0};
    const R dion[6] = {0, 1, 0, 1, 4, 0};
    for (int k = 0; k < 6; k++) {
        R *col = J + 6 * k;
        col[0] = dw3[k] - dw0[k];
        col[1] = dw0[k] - dw3[k];
        col[2] = dw4[k] - dw1[k];
        col[3] = (dw1[k] + dw5[k]) - (dw2[k] + dw4[k]);
        col[4] = dw2[k] - dw5[k];
        R heat = zero;
        if (k == 0) heat = s.G[0] * (R)E0;
        if (k == 2) heat = s.G[1] * (R)E1;
        if (k == 3) heat = s.G[2] * (R)E2;
        R dcci = (dc0[k] * (R)CHI0 + dc1[k] * (R)CHI1) + dc2[k] * (R)CHI2;
        R dcrec = ((dw3[k] + dw4[k]) + dw5[k]) * recT + W * ((R)RECC * (s.dT[k] * (R)1.0e-4f));
        R dcff = (LN * (dsqc * s.dT[k])) * xion + LNs * (dxe[k] * ion + xe * dion[k]);
        col[5] = heat - (((dcci + dcrec) + dcff) * s.phi + cool * s.dphi[k]);
    }
}

// Mass-action network (authored for this project): R two-body reactions
// a + b -> c + d with rate w = ((k_r * s) * y_a) * y_b, s = c[0] the
// per-system rate scale.  RHS and Jacobian are accumulated reaction by
// reaction in ascending r, updating a, b, c, d in that order; the
// Jacobian updates column a then column b for each of them.
template <class R>
void mass_action_rhs(const oracle_net *net, const R *y, const R *c, R *f) {
    const int N = net->N;
    for (int i = 0; i < N; i++) f[i] = (R)0.0f;
    for (int r = 0; r < net->R; r++) {
        const int32_t *x = net->rx + 4 * r;
        R w = (((R)net->k[r] * c[0]) * y[x[0]]) * y[x[1]];
        f[x[0]] -= w;
        f[x[1]] -= w;
        f[x[2]] += w;
        f[x[3]] += w;
    }
}

template <class R>
void mass_action_jac(const oracle_net *net, const R *y, const R *c, R *J) {
    const int N = net->N;
    for (int r = 0; r < net->R; r++) {
        const int32_t *x = net->rx + 4 * r;
        R ks = (R)net->k[r] * c[0];
        R da = ks * y[x[1]];  // d w / d y_a
        R db = ks * y[x[0]];  // d w / d y_b
        const int a = x[0], b = x[1];
        J[a * N + x[0]] -= da;
        J[b * N + x[0]] -= db;
        J[a * N + x[1]] -= da;
        J[b * N + x[1]] -= db;
        J[a * N + x[2]] += da;
        J[b * N + x[2]] += db;
        J[a * N + x[3]] += da;
        J[b * N + x[3]] += db;
    }
}

template <class R>
void net_rhs(const oracle_net *net, R /*t*/, const R *y, const R *c, R *f) {
    switch (net->id) {
    case ORACLE_NET_KATZ96: katz96_rhs(y, c, f); break;
    case ORACLE_NET_NR_TEST: nr_test_rhs(y, c, f); break;
    case ORACLE_NET_HHE6: hhe6_rhs(y, c, net->params, f); break;
    case ORACLE_NET_MASS_ACTION: mass_action_rhs(net, y, c, f); break;
    case ORACLE_NET_GENERATED:
        if constexpr (std::is_same<R, float>::value)
            net->gen_rhs_f32(y, c, f);
        else
            net->gen_rhs_f64(y, c, f);
        break;
    default: std::abort();
    }
}

template <class R>
void net_jac(const oracle_net *net, const R *y, const R *c, R *J) {
    switch (net->id) {
    case ORACLE_NET_KATZ96: katz96_jac(y, c, J); break;
    case ORACLE_NET_NR_TEST: nr_test_jac(y, c, J); break;
    case ORACLE_NET_HHE6: hhe6_jac(y, c, net->params, J); break;
    case ORACLE_NET_MASS_ACTION: mass_action_jac(net, y, c, J); break;
    case ORACLE_NET_GENERATED:
        if constexpr (std::is_same<R, float>::value)
            net->gen_jac_f32(y, c, J);
        else
            net->gen_jac_f64(y, c, J);
        break;
    default: std::abort();
    }
}

// ---------------------------------------------------------------------------
// Linear algebra: gje_gold.c and matrix_ops.c.
// ---------------------------------------------------------------------------

// gje_gold.c:18-32 scaleRow with read_scale_factor = 1 / 0.
template <class R>
void scale_row(R *row, int n, R scale) {
    for (int t = 0; t < n; t++) row[t] /= scale;
}

// gje_gold.c:34-54 subtractRows.
template <class R>
void subtract_rows(R *target, const R *row, int n, R scale) {
    for (int t = 0; t < n; t++) target[t] -= scale * row[t];
}

// gje_gold.c:136-144 gjeInvertMatrix = setIdentity (:5-16), gjeUFactor
// (:56-100), gjeLFactor (:102-134).  The matrix is destroyed.
template <class R>
void gje_invert(R *M, R *Inv, int n) {
    std::memset(Inv, 0, sizeof(R) * n * n);
    for (int i = 0; i < n; i++) Inv[i * (n + 1)] = (R)1;
    for (int r = 0; r < n; r++) {
        R s = M[r * n + r];
        scale_row(M + r * n, n, s);
        scale_row(Inv + r * n, n, s);
        for (int q = r + 1; q < n; q++) {
            R sq = M[q * n + r];
            subtract_rows(M + q * n, M + r * n, n, sq);
            subtract_rows(Inv + q * n, Inv + r * n, n, sq);
        }
    }
    for (int r = 0; r < n; r++) {
        int br = n - 1 - r;
        for (int q = r + 1; q < n; q++) {
            int bq = n - 1 - q;
            R sq = M[bq * n + br];
            subtract_rows(M + bq * n, M + br * n, n, sq);
            subtract_rows(Inv + bq * n, Inv + br * n, n, sq);
        }
    }
}

// matrix_ops.c:6-27: out[r] = sum_c M[c*n + r] * v[c], c ascending from 0.
template <class R>
void matvec(const R *M, R *v, int n, R *out) {
    for (int i = 0; i < n; i++) out[i] = (R)0;
    for (int c = 0; c < n; c++)
        for (int r = 0; r < n; r++) out[r] += M[c * n + r] * v[c];
    for (int r = 0; r < n; r++) v[r] = out[r];
}

// ---------------------------------------------------------------------------
// Integrator: common_gold.c, sie_gold.c, rk2_gold.c.
// ---------------------------------------------------------------------------

// common_gold.c:8-46.  abs_error and rel_error are doubles; the float
// difference y2 - y1 is rounded to the state type first.  RK2 divides by
// 2*y2 - y1 (evaluated in the state type) + 1e-12.
template <class R>
int check_error(int method, const R *y1, const R *y2, int n, R ABS, R REL) {
    int flag = 0;
    for (int i = 0; i < n; i++) {
        R d = y2[i] - y1[i];
        double abs_error = std::fabs((double)d);
        if (abs_error >= (double)ABS) flag = 1;
        double rel_error;
        if (method == ORACLE_RK2) {
            R den = 2 * y2[i] - y1[i];
            rel_error = std::fabs((double)d / ((double)den + 1e-12));
        } else {
            rel_error = std::fabs((double)d / ((double)y2[i] + 1e-12));
        }
        if (rel_error >= (double)REL && std::fabs((double)y1[i]) > (double)ABS &&
            std::fabs((double)y2[i]) > (double)ABS)
            flag = 1;
    }
    return flag;
}

template <class R>
struct Work {
    std::vector<R> J, Inv, f, out;
    explicit Work(int n) : J(n * n), Inv(n * n), f(n), out(n) {}
};

// sie_gold.c:38-79 (SIE) and rk2_gold.c:18-46 (RK2) take_step.
template <class R>
int take_step(int method, const oracle_net *net, R tnow, R tend, int nsub, R *y, const R *c,
              int n, Work<R> &w) {
    R h = (tend - tnow) / nsub;
    for (int s = 0; s < nsub; s++) {
        net_rhs<R>(net, tnow, y, c, w.f.data());
        if (method == ORACLE_SIE) {
            std::memset(w.J.data(), 0, sizeof(R) * n * n);
            net_jac<R>(net, y, c, w.J.data());
            // scaleAndInvertJacobians, sie_gold.c:17-36
            for (int e = 0; e < n * n; e++) w.J[e] = -h * w.J[e];
            for (int i = 0; i < n; i++) w.J[i * (n + 1)] += 1;
            gje_invert(w.J.data(), w.Inv.data(), n);
            matvec(w.Inv.data(), w.f.data(), n, w.out.data());
        }
        for (int i = 0; i < n; i++) y[i] += h * w.f[i];
        tnow += h;
    }
    return nsub;
}

template <class R>
void record(oracle_trace *tr, int n, R tnow, R h, int flag, int accept, const R *y1, const R *y2) {
    if (!tr) return;
    int a = tr->count++;
    if (a >= tr->capacity) return;
    if (tr->tnow) tr->tnow[a] = (float)tnow;
    if (tr->h) tr->h[a] = (float)h;
    if (tr->flag) tr->flag[a] = flag;
    if (tr->accept) tr->accept[a] = accept;
    for (int i = 0; i < n; i++) {
        if (tr->y1) tr->y1[(size_t)a * n + i] = (float)y1[i];
        if (tr->y2) tr->y2[(size_t)a * n + i] = (float)y2[i];
    }
}

// common_gold.c:48-153 integrateSystem (ADAPTIVE_TIMESTEP defined, as in
// every reference build), with acceptSolution from sie_gold.c:10-15 /
// rk2_gold.c:8-16.
//
// Termination guards (not in the reference, which loops forever in these
// cases): an accepted step that does not advance tnow stops the system with
// ORACLE_STATUS_STALLED when the next trial would repeat it exactly (SIE, or
// RK2 with h == 0); max_attempts > 0 stops it with ORACLE_STATUS_EXHAUSTED.
//
// Controller ORACLE_CTRL_ADAPTIVE (an extension, not in the reference) keeps
// the step-doubling pair and replaces the timestep logic: err = max_i
// |y2 - y1| / (ABS + REL |y2|) in double, accept iff err <= 1 (NaN fails),
// accepted value 2 y2 - y1, next step h * clamp(0.9/sqrt(err), 0.2, 4).
template <class R>
int integrate_system(int method, const oracle_net *net, R tnow, R tend, R h, R *y, const R *c,
                     int n, R ABS, R REL, Work<R> &w, oracle_trace *tr,
                     long long max_attempts = 0, int *status = nullptr, int controller = 0) {
    std::vector<R> y1(y, y + n), y2(y, y + n);
    int nsteps = 0, unsolved = 0;
    long long attempts = 0;
    if (status) *status = ORACLE_STATUS_OK;
    while (tnow < tend) {
        if (max_attempts > 0 && attempts >= max_attempts) {
            if (status) *status = ORACLE_STATUS_EXHAUSTED;
            break;
        }
        attempts++;
        h = (R)std::fmin((double)(tend - tnow), (double)h);
        R tb = tnow + h;
        nsteps += take_step<R>(method, net, tnow, tb, 1, y1.data(), c, n, w);
        nsteps += take_step<R>(method, net, tnow, tb, 2, y2.data(), c, n, w);
        if (controller == ORACLE_CTRL_ADAPTIVE) {
            double err = 0.0;
            for (int i = 0; i < n; i++) {
                double e = std::fabs((double)y2[i] - (double)y1[i]) /
                           ((double)ABS + (double)REL * std::fabs((double)y2[i]));
                if (e > err || e != e || err != err) err = (err != err) ? err : e;
            }
            const bool ok = err <= 1.0;
            record(tr, n, tnow, h, ok ? 0 : 1, ok ? 1 : 0, y1.data(), y2.data());
            double fac;
            if (ok)
                fac = err > 0.0 ? 0.9 / std::sqrt(err) : 4.0;
            else
                fac = (err == err) ? 0.9 / std::sqrt(err) : 0.2;
            fac = std::fmin(4.0, std::fmax(0.2, fac));
            R hh = h;
            h = (R)((double)hh * fac);
            if (ok) {
                for (int i = 0; i < n; i++) y[i] = 2 * y2[i] - y1[i];
                tnow += hh;
            }
            std::memcpy(y1.data(), y, sizeof(R) * n);
            std::memcpy(y2.data(), y, sizeof(R) * n);
            if (h == (R)0 && tnow < tend) {
                if (status) *status = ORACLE_STATUS_STALLED;
                break;
            }
            continue;
        }
        int flag = check_error<R>(method, y1.data(), y2.data(), n, ABS, REL);
        record(tr, n, tnow, h, flag, (flag && unsolved < 10) ? 0 : 1, y1.data(), y2.data());
        if (flag && unsolved < 10) {
            h = (R)((double)h / 2.0);
            unsolved++;
            std::memcpy(y1.data(), y, sizeof(R) * n);
            std::memcpy(y2.data(), y, sizeof(R) * n);
        } else {
            if (method == ORACLE_SIE) {
                std::memcpy(y, y2.data(), sizeof(R) * n);
                std::memcpy(y1.data(), y, sizeof(R) * n);
            } else {
                for (int i = 0; i < n; i++) y[i] = 2 * y2[i] - y1[i];
                std::memcpy(y1.data(), y, sizeof(R) * n);
                std::memcpy(y2.data(), y, sizeof(R) * n);
            }
            R tprev = tnow;
            tnow += h;
            if (method == ORACLE_RK2)
                h *= 2;
            else
                h = (tend - tnow);
            unsolved = 0;
            if (tnow == tprev && (method == ORACLE_SIE || h == (R)0)) {
                if (status) *status = ORACLE_STATUS_STALLED;
                break;
            }
        }
    }
    return nsteps;
}

// sie_gold.c:81-119 / rk2_gold.c:48-79 goldIntegrateSystem.
template <class R>
int gold_integrate(int method, const oracle_net *net, R tnow, R tend, int nfix, const R *C,
                   int cstride, R *Y, int ns, int n, R ABS, R REL, int32_t *steps_out,
                   long long max_attempts = 0, int32_t *status_out = nullptr, int controller = 0) {
    Work<R> w(n);
    int total = 0;
    for (int s = 0; s < ns; s++) {
        int st = 0;
        int k = integrate_system<R>(method, net, tnow, tend, (tend - tnow) / nfix, Y + (size_t)n * s,
                                    C + (size_t)cstride * s, n, ABS, REL, w, nullptr,
                                    max_attempts, &st, controller);
        if (steps_out) steps_out[s] = k;
        if (status_out) status_out[s] = st;
        total += k;
    }
    return total;
}

// Lockstep SIE, restating cuda/SIE_host/harness_kernel.cu:248-425 (errorLoop):
// one timestep shared by all systems; each attempt takes the full step and the
// two half steps (solveSystem with n = 1 and 2, :296-332) for every system;
// the lockstep error test (vector_kernels.cu:45-70: strict '>', relative test
// gated on y1 > ABS and y2 > ABS without absolute values) failing for any
// system halves the step for all; acceptance copies the half-step state and
// doubles the step; the loop gives up after 9 consecutive failures.  The
// sub-step arithmetic is c_baseline's (take_step above).  Returns
// 3 * attempts * Nsystems (:295, :424).
long long lockstep(const oracle_net *net, float tnow, float tend, int nfix, const float *C, int cstride,
                   float *Y, int ns, int n, float ABS, float REL) {
    Work<float> w(n);
    std::vector<float> Y2((size_t)ns * n), y1(n);
    long long nsteps = 0;
    int unsolved = 0;
    float timestep = (tend - tnow) / nfix;
    while (tnow < tend && unsolved < 9) {
        timestep = (float)std::fmin((double)timestep, (double)(tend - tnow));
        nsteps += 3;
        int flag = 0;
        for (int s = 0; s < ns; s++) {
            const float *c = C + (size_t)cstride * s;
            float *y2 = Y2.data() + (size_t)n * s;
            std::memcpy(y1.data(), Y + (size_t)n * s, sizeof(float) * n);
            std::memcpy(y2, Y + (size_t)n * s, sizeof(float) * n);
            take_step<float>(ORACLE_SIE, net, tnow, tnow + timestep, 1, y1.data(), c, n, w);
            take_step<float>(ORACLE_SIE, net, tnow, tnow + timestep, 2, y2, c, n, w);
            for (int i = 0; i < n; i++) {
                float d = y1[i] - y2[i];
                if (std::fabs(d) > ABS) flag = 1;
                if (std::fabs((double)d / ((double)y2[i] + 1e-12)) > (double)REL && y1[i] > ABS && y2[i] > ABS)
                    flag = 1;
            }
        }
        if (flag) {
            unsolved++;
            timestep /= 2;
        } else {
            unsolved = 0;
            std::memcpy(Y, Y2.data(), sizeof(float) * (size_t)ns * n);
            tnow += timestep;
            timestep *= 2;
        }
    }
    return nsteps * ns;
}

}  // namespace

extern "C" {

long long oracle_lockstep(const oracle_net *net, float tnow, float tend, int n_integration_steps,
                          const float *constants, int const_stride, float *equations, int Nsystems,
                          int Neqn, float ABSOLUTE, float RELATIVE) {
    return lockstep(net, tnow, tend, n_integration_steps, constants, const_stride, equations, Nsystems,
                    Neqn, ABSOLUTE, RELATIVE);
}

float oracle_wexpf(float x) { return wexpf(x); }
float oracle_wlogf(float x) { return wlogf(x); }

int oracle_gold_integrate(int method, const oracle_net *net, float tnow, float tend,
                          int n_integration_steps, const float *constants, int const_stride,
                          float *equations, int Nsystems, int Neqn, float ABSOLUTE,
                          float RELATIVE, int32_t *steps_out) {
    return gold_integrate<float>(method, net, tnow, tend, n_integration_steps, constants,
                                 const_stride, equations, Nsystems, Neqn, ABSOLUTE, RELATIVE,
                                 steps_out);
}

int oracle_gold_integrate_ex(int method, const oracle_net *net, float tnow, float tend,
                             int n_integration_steps, const float *constants, int const_stride,
                             float *equations, int Nsystems, int Neqn, float ABSOLUTE,
                             float RELATIVE, int32_t *steps_out, long long max_attempts,
                             int32_t *status_out, int controller) {
    return gold_integrate<float>(method, net, tnow, tend, n_integration_steps, constants,
                                 const_stride, equations, Nsystems, Neqn, ABSOLUTE, RELATIVE,
                                 steps_out, max_attempts, status_out, controller);
}

int oracle_gold_integrate_f64(int method, const oracle_net *net, double tnow, double tend,
                              int n_integration_steps, const double *constants, int const_stride,
                              double *equations, int Nsystems, int Neqn, double ABSOLUTE,
                              double RELATIVE, int32_t *steps_out, int controller) {
    return gold_integrate<double>(method, net, tnow, tend, n_integration_steps, constants,
                                  const_stride, equations, Nsystems, Neqn, ABSOLUTE, RELATIVE,
                                  steps_out, 0, nullptr, controller);
}

int oracle_gold_integrate_mt(int method, const oracle_net *net, float tnow, float tend,
                             int n_integration_steps, const float *constants, int const_stride,
                             float *equations, int Nsystems, int Neqn, float ABSOLUTE,
                             float RELATIVE, int32_t *steps_out, int nthreads, int controller) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > Nsystems) nthreads = Nsystems > 0 ? Nsystems : 1;
    std::vector<int> part(nthreads, 0);
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
        int lo = (int)((long long)Nsystems * t / nthreads);
        int hi = (int)((long long)Nsystems * (t + 1) / nthreads);
        th.emplace_back([=, &part]() {
            part[t] = gold_integrate<float>(method, net, tnow, tend, n_integration_steps,
                                            constants + (size_t)const_stride * lo, const_stride,
                                            equations + (size_t)Neqn * lo, hi - lo, Neqn, ABSOLUTE,
                                            RELATIVE, steps_out ? steps_out + lo : nullptr, 0, nullptr,
                                            controller);
        });
    }
    int total = 0;
    for (int t = 0; t < nthreads; t++) {
        th[t].join();
        total += part[t];
    }
    return total;
}

int oracle_integrate_system_traced(int method, const oracle_net *net, float tnow, float tend,
                                   float timestep, float *y, const float *c, int N,
                                   float ABSOLUTE, float RELATIVE, oracle_trace *trace,
                                   long long max_attempts, int32_t *status, int controller) {
    Work<float> w(N);
    if (trace) trace->count = 0;
    int st = 0;
    int n = integrate_system<float>(method, net, tnow, tend, timestep, y, c, N, ABSOLUTE, RELATIVE,
                                    w, trace, max_attempts, &st, controller);
    if (status) *status = st;
    return n;
}

void oracle_integrate_traced_batch(int method, const oracle_net *net, float tnow, float tend, int nfix,
                                   const float *constants, int const_stride, float *equations, int ns, int n,
                                   float ABSOLUTE, float RELATIVE, int capacity, float *h, int32_t *accept,
                                   int32_t *flag, float *y1, float *y2, int32_t *count, int32_t *steps_out,
                                   int32_t *status_out, long long max_attempts, int controller, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > ns) nthreads = ns > 0 ? ns : 1;
    auto run = [&](int lo, int hi) {
        Work<float> w(n);
        for (int s = lo; s < hi; s++) {
            const size_t base = (size_t)s * capacity;
            oracle_trace tr;
            tr.capacity = capacity;
            tr.count = 0;
            tr.tnow = nullptr;
            tr.h = h ? h + base : nullptr;
            tr.flag = flag ? flag + base : nullptr;
            tr.accept = accept ? accept + base : nullptr;
            tr.y1 = y1 ? y1 + base * n : nullptr;
            tr.y2 = y2 ? y2 + base * n : nullptr;
            int st = 0;
            const int k = integrate_system<float>(method, net, tnow, tend, (tend - tnow) / nfix,
                                                  equations + (size_t)n * s, constants + (size_t)const_stride * s,
                                                  n, ABSOLUTE, RELATIVE, w, &tr, max_attempts, &st, controller);
            if (count) count[s] = tr.count <= capacity ? tr.count : -tr.count;
            if (steps_out) steps_out[s] = k;
            if (status_out) status_out[s] = st;
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
        const int lo = (int)((long long)ns * t / nthreads), hi = (int)((long long)ns * (t + 1) / nthreads);
        th.emplace_back(run, lo, hi);
    }
    for (auto &x : th) x.join();
}

}  // extern "C"

// Attempt replay (an extension for parity checks; no reference counterpart):
// the step sequence of a recorded run of integrate_system, on its float
// time grid, in R arithmetic.
template <class R>
void replay_batch(int method, const oracle_net *net, float tnow, const float *constants, int const_stride,
                  R *equations, int ns, int n, int capacity, const float *h, const int32_t *accept,
                  const int32_t *count, const float *ystart, R *y1o, R *y2o, int controller, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > ns) nthreads = ns > 0 ? ns : 1;
    const bool keep_y2 = controller != ORACLE_CTRL_ADAPTIVE && method == ORACLE_SIE;
    auto run = [&](int lo, int hi) {
        Work<R> w(n);
        std::vector<R> c(const_stride > 0 ? const_stride : 1), y(n), y1(n), y2(n);
        for (int s = lo; s < hi; s++) {
            for (int k = 0; k < const_stride; k++) c[k] = (R)constants[(size_t)const_stride * s + k];
            R *ys = equations ? equations + (size_t)n * s : nullptr;
            if (ys) std::memcpy(y.data(), ys, sizeof(R) * n);
            float t = tnow;
            const int cnt = count[s] < capacity ? count[s] : capacity;
            for (int a = 0; a < cnt; a++) {
                const size_t at = (size_t)s * capacity + a;
                if (ystart)
                    for (int i = 0; i < n; i++) y[i] = (R)ystart[at * n + i];
                const float tb = t + h[at];
                // the float run's full step is fl32(tb - t); in double the
                // interval [t, t + that] reproduces it (and its halves) exactly
                const R te = std::is_same<R, float>::value ? (R)tb : (R)t + (R)(float)(tb - t);
                std::memcpy(y1.data(), y.data(), sizeof(R) * n);
                std::memcpy(y2.data(), y.data(), sizeof(R) * n);
                take_step<R>(method, net, (R)t, te, 1, y1.data(), c.data(), n, w);
                take_step<R>(method, net, (R)t, te, 2, y2.data(), c.data(), n, w);
                if (y1o)
                    for (int i = 0; i < n; i++) y1o[at * n + i] = y1[i];
                if (y2o)
                    for (int i = 0; i < n; i++) y2o[at * n + i] = y2[i];
                if (accept[at]) {
                    for (int i = 0; i < n; i++) y[i] = keep_y2 ? y2[i] : 2 * y2[i] - y1[i];
                    t = tb;
                }
            }
            if (ys && !ystart) std::memcpy(ys, y.data(), sizeof(R) * n);
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
        const int lo = (int)((long long)ns * t / nthreads), hi = (int)((long long)ns * (t + 1) / nthreads);
        th.emplace_back(run, lo, hi);
    }
    for (auto &x : th) x.join();
}

extern "C" {

void oracle_replay_batch(int method, const oracle_net *net, float tnow, const float *constants, int const_stride,
                         float *equations, int ns, int n, int capacity, const float *h, const int32_t *accept,
                         const int32_t *count, const float *ystart, float *y1, float *y2, int controller,
                         int nthreads) {
    replay_batch<float>(method, net, tnow, constants, const_stride, equations, ns, n, capacity, h, accept, count,
                        ystart, y1, y2, controller, nthreads);
}

void oracle_replay_batch_f64(int method, const oracle_net *net, float tnow, const float *constants,
                             int const_stride, double *equations, int ns, int n, int capacity, const float *h,
                             const int32_t *accept, const int32_t *count, const float *ystart, double *y1,
                             double *y2, int controller, int nthreads) {
    replay_batch<double>(method, net, tnow, constants, const_stride, equations, ns, n, capacity, h, accept, count,
                         ystart, y1, y2, controller, nthreads);
}

void oracle_gje_invert(float *matrix, float *inverse, int N) { gje_invert<float>(matrix, inverse, N); }

void oracle_matvec(const float *matrix, float *vector, int N) {
    std::vector<float> out(N);
    matvec<float>(matrix, vector, N, out.data());
}

int oracle_check_error(int method, const float *y1, const float *y2, int N, float ABSOLUTE,
                       float RELATIVE) {
    return check_error<float>(method, y1, y2, N, ABSOLUTE, RELATIVE);
}

int oracle_take_step(int method, const oracle_net *net, float tnow, float tend, int n, float *y,
                     const float *c, int N) {
    Work<float> w(N);
    return take_step<float>(method, net, tnow, tend, n, y, c, N, w);
}

void oracle_rhs(const oracle_net *net, float t, const float *y, const float *c, float *f) {
    net_rhs<float>(net, t, y, c, f);
}

void oracle_jac(const oracle_net *net, const float *y, const float *c, float *J) {
    std::memset(J, 0, sizeof(float) * net->N * net->N);
    net_jac<float>(net, y, c, J);
}

void oracle_rhs_f64(const oracle_net *net, double t, const double *y, const double *c, double *f) {
    net_rhs<double>(net, t, y, c, f);
}

void oracle_jac_f64(const oracle_net *net, const double *y, const double *c, double *J) {
    std::memset(J, 0, sizeof(double) * net->N * net->N);
    net_jac<double>(net, y, c, J);
}

}  // extern "C"
