"""Checks on the oracle that the mathematics fixes (closed forms, invariants,
independent float64 statements), chosen so that a dropped term, a wrong sign
or index, or a transposed operand fails one of them."""
import numpy as np
import pytest

import hhe6_f64 as H
import oracle_lib as O
from wind_amd import networks as W


def test_gje_2x2_closed_form():
    rng = np.random.default_rng(0)
    for _ in range(50):
        a, d = rng.uniform(2.0, 3.0, 2).astype(np.float32)
        b, c = rng.uniform(-1.0, 1.0, 2).astype(np.float32)
        M = np.array([[a, b], [c, d]], np.float32)
        Inv = np.zeros_like(M)
        Mw = M.copy()
        O.lib().oracle_gje_invert(Mw.ctypes.data, Inv.ctypes.data, 2)
        det = float(a) * float(d) - float(b) * float(c)
        want = np.array([[d, -b], [-c, a]], np.float64) / det
        assert np.allclose(Inv, want, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("N", [3, 6, 17, 40])
def test_gje_inverse_property(N):
    rng = np.random.default_rng(N)
    A = (np.eye(N) * N + rng.standard_normal((N, N))).astype(np.float32)
    Inv = np.zeros_like(A)
    Aw = A.copy()
    O.lib().oracle_gje_invert(Aw.ctypes.data, Inv.ctypes.data, N)
    assert np.allclose(A.astype(np.float64) @ Inv, np.eye(N), atol=2e-5)
    assert np.allclose(Inv, np.linalg.inv(A.astype(np.float64)), rtol=1e-4, atol=1e-6)


def test_matvec_transposed_brute_force():
    rng = np.random.default_rng(1)
    N = 7
    M = rng.standard_normal((N, N)).astype(np.float32)
    v = rng.standard_normal(N).astype(np.float32)
    out = v.copy()
    O.lib().oracle_matvec(M.ctypes.data, out.ctypes.data, N)
    # out_r = sum_c M[c*N + r] v_c: the stored matrix is used transposed
    assert np.allclose(out, M.T.astype(np.float64) @ v, rtol=1e-5, atol=1e-5)


def test_take_step_reduces_to_explicit_euler_and_diagonal_form():
    """J = 0 (NR_test with c1 = c2 = 0 and c0 = c3 = 0 gives f = 0, J = 0):
    y unchanged.  Diagonal J (c1 = c2 = 0): SIE gives y_i + h f_i / (1 - h J_ii);
    RK2 (explicit Euler) gives y_i + h f_i."""
    net = O.nr_test()
    y0 = np.array([0.7, -0.3], np.float32)
    c = np.array([-3.0, 0.0, 0.0, -50.0], np.float32)
    h = 0.01
    for method in (O.SIE, O.RK2):
        y = y0.copy()
        O.lib().oracle_take_step(method, net.ref, 0.0, h, 1, y.ctypes.data, c.ctypes.data, 2)
        f = np.array([c[0] * y0[0], c[3] * y0[1]], np.float64)
        if method == O.SIE:
            want = y0 + h * f / (1 - h * np.array([c[0], c[3]], np.float64))
        else:
            want = y0 + h * f
        assert np.allclose(y, want, rtol=1e-6)
    z = np.zeros(4, np.float32)
    y = y0.copy()
    O.lib().oracle_take_step(O.SIE, net.ref, 0.0, h, 2, y.ctypes.data, z.ctypes.data, 2)
    assert np.array_equal(y, y0)


def test_nr_test_against_closed_form():
    """NR_test over t in [0, 1] against y1 = 2e^{-at} - e^{-1000at},
    y2 = -e^{-at} + e^{-1000at}.  The reference controller at tol 1e-3 is
    first-order accurate (about 1.3% at t = 1); the adaptive controller at
    tol 1e-7 is within 1e-3 (float32 time and state)."""
    y0, c, a = W.nr_test_systems(256, seed=9)
    exact = W.nr_test_exact(a, 1.0)
    y = y0.copy()
    _, steps, st = O.gold_integrate(O.SIE, O.nr_test(), 0.0, 1.0, 1000, c, y, 256, 1e-3, 1e-3, const_stride=4)
    assert (st == 0).all()
    rel = np.abs(y.reshape(-1, 2) - exact) / np.abs(exact)
    assert rel.max() < 0.02
    y = y0.copy()
    _, _, st = O.gold_integrate(O.SIE, O.nr_test(), 0.0, 1.0, 1000, c, y, 256, 1e-7, 1e-7, const_stride=4,
                                controller=1, max_attempts=100000)
    assert (st == 0).all()
    assert (np.abs(y.reshape(-1, 2) - exact) / np.abs(exact)).max() < 1e-3


def test_check_error_boundaries():
    """>= at the absolute threshold; the relative test only where |y1|, |y2| > ABS."""
    L = O.lib()
    def f(m, a, b, A, R):
        y1 = np.array([a], np.float32)
        y2 = np.array([b], np.float32)
        return L.oracle_check_error(m, y1.ctypes.data, y2.ctypes.data, 1, A, R)
    assert f(O.SIE, 0.5, 0.5, 1e-3, 1e-3) == 0
    assert f(O.SIE, 0.0, 0.25, 0.25, 10.0) == 1        # |d| == ABS fails (>=)
    assert f(O.SIE, 1e-4, 2e-4, 1e-3, 1e-3) == 0       # relative test skipped below ABS
    assert f(O.SIE, 1.0, 1.01, 1.0, 0.005) == 0        # |y1| = ABS: relative test skipped
    assert f(O.SIE, 1.5, 1.51, 1.0, 0.005) == 1        # relative error 0.0066 >= 0.005
    # SIE divides by y2 (0.1 / 1.1 = 0.091), RK2 by 2 y2 - y1 (0.1 / 1.0)
    assert f(O.SIE, 1.2, 1.1, 1.0, 0.095) == 0 and f(O.RK2, 1.2, 1.1, 1.0, 0.095) == 1


@pytest.mark.parametrize("x", list(np.linspace(-87, 88, 41)) + [-1e-3, 0.0, 1e-7, 0.5, 1.0])
def test_wexpf_against_libm(x):
    v = O.lib().oracle_wexpf(float(x))
    want = np.exp(np.float64(np.float32(x)))
    assert abs(v - want) <= 2.5e-7 * want + 1e-44


@pytest.mark.parametrize("x", [1e-30, 1e-10, 0.3, 0.7071, 1.0, 1.4142, 2.0, 10.0, 123.4, 1e8, 3e38])
def test_wlogf_against_libm(x):
    v = O.lib().oracle_wlogf(float(x))
    want = np.log(np.float64(np.float32(x)))
    assert abs(v - want) <= 3e-7 * max(abs(want), 1.0)


def hhe6_state_samples(n=24, seed=3):
    y0, nH = W.hhe6_systems(n, seed=seed)
    return y0.reshape(-1, 6).astype(np.float64), nH.astype(np.float64)


def test_hhe6_rhs_matches_independent_f64_statement():
    p = W.hhe6_params()
    net = O.hhe6(p)
    Y, nH = hhe6_state_samples()
    for y, n in zip(Y, nH):
        f = np.zeros(6)
        c = np.array([n])
        yc = np.ascontiguousarray(y)
        O.lib().oracle_rhs_f64(net.ref, 0.0, yc.ctypes.data, c.ctypes.data, f.ctypes.data)
        want = H.rhs(0.0, y, n, p)
        assert np.allclose(f, want, rtol=1e-12, atol=1e-300)
        # float version agrees to float rounding
        f32 = np.zeros(6, np.float32)
        y32, c32 = y.astype(np.float32), c.astype(np.float32)
        O.lib().oracle_rhs(net.ref, 0.0, y32.ctypes.data, c32.ctypes.data, f32.ctypes.data)
        scale = np.max(np.abs(want))
        assert np.allclose(f32, want, rtol=2e-4, atol=2e-5 * scale)


def test_hhe6_jacobian_finite_difference_and_conservation():
    p = W.hhe6_params()
    net = O.hhe6(p)
    Y, nH = hhe6_state_samples()
    for y, n in zip(Y, nH):
        J = np.zeros(36)
        c = np.array([n])
        yc = np.ascontiguousarray(y)
        O.lib().oracle_jac_f64(net.ref, yc.ctypes.data, c.ctypes.data, J.ctypes.data)
        Jr = J.reshape(6, 6).T  # stored column-major: J[j*N + i] = df_i/dy_j
        Jfd = H.jac_fd(y, n, p)
        assert np.allclose(Jr, Jfd, rtol=1e-4, atol=1e-6 * np.max(np.abs(Jfd)))
        # sum H and sum He are conserved: the corresponding rows sum to zero
        assert np.allclose(Jr[0] + Jr[1], 0, atol=1e-12 * np.max(np.abs(Jr)))
        assert np.allclose(Jr[2] + Jr[3] + Jr[4], 0, atol=1e-9 * np.max(np.abs(Jr)))


def ma_rhs_numpy(N, rx, k, y, s):
    rx = rx.reshape(-1, 4)
    f = np.zeros(N)
    for (a, b, cc, d), kr in zip(rx, k):
        w = kr * s * y[a] * y[b]
        f[a] -= w
        f[b] -= w
        f[cc] += w
        f[d] += w
    return f


@pytest.mark.parametrize("N", [5, 12])
def test_mass_action_rhs_jac_and_conservation(N):
    rx, k = W.mass_action_network(N, seed=N)
    net = O.mass_action(N, rx, k)
    y0, sc = W.mass_action_systems(6, N, seed=N + 1, scale_range=(0.5, 2.0))
    for y, s in zip(y0.reshape(-1, N).astype(np.float64), sc.astype(np.float64)):
        f = np.zeros(N)
        J = np.zeros(N * N)
        c = np.array([s])
        yy = np.ascontiguousarray(y)
        O.lib().oracle_rhs_f64(net.ref, 0.0, yy.ctypes.data, c.ctypes.data, f.ctypes.data)
        O.lib().oracle_jac_f64(net.ref, yy.ctypes.data, c.ctypes.data, J.ctypes.data)
        assert np.allclose(f, ma_rhs_numpy(N, rx, k, y, s), rtol=1e-12, atol=1e-300)
        Jr = J.reshape(N, N).T
        # independent statement: J = sum_r nu_r (grad w_r)^T, nu_r = -e_a - e_b + e_c + e_d,
        # grad w_r = k_r s (y_b e_a + y_a e_b)
        Jw = np.zeros((N, N))
        for (a, b, cc, d), kr in zip(rx.reshape(-1, 4), k.astype(np.float64)):
            nu = np.zeros(N)
            np.add.at(nu, [a, b, cc, d], [-1.0, -1.0, 1.0, 1.0])
            g = np.zeros(N)
            np.add.at(g, [a, b], [kr * s * y[b], kr * s * y[a]])
            Jw += np.outer(nu, g)
        assert np.allclose(Jr, Jw, rtol=1e-10, atol=1e-12 * np.max(np.abs(Jw)) + 1e-300)
        # a + b -> c + d conserves sum(y): columns of J sum to zero, f sums to zero
        assert abs(f.sum()) <= 1e-12 * np.abs(f).sum() + 1e-300
        assert np.allclose(Jr.sum(axis=0), 0, atol=1e-12 * np.max(np.abs(Jr)))


def test_mass_action_integration_conserves_total():
    N = 10
    rx, k = W.mass_action_network(N, seed=77)
    y0, sc = W.mass_action_systems(16, N, seed=78, scale_range=(0.5, 2.0))
    y = y0.copy()
    _, _, st = O.gold_integrate(O.SIE, O.mass_action(N, rx, k), 0.0, W.MASS_ACTION_TEND, 1, sc, y, 16, 1e-3, 1e-3,
                                const_stride=1, max_attempts=20000)
    tot0 = y0.reshape(-1, N).astype(np.float64).sum(1)
    tot1 = y.reshape(-1, N).astype(np.float64).sum(1)
    ok = st == 0
    assert ok.any()
    assert np.allclose(tot1[ok], tot0[ok], rtol=1e-4)


def test_fp64_oracle_adaptive_close_to_truth():
    """The double-precision oracle with the adaptive controller at tight
    tolerance lands on the FP64 Radau truth of the benchmark sample."""
    import os

    T = np.load(os.path.join(os.path.dirname(__file__), "golden", "truth_hhe6.npz"))
    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(int(T["nsys"]), seed=int(T["seed"]))
    pick = T["idx"][:6]
    y = y0.reshape(-1, 6)[pick].astype(np.float64).reshape(-1).copy()
    c = nH[pick].astype(np.float64)
    O.gold_integrate_f64(O.SIE, O.hhe6(p), 0.0, float(T["tend"]), 1, c, y, len(pick), 1e-7, 1e-7, const_stride=1,
                         controller=1)
    err = np.abs(y.reshape(-1, 6) - T["y"][:6]) / np.maximum(np.abs(T["y"][:6]), 1e-3)
    assert err.max() < 1e-4
