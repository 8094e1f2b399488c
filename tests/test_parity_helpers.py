"""The float64 parity helpers: an independent numpy float64
integrator against the oracle's double build, and the Katz96 network's
long-time attractor (the ionisation equilibrium of katz96.py:29-100)
reached by the oracle and, on the GPU, by the library."""
import numpy as np
import pytest

import katz96_eqm as K
import oracle_lib as O
import pysolvers_f64 as P
from wind_amd import api
from wind_amd import networks as W


def _katz96_fns(c):
    """float64 Katz96 RHS / Jacobian (J[i, j] = d f_i / d y_j) from the network's expressions."""
    def rhs(y):
        ne = y[1] + y[3] + 2.0 * y[4]
        return np.array([
            c[2] * ne * y[1] - (c[0] * ne + c[1]) * y[0],
            -c[2] * ne * y[1] + (c[0] * ne + c[1]) * y[0],
            (c[7] + c[8]) * ne * y[3] - (c[3] * ne + c[4]) * y[2],
            (c[3] * ne + c[4]) * y[2] + c[9] * ne * y[4] - (c[5] * ne + c[6]) * y[3] - (c[7] + c[8]) * ne * y[3],
            (c[5] * ne + c[6]) * y[3] - c[9] * ne * y[4]])

    def jac(y):
        eps = 1e-7 * np.maximum(np.abs(y), 1e-12)
        J = np.zeros((5, 5))
        for j in range(5):
            e = np.zeros(5)
            e[j] = eps[j]
            J[:, j] = (rhs(y + e) - rhs(y - e)) / (2 * eps[j])
        return J
    return rhs, jac


def test_katz96_rhs_matches_oracle_f64():
    rng = np.random.default_rng(2)
    net = O.katz96()
    for _ in range(10):
        c = K.constants(10 ** rng.uniform(-2, 2), 10 ** rng.uniform(4, 6)).astype(np.float64)
        y = rng.uniform(0.01, 1.0, 5)
        f = np.zeros(5)
        import ctypes as C
        O.lib().oracle_rhs_f64(net.ref, C.c_double(0.0), y.ctypes.data, c.ctypes.data, f.ctypes.data)
        rhs, _ = _katz96_fns(c)
        assert np.allclose(f, rhs(y), rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("method", [0, 1])
def test_numpy_f64_integrator_matches_oracle_double(method):
    """NR_test (stiff linear, exact Jacobian), adaptive controller at a
    tolerance loose enough that no decision sits near its threshold: the
    two float64 implementations (different solve and operation order)
    agree to 1e-10 with the same sub-step counts."""
    y0, c, a = W.nr_test_systems(4, seed=7)
    net = O.nr_test()
    for s in range(4):
        cs = c[4 * s:4 * s + 4].astype(np.float64)
        A = np.array([[cs[0], cs[1]], [cs[2], cs[3]]])
        yp, nsub = P.integrate_adaptive(method, lambda y: A @ y, lambda y: A, y0[2 * s:2 * s + 2], 0.0, 1.0, 1000,
                                        1e-2 if method == 0 else 1e-1, 1e-2 if method == 0 else 1e-1)
        yo = y0[2 * s:2 * s + 2].astype(np.float64).copy()
        tot, steps = O.gold_integrate_f64(method, net, 0.0, 1.0, 1000, cs.copy(), yo, 1, 1e-2 if method == 0 else 1e-1,
                                          1e-2 if method == 0 else 1e-1, const_stride=4, controller=1)
        assert tot == nsub, (s, tot, nsub)
        assert np.allclose(yo, yp, rtol=1e-10, atol=1e-14)


GRID = [(nH, T) for nH in (1e-2, 1.0, 1e2) for T in (1.5e4, 3e4, 1e5, 1e6)]


def _katz_inputs(y_he=0.25):
    c = np.concatenate([K.constants(nH, T) for nH, T in GRID])
    y0 = np.tile(np.array([1.0, 0.0, y_he / 4, 0.0, 0.0], np.float32), len(GRID))
    eq = np.array([K.equilibrium(nH, T, y_he) for nH, T in GRID])
    return c, y0, eq


def _eq_err(y, eq):
    return np.max(np.abs(y - eq) / np.maximum(np.abs(eq), 1e-6), axis=1)


def test_katz96_oracle_reaches_equilibrium():
    c, y0, eq = _katz_inputs()
    y = y0.copy()
    O.gold_integrate(O.SIE, O.katz96(), 0.0, 1e9, 1, c, y, len(GRID), 1e-7, 1e-7, const_stride=10, controller=1,
                     max_attempts=200000)
    e = _eq_err(y.reshape(-1, 5), eq)
    print("oracle: max rel err vs equilibrium", e.max())
    assert (e <= 1e-3).all()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [api.MODE_EXACT, api.MODE_FAST])
def test_katz96_gpu_reaches_equilibrium(mode):
    c, y0, eq = _katz_inputs()
    y = y0.copy()
    st = np.zeros(len(GRID), np.int32)
    api.integrate(api.Network.builtin(api.NET_KATZ96), 0.0, 1e9, 1, c, y, len(GRID), 1e-7, 1e-7, mode=mode,
                  controller=api.CTRL_ADAPTIVE, const_stride=10, status_out=st, max_attempts=200000)
    e = _eq_err(y.reshape(-1, 5), eq)
    print("gpu: max rel err vs equilibrium", e.max())
    assert (st == 0).all()
    assert (e <= 1e-3).all()
