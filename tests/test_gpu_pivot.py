"""Partial pivoting in the fast-mode LU solves (WIND_FLAG_PIVOT).

A linear system whose Jacobian is [[1, 1], [1, -3]] has I - hJ with a zero
leading pivot at h = 1.  One attempt of h = 1 (accepted: huge tolerances,
adaptive controller, so the kept state 2 y2 - y1 uses the full step's
solve):
* with pivoting the kernels give the float64 pivoted solution (numpy) to
  float rounding: the lane kernel of the compiled-in NR_test network, the
  generated network's lane kernel (N = 2) and its one-workgroup-per-system
  kernel with the wave-shuffle pivot search (tiled, N = 20);
* without pivoting the zero pivot makes the full step non-finite, the
  attempt fails and, with one attempt allowed, the system reports
  EXHAUSTED with its state unchanged.
On regular workloads the pivoted solvers agree with the unpivoted ones to
the integration tolerance."""
import numpy as np
import pytest

from wind_amd import api
from wind_amd import networks as W

pytestmark = pytest.mark.gpu

J2 = np.array([[1.0, 1.0], [1.0, -3.0]])
Y0 = np.array([1.0, 1.0])


def f64_attempt(h):
    """2 y2 - y1 of one SIE attempt of size h, float64, pivoted solves."""
    def sub(y, hs):
        f = J2 @ y
        return y + hs * np.linalg.solve(np.eye(2) - hs * J2, f)

    y1 = sub(Y0, h)
    y2 = sub(sub(Y0, h / 2), h / 2)
    return 2 * y2 - y1


def run(net, T, c, flags, kernel=api.KERNEL_AUTO):
    N = 2 * T
    y = np.tile(Y0.astype(np.float32), T)
    steps = np.zeros(1, np.int32)
    status = np.full(1, -1, np.int32)
    lib = api.load()
    import ctypes as C
    o = api.WindOptions()
    lib.wind_options_init(C.byref(o))
    o.mode, o.controller, o.kernel, o.flags, o.max_attempts = api.MODE_FAST, api.CTRL_ADAPTIVE, kernel, flags, 1
    o.steps_out, o.status_out = steps.ctypes.data, status.ctypes.data
    cc = np.asarray(c, np.float32)
    rc = lib.wind_integrate(net.handle, 0.0, 1.0, 1, cc.ctypes.data, y.ctypes.data, 1, N, 1e30, 1e30, C.byref(o))
    assert rc == 0, lib.wind_last_error().decode()
    return y.reshape(T, 2), int(status[0])


CASES = [("nr_test builtin", lambda: api.Network.builtin(api.NET_NR_TEST), 1, [1.0, 1.0, 1.0, -3.0]),
         ("pivot_demo (lane)", lambda: api.Network.plugin("pivot_demo"), 1, [1.0, 1.0]),
         ("pivot_demo_t10 (workgroup)", lambda: api.Network.plugin("pivot_demo_t10"), 10, [1.0, 1.0])]


@pytest.mark.parametrize("name,make,T,c", CASES)
def test_zero_leading_pivot(name, make, T, c):
    net = make()
    want = f64_attempt(1.0)
    y, st = run(net, T, c, api.FLAG_PIVOT)
    print(f"{name}: pivoted {y[0]}, float64 {want}", flush=True)
    assert st == api.STATUS_OK
    assert np.allclose(y.astype(np.float64), want[None, :], rtol=1e-6, atol=1e-6)
    y, st = run(net, T, c, 0)
    print(f"{name}: unpivoted status {st}, state {y[0]}", flush=True)
    assert st == api.STATUS_EXHAUSTED
    assert np.array_equal(y, np.tile(Y0.astype(np.float32), (T, 1)))


def test_pivot_rejected_in_exact_mode():
    net = api.Network.builtin(api.NET_NR_TEST)
    y = Y0.astype(np.float32).copy()
    with pytest.raises(api.WindError, match="fast mode"):
        import ctypes as C
        lib = api.load()
        o = api.WindOptions()
        lib.wind_options_init(C.byref(o))
        o.flags = api.FLAG_PIVOT
        cc = np.array([1.0, 1.0, 1.0, -3.0], np.float32)
        api._check(lib.wind_integrate(net.handle, 0.0, 1.0, 1, cc.ctypes.data, y.ctypes.data, 1, 2, 1e-3, 1e-3,
                                      C.byref(o)))


def test_pivoted_solvers_on_regular_workloads():
    """hhe6 (lane kernel) and a mass-action network at N = 32 (the LDS LU
    instead of the blocked MFMA solver): pivoted and unpivoted fast runs end
    within 1e-3 of each other (each is within the tolerance of the
    solution)."""
    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(2048, seed=44)
    net = api.Network.builtin(api.NET_HHE6, p)
    out = []
    for flags in (0, api.FLAG_PIVOT):
        y = y0.copy()
        api.integrate(net, 0.0, 1.0e5, 1, nH, y, 2048, 1e-5, 1e-5, mode=api.MODE_FAST, controller=api.CTRL_ADAPTIVE,
                      flags=flags)
        out.append(y.reshape(-1, 6).astype(np.float64))
    err = np.abs(out[0] - out[1]) / np.maximum(np.abs(out[1]), 1e-3)
    assert err.max() <= 1e-3, err.max()
    rx, k = W.mixed_network(32)
    ys, sc = W.mass_action_systems(256, 32, seed=45, scale_range=W.MIXED_SCALE_RANGE)
    net = api.Network.mass_action(32, rx, k)
    out = []
    for flags in (0, api.FLAG_PIVOT):
        y = ys.copy()
        api.integrate(net, 0.0, 1.0e8, 1, sc, y, 256, 1e-10, 8e-5, mode=api.MODE_FAST, controller=api.CTRL_ADAPTIVE,
                      flags=flags)
        out.append(y.reshape(-1, 32).astype(np.float64))
    tot = np.abs(out[1]).sum(axis=1, keepdims=True)
    err = np.abs(out[0] - out[1]) / np.maximum(np.abs(out[1]), 1e-6 * tot)
    assert err.max() <= 1e-3, err.max()
