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
On regular workloads the pivoted solvers meet the same accuracy criterion
as the unpivoted ones."""
import os

import numpy as np
import pytest

from wind_amd import api
from wind_amd import networks as W

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

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
    """Pivoting changes only the rounding of well-conditioned solves.
    hhe6 (lane kernel) on the 512 FP64-truth systems at the benchmark's
    tolerance (2e-6): the pivoted and unpivoted runs each meet the
    benchmark's criterion (final-state relative error <= 1e-3 against FP64,
    floor 1e-3) on every system, and end within 1e-3 of each other on at
    least 99% of them (two float runs whose rounding differs take different
    step sequences).  A mass-action network at N = 32 (the LDS LU instead of
    the blocked MFMA solver) at the mixed workload's tolerances: both runs
    end within 1e-3 of each other (floor 1e-6 of the system's total) on at
    least 99% of the systems."""
    T = np.load(os.path.join(ROOT, "tests", "golden", "truth_hhe6.npz"))
    idx, yt = T["idx"], T["y"]
    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(int(T["nsys"]), seed=int(T["seed"]))
    net = api.Network.builtin(api.NET_HHE6, p)
    n = len(idx)
    out = []
    for flags in (0, api.FLAG_PIVOT):
        y = y0.reshape(-1, 6)[idx].reshape(-1).copy()
        st = np.zeros(n, np.int32)
        api.integrate(net, 0.0, float(T["tend"]), 1, nH[idx].copy(), y, n, 2e-6, 2e-6, mode=api.MODE_FAST,
                      controller=api.CTRL_ADAPTIVE, flags=flags, max_attempts=100000, status_out=st)
        assert (st == 0).all()
        out.append(y.reshape(-1, 6).astype(np.float64))
        e = np.max(np.abs(out[-1] - yt) / np.maximum(np.abs(yt), 1e-3), axis=1)
        print(f"hhe6 flags {flags}: max error vs FP64 {e.max():.2e}", flush=True)
        assert e.max() <= 1e-3
    d = np.max(np.abs(out[0] - out[1]) / np.maximum(np.abs(out[1]), 1e-3), axis=1)
    print(f"hhe6 pivoted vs unpivoted: median {np.median(d):.2e}, max {d.max():.2e}", flush=True)
    assert np.mean(d <= 1e-3) >= 0.99
    rx, k = W.mixed_network(32)
    ys, sc = W.mass_action_systems(256, 32, seed=45, scale_range=W.MIXED_SCALE_RANGE)
    net = api.Network.mass_action(32, rx, k)
    out = []
    for flags in (0, api.FLAG_PIVOT):
        y = ys.copy()
        api.integrate(net, 0.0, W.MASS_ACTION_TEND, 1, sc, y, 256, 1e-10, 8e-5, mode=api.MODE_FAST,
                      controller=api.CTRL_ADAPTIVE, flags=flags)
        out.append(y.reshape(-1, 32).astype(np.float64))
    tot = np.abs(out[1]).sum(axis=1, keepdims=True)
    d = np.max(np.abs(out[0] - out[1]) / np.maximum(np.abs(out[1]), 1e-6 * tot), axis=1)
    print(f"mass action N=32 pivoted vs unpivoted: median {np.median(d):.2e}, max {d.max():.2e}", flush=True)
    assert np.all(np.isfinite(d))
    assert np.mean(d <= 1e-3) >= 0.99
