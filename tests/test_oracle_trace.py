"""The oracle's per-attempt trace (the quantities the reference prints under
DEBUGBLOCK, cuda/c_baseline/common/common_gold.c:102-113) is consistent
with its untraced integration, which the golden vectors pin to the
reference bit for bit:

* the traced batch ends in the same states, step counts and status;
* the accepted steps tile [tnow, tend] (float sums, in order);
* the reference controller's accept decision is the refinement rule of
  common_gold.c:116-148 (reject iff the error test fails and fewer than 10
  refinements were made in a row), and the adaptive one's is err <= 1;
* replaying the recorded sequence through the reference's own compiled
  take_step and its acceptSolution rule reproduces each attempt's y1, y2
  and the final state bit for bit."""
import numpy as np
import pytest

import oracle_lib as O
from wind_amd import networks as W


def nr_case(n=64):
    y0, c, _ = W.nr_test_systems(n, seed=5)
    return O.nr_test(), y0, c, 4, 1.0, 1000, 1e-3


def hhe6_case(n=48):
    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(n, seed=6)
    return O.hhe6(p), y0, nH, 1, 1.0e4, 1, 1e-4


@pytest.mark.parametrize("case", [nr_case, hhe6_case])
@pytest.mark.parametrize("method", [O.SIE, O.RK2])
@pytest.mark.parametrize("controller", [0, 1])
def test_traced_batch_equals_untraced(case, method, controller):
    net, y0, c, stride, tend, nfix, tol = case()
    n = len(y0) // net.N
    ya = y0.copy()
    cap = 30000
    tot, steps, status = O.gold_integrate(method, net, 0.0, tend, nfix, c, ya, n, tol, tol, const_stride=stride,
                                          controller=controller, max_attempts=cap)
    yb = y0.copy()
    tr = O.traced_batch(method, net, 0.0, tend, nfix, c, yb, n, tol, tol, capacity=cap, const_stride=stride,
                        controller=controller, max_attempts=cap, threads=4)
    assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32))
    assert np.array_equal(tr["steps"], steps) and np.array_equal(tr["status"], status)
    assert np.all(tr["count"] > 0), "capacity too small for the case"
    assert np.array_equal(3 * tr["count"], steps)
    for s in range(n):
        k = tr["count"][s]
        acc = tr["accept"][s, :k].astype(bool)
        h = tr["h"][s, :k]
        t = np.float32(0.0)
        for hv in h[acc]:
            t = np.float32(t + hv)
        if status[s] == 0:
            assert t == np.float32(tend), (s, t)
        if controller == 0:
            fl = tr["flag"][s, :k]
            unsolved = 0
            for a, f in zip(acc, fl):
                assert a == (not (f and unsolved < 10))
                unsolved = unsolved + 1 if not a else 0
        else:
            assert np.array_equal(acc, tr["flag"][s, :k] == 0)


def test_trace_attempts_replay_through_reference_take_step():
    """SIE and RK2 on Katz96 (the network compiled into the reference's
    c_baseline): each recorded attempt's y1 and y2 equal the reference's own
    take_step (compiled from source) with n = 1 and 2, started from the
    state the recorded accept decisions leave (acceptSolution: y2 for SIE,
    2 y2 - y1 for RK2)."""
    if not O.ref_available():
        pytest.skip("reference c_baseline not built (oracle/_ref)")
    import ctypes as C
    import os

    G = np.load(os.path.join(os.path.dirname(__file__), "golden", "c_baseline_kats.npz"))
    ys, cs = G["sys_y"][:12], G["sys_c"][:12]
    nk = len(ys)
    J = np.zeros(25, np.float32)
    Inv = np.zeros(25, np.float32)
    f = np.zeros(5, np.float32)
    for method in (O.SIE, O.RK2):
        R = O.ref(method)
        yk = ys.reshape(-1).copy()
        tr = O.traced_batch(method, O.katz96(), 0.0, 50.0, 1, cs.reshape(-1).copy(), yk, nk, 1e-4, 1e-4,
                            capacity=20000, const_stride=10)
        assert np.all(tr["count"] > 0)
        for s in range(nk):
            y = ys[s].copy()
            c = np.ascontiguousarray(cs[s], np.float32)
            tnow = np.float32(0.0)
            for k in range(tr["count"][s]):
                tb = np.float32(tnow + tr["h"][s, k])
                y1 = y.copy()
                y2 = y.copy()
                R.take_step(C.c_float(tnow), C.c_float(tb), 1, y1.ctypes.data, c.ctypes.data, f.ctypes.data,
                            J.ctypes.data, Inv.ctypes.data, 5)
                R.take_step(C.c_float(tnow), C.c_float(tb), 2, y2.ctypes.data, c.ctypes.data, f.ctypes.data,
                            J.ctypes.data, Inv.ctypes.data, 5)
                assert np.array_equal(y1.view(np.uint32), tr["y1"][s, k].view(np.uint32)), (method, s, k)
                assert np.array_equal(y2.view(np.uint32), tr["y2"][s, k].view(np.uint32)), (method, s, k)
                if tr["accept"][s, k]:
                    y = y2 if method == O.SIE else (np.float32(2) * y2 - y1).astype(np.float32)
                    tnow = tb
            assert np.array_equal(y.view(np.uint32), yk[5 * s:5 * s + 5].view(np.uint32))
