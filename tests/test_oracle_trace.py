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


def replay_case(case, method, controller, n):
    net, y0, c, stride, tend, _, tol = case(n)
    y = y0.copy()
    cap = 30000
    tr = O.traced_batch(method, net, 0.0, tend, 1, c, y, n, tol, tol, capacity=cap, const_stride=stride,
                        controller=controller, max_attempts=cap, threads=4)
    assert np.all(tr["count"] > 0), "capacity too small for the case"
    tr["capacity"] = cap
    return net, y0, c, tr, y


@pytest.mark.parametrize("case", [nr_case, hhe6_case])
@pytest.mark.parametrize("method", [O.SIE, O.RK2])
@pytest.mark.parametrize("controller", [0, 1])
def test_float_replay_reproduces_the_run(case, method, controller):
    """Replaying a recorded run in float (one output interval) gives its
    final states and, attempt by attempt from the recorded start states, its
    y1 and y2 bit for bit: the replay takes the run's steps on its time
    grid."""
    n = 24
    net, y0, c, tr, yend = replay_case(case, method, controller, n)
    N = net.N
    yr = O.replay_batch(method, net, 0.0, c, y0, n, tr, controller=controller, threads=4, f64=False)
    assert np.array_equal(yr.reshape(-1).view(np.uint32), yend.view(np.uint32))
    ys = O.start_states(tr, y0, n, N, not (controller == 0 and method == O.SIE))
    y1, y2 = O.replay_batch(method, net, 0.0, c, None, n, tr, ystart=ys, controller=controller, threads=4,
                            f64=False)
    live = np.arange(tr["capacity"])[None, :] < tr["count"][:, None]
    for got, ref in ((y1, tr["y1"]), (y2, tr["y2"])):
        a, b = got[live], ref[live]
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a[~np.isnan(a)].view(np.uint32), b[~np.isnan(b)].view(np.uint32))


@pytest.mark.parametrize("method", [O.SIE, O.RK2])
def test_f64_replay_closed_form(method):
    """NR_test is linear (f = J y, J constant): an attempt from y over
    [t, t'] is y + h (I - h J)^-1 J y (SIE) or y + h J y (RK2) per sub-step.
    The float64 replay equals that closed form, evaluated in numpy float64
    with the step sizes of the float run's time grid (full step
    fl32(t' - t), half steps of half that), per attempt and over the whole
    run."""
    n = 16
    net, y0, c, tr, _ = replay_case(nr_case, method, 1, n)
    N = 2
    ys = O.start_states(tr, y0, n, N, True)
    y1, y2 = O.replay_batch(method, net, 0.0, c, None, n, tr, ystart=ys, controller=1, threads=4)
    yw = O.replay_batch(method, net, 0.0, c, y0, n, tr, controller=1, threads=4)
    cc = c.reshape(n, 4).astype(np.float64)
    f32 = np.float32

    def sub(J, y, h):
        if method == O.SIE:
            return y + h * np.linalg.solve(np.eye(2) - h * J, J @ y)
        return y + h * (J @ y)

    worst = 0.0
    for s in range(n):
        J = np.array([[cc[s, 0], cc[s, 1]], [cc[s, 2], cc[s, 3]]])
        t = f32(0.0)
        yc = y0.reshape(n, N)[s].astype(np.float64)
        for k in range(tr["count"][s]):
            tb = f32(t + tr["h"][s, k])
            hf = float(f32(tb - t))
            a1 = sub(J, ys[s, k].astype(np.float64), hf)
            a2 = sub(J, sub(J, ys[s, k].astype(np.float64), hf / 2), hf / 2)
            w1 = sub(J, yc, hf)
            w2 = sub(J, sub(J, yc, hf / 2), hf / 2)
            sc = max(1.0, float(np.abs(a1).max()), float(np.abs(a2).max()))
            worst = max(worst, float(np.abs(y1[s, k] - a1).max()) / sc, float(np.abs(y2[s, k] - a2).max()) / sc)
            if tr["accept"][s, k]:
                yc = 2 * w2 - w1
                t = tb
        assert np.allclose(yw[s], yc, rtol=1e-12, atol=1e-13), s
    assert worst <= 1e-11, worst  # float64 rounding times the condition of I - hJ (~1e3)
