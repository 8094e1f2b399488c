"""Per-attempt trace and replay on the GPU, through the C-ABI (wind_trace).

The reference's own parity method is the per-attempt trace: its DEBUGBLOCK
builds print the full-step and two-half-step states of every attempt
(cuda/common/kernel.cu:214-249 against cuda/c_baseline/common/common_gold.c:
102-113).  Here:

* exact mode records the oracle's trace bit for bit (y1, y2 and the error
  test of every attempt), and replaying the oracle's step sequence gives the
  oracle's final states bit for bit;
* fast mode (the benchmarked kernels: the lane kernel on hhe6, SIE and RK2,
  and the dense LDS/MFMA solver on mass-action networks at N = 32, 64, 157)
  replays the oracle's float32 run - the same step sizes and the same
  accept decisions - and is compared with it per species as
  |y_gpu - y_oracle| / max(|y_oracle|, floor), the floor being the
  controller's absolute tolerance ABS for hhe6 and 1e-6 of the system's
  total abundance for mass action (the accuracy floor of the benchmark's
  criterion):
  - per attempt: every attempt starts from the oracle's state before it
    (wind_trace.y_start); the full-step and two-half-step states y1, y2 of
    every accepted attempt are within 1e-5 of the oracle's, relative to the
    attempt's own magnitudes max(|y_oracle|, |y_start|, |y_oracle - y_start|,
    floor) (an attempt computes y_start + (its increment), so its rounding
    scales with both; for explicit RK2 the increment nearly cancels the
    state in places); the kernel's own error test agrees with the oracle's on
    at least 99.9% of all attempts;
  - whole runs: the kernel carries its own state through the oracle's
    accepted-step sequence; every system ends within 5e-5 of the oracle and
    99.9% of the species within 1e-5 (hhe6's rates k = A T^b exp(-C/T) use
    the hardware exp2 in fast mode, whose relative difference from the
    oracle's portable exp reaches ~1e-5 per evaluation near the float
    exponent's range, and a state integrates such differences over its
    steps).  Explicit RK2 on this stiff network runs at its stability limit,
    where replaying steps without the controller's feedback amplifies any
    rounding difference, so RK2 is compared per attempt only.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from wind_amd import api
from wind_amd import networks as W

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
REL_BOUND = 1e-5


def hhe6_sample():
    """The 512 systems of the 2^20-system hhe6 workload that the FP64 truth
    fixture holds (every 2048th)."""
    d = np.load(os.path.join(ROOT, "tests", "golden", "truth_hhe6.npz"))
    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(int(d["nsys"]), seed=int(d["seed"]))
    idx = d["idx"]
    return p, np.ascontiguousarray(y0.reshape(-1, 6)[idx].reshape(-1)), np.ascontiguousarray(nH[idx]), float(d["tend"])


def oracle_trace(method, onet, tend, nfix, c, y0, n, abs_, rel, controller, stride, max_attempts=0,
                 with_states=False):
    """The oracle's float32 run with every attempt recorded (capacity = its
    largest attempt count)."""
    if max_attempts:
        cap = max_attempts
    else:
        y = y0.copy()
        _, steps, _ = O.gold_integrate(method, onet, 0.0, tend, nfix, c, y, n, abs_, rel, const_stride=stride,
                                       controller=controller, threads=THREADS)
        cap = int(steps.max()) // 3
    y = y0.copy()
    tr = O.traced_batch(method, onet, 0.0, tend, nfix, c, y, n, abs_, rel, capacity=cap, const_stride=stride,
                        max_attempts=max_attempts, controller=controller, threads=THREADS, with_states=with_states)
    assert np.all(tr["count"] >= 0)
    tr["y"] = y
    tr["capacity"] = cap
    return tr


def replay(net, method, mode, tend, nfix, c, y0, n, abs_, rel, controller, tr, record=False, stride=0):
    cap = tr["capacity"]
    y = y0.copy()
    steps = np.zeros(n, np.int32)
    t = {"capacity": cap, "h": tr["h"], "accept": tr["accept"], "count": tr["count"]}
    if record:
        t["y1"] = np.zeros((n, cap, net.N), np.float32)
        t["y2"] = np.zeros((n, cap, net.N), np.float32)
    t["flag"] = np.full((n, cap), -1, np.int32)
    api.integrate(net, 0.0, tend, nfix, c, y, n, abs_, rel, method=method, mode=mode, controller=controller,
                  steps_out=steps, trace=t, const_stride=stride)
    assert np.array_equal(steps, 3 * tr["count"])
    return y, t


def start_states(tr, y0, n, N, extrap):
    """The oracle's state before each recorded attempt: the initial state,
    then the value kept at the last accepted attempt (2 y2 - y1 in float32,
    or y2 for the reference controller's SIE)."""
    cap = tr["capacity"]
    acc = tr["accept"].astype(bool) & (np.arange(cap)[None, :] < tr["count"][:, None])
    kept = (np.float32(2) * tr["y2"] - tr["y1"]).astype(np.float32) if extrap else tr["y2"]
    idx = np.where(acc, np.arange(cap)[None, :], -1)
    last = np.maximum.accumulate(idx, axis=1)
    prev = np.concatenate([np.full((n, 1), -1), last[:, :-1]], axis=1)
    ys = np.take_along_axis(kept, np.maximum(prev, 0)[:, :, None], axis=1)
    init = np.broadcast_to(y0.reshape(n, 1, N), ys.shape)
    return np.ascontiguousarray(np.where(prev[:, :, None] >= 0, ys, init), dtype=np.float32)


def rel_dev(y, yr, floor):
    y = np.asarray(y, np.float64)
    yr = np.asarray(yr, np.float64)
    return np.abs(y - yr) / np.maximum(np.abs(yr), floor)


def decision_mismatch(t, tr):
    """Fraction of replayed attempts where the kernel's own error test
    disagrees with the oracle's."""
    cap = tr["capacity"]
    mask = np.arange(cap)[None, :] < tr["count"][:, None]
    return float(np.mean(t["flag"][mask] != tr["flag"][mask]))


# ------------------------------------------------------------- exact mode
@pytest.mark.parametrize("method", [api.SIE, api.RK2])
@pytest.mark.parametrize("controller", [api.CTRL_REFERENCE, api.CTRL_ADAPTIVE])
def test_exact_trace_and_replay_bit_identical(method, controller):
    """Exact mode (lane kernel, hhe6): recording its own run gives the
    oracle's trace bit for bit; replaying the oracle's sequence gives the
    oracle's final states bit for bit."""
    p, y0, c, _ = hhe6_sample()
    n = 128
    y0, c = y0[: 6 * n].copy(), c[:n].copy()
    tend, tol = 1.0e4, 1e-4
    onet = O.hhe6(p)
    tr = oracle_trace(method, onet, tend, 1, c, y0, n, tol, tol, controller, 1, max_attempts=20000, with_states=True)
    net = api.Network.builtin(api.NET_HHE6, p)
    cap = tr["capacity"]
    # record only: the kernel's own controller
    y = y0.copy()
    steps = np.zeros(n, np.int32)
    t = {"capacity": cap, "y1": np.zeros((n, cap, 6), np.float32), "y2": np.zeros((n, cap, 6), np.float32),
         "flag": np.zeros((n, cap), np.int32)}
    api.integrate(net, 0.0, tend, 1, c, y, n, tol, tol, method=method, controller=controller, steps_out=steps,
                  trace=t, max_attempts=20000)
    assert np.array_equal(steps, tr["steps"])
    mask = np.arange(cap)[None, :] < tr["count"][:, None]
    for k in ("y1", "y2"):
        a, b = t[k][mask], tr[k][mask]
        assert np.array_equal(np.isnan(a), np.isnan(b)), k
        assert np.array_equal(a[~np.isnan(a)].view(np.uint32), b[~np.isnan(b)].view(np.uint32)), k
    assert np.array_equal(t["flag"][mask], tr["flag"][mask])
    # replay
    yr, _ = replay(net, method, api.MODE_EXACT, tend, 1, c, y0, n, tol, tol, controller, tr)
    fin = np.isfinite(tr["y"])
    assert np.array_equal(np.isfinite(yr), fin)
    assert np.array_equal(yr[fin].view(np.uint32), tr["y"][fin].view(np.uint32))


# ------------------------------------------------------------ per attempt
def check_attempts(tr, t, ystart, floor):
    """Per-attempt parity of the recorded y1 / y2 (accepted attempts) and
    agreement of the own error tests (all attempts); returns (worst, agree)."""
    K = tr["capacity"]
    live = np.arange(K)[None, :] < tr["count"][:, None]
    acc = tr["accept"].astype(bool) & live
    worst = 0.0
    for k in ("y1", "y2"):
        ref = tr[k][acc].astype(np.float64)
        got = t[k][acc].astype(np.float64)
        st = ystart[acc].astype(np.float64)
        assert np.all(np.isfinite(ref)) and np.all(np.isfinite(got)), k
        fl = floor if np.ndim(floor) == 0 else np.broadcast_to(floor[:, None, :], tr[k].shape)[acc]
        scale = np.maximum.reduce([np.abs(ref), np.abs(st), np.abs(ref - st), np.broadcast_to(fl, ref.shape)])
        dev = np.abs(got - ref) / scale
        worst = max(worst, float(dev.max()))
        assert dev.max() <= REL_BOUND, (k, np.unravel_index(np.argmax(dev), dev.shape), float(dev.max()))
    agree = 1.0 - decision_mismatch(t, tr)
    assert agree >= 0.999
    return worst, agree


def traced_call(net, method, tend, c, y0, n, abs_, rel, tr, ystart):
    K = tr["capacity"]
    t = {"capacity": K, "h": tr["h"], "accept": tr["accept"], "count": tr["count"], "y_start": ystart,
         "y1": np.zeros((n, K, net.N), np.float32), "y2": np.zeros((n, K, net.N), np.float32),
         "flag": np.full((n, K), -1, np.int32)}
    steps = np.zeros(n, np.int32)
    api.integrate(net, 0.0, tend, 1, c, y0.copy(), n, abs_, rel, method=method, mode=api.MODE_FAST,
                  controller=api.CTRL_ADAPTIVE, steps_out=steps, trace=t)
    assert np.array_equal(steps, 3 * tr["count"])
    return t


@pytest.mark.parametrize("method,tol,cap", [(api.SIE, 2e-6, 0), (api.RK2, 1e-4, 3000)])
def test_fast_lane_attempts_hhe6(method, tol, cap):
    """The benchmarked fast lane kernel on the 512 FP64-fixture systems over
    1 Myr (SIE at the benchmark's tolerance 2e-6, every attempt; RK2 at 1e-4,
    the first 3000 attempts of each system), each attempt started from the
    oracle's state."""
    p, y0, c, tend = hhe6_sample()
    n = len(c)
    tr = oracle_trace(method, O.hhe6(p), tend, 1, c, y0, n, tol, tol, api.CTRL_ADAPTIVE, 1, max_attempts=cap,
                      with_states=True)
    ystart = start_states(tr, y0, n, 6, True)
    net = api.Network.builtin(api.NET_HHE6, p)
    t = traced_call(net, method, tend, c, y0, n, tol, tol, tr, ystart)
    worst, agree = check_attempts(tr, t, ystart, tol)
    print(f"hhe6 fast per-attempt m{method}: {int(tr['count'].sum())} attempts of {n} systems "
          f"({int(tr['accept'].sum())} accepted); max rel dev {worst:.3e}; own-test decisions agreeing {agree:.6f}",
          flush=True)


DENSE_ATTEMPT_CASES = [(32, 48, 400), (64, 32, 240), (157, 8, 150)]


def dense_case(N, ns):
    if N == 157:
        rx, k = W.mass_action_network(157, seed=W.DENSE_NET_SEED)
    else:
        rx, k = W.mixed_network(N)
    y0, sc = W.mass_action_systems(ns, N, seed=3000 + N, scale_range=W.MIXED_SCALE_RANGE)
    return rx, k, y0, sc


@pytest.mark.parametrize("N,ns,cap", DENSE_ATTEMPT_CASES)
def test_fast_dense_attempts(N, ns, cap):
    """The dense LDS/MFMA solver (FP32 MFMA, no refinement) per attempt, on
    the first attempts of the mixed workload's networks (N = 157: the dense
    workload's network) at the benchmark's tolerances (abs 1e-10, rel 8e-5);
    floor 1e-6 of the system's total abundance."""
    rx, k, y0, sc = dense_case(N, ns)
    tend = W.MASS_ACTION_TEND
    tr = oracle_trace(api.SIE, O.mass_action(N, rx, k), tend, 1, sc, y0, ns, 1e-10, 8e-5, api.CTRL_ADAPTIVE, 1,
                      max_attempts=cap, with_states=True)
    ystart = start_states(tr, y0, ns, N, True)
    net = api.Network.mass_action(N, rx, k)
    t = traced_call(net, api.SIE, tend, sc, y0, ns, 1e-10, 8e-5, tr, ystart)
    floor = 1e-6 * np.abs(ystart.astype(np.float64)).sum(axis=2)  # [ns, K]
    floor = np.broadcast_to(floor[:, :, None], ystart.shape)
    K = tr["capacity"]
    live = np.arange(K)[None, :] < tr["count"][:, None]
    acc = tr["accept"].astype(bool) & live
    worst = 0.0
    for key in ("y1", "y2"):
        ref = tr[key][acc].astype(np.float64)
        got = t[key][acc].astype(np.float64)
        st = ystart[acc].astype(np.float64)
        scale = np.maximum.reduce([np.abs(ref), np.abs(st), np.abs(ref - st), floor[acc]])
        dev = np.abs(got - ref) / scale
        worst = max(worst, float(dev.max()))
        assert dev.max() <= REL_BOUND, (key, np.unravel_index(np.argmax(dev), dev.shape), float(dev.max()))
    agree = 1.0 - decision_mismatch(t, tr)
    assert agree >= 0.999
    print(f"dense N={N} fast per-attempt: {int(tr['count'].sum())} attempts of {ns} systems; max rel dev {worst:.3e}; "
          f"own-test decisions agreeing {agree:.6f}", flush=True)


# -------------------------------------------------------------- whole runs
def test_fast_lane_replay_hhe6_sie():
    """The benchmarked fast lane kernel (SIE, tol 2e-6) carries its own state
    through the oracle's accepted-step sequence over 1 Myr."""
    p, y0, c, tend = hhe6_sample()
    n, tol = len(c), 2e-6
    tr = oracle_trace(api.SIE, O.hhe6(p), tend, 1, c, y0, n, tol, tol, api.CTRL_ADAPTIVE, 1)
    net = api.Network.builtin(api.NET_HHE6, p)
    y, t = replay(net, api.SIE, api.MODE_FAST, tend, 1, c, y0, n, tol, tol, api.CTRL_ADAPTIVE, tr)
    dev = rel_dev(y, tr["y"], tol)
    accepted = tr["accept"].sum(axis=1)
    frac = float(np.mean(dev <= REL_BOUND))
    print(f"hhe6 fast replay SIE: {n} systems, attempts mean {tr['count'].mean():.0f} max {tr['count'].max()}, "
          f"accepted steps min {accepted.min()}; rel dev max {dev.max():.3e}, median {np.median(dev):.3e}, "
          f"within 1e-5: {frac:.5f}; own-test decisions differing {decision_mismatch(t, tr):.2e}", flush=True)
    assert np.all(np.isfinite(y))
    assert frac >= 0.999
    assert dev.max() <= 5e-5, np.unravel_index(np.argmax(dev), dev.shape)


DENSE_CASES = [(32, 96, 0), (64, 32, 240), (157, 8, 150)]


@pytest.mark.parametrize("N,ns,cap", DENSE_CASES)
def test_fast_dense_replay(N, ns, cap):
    """The dense LDS/MFMA solver carries its own state through the oracle's
    accepted-step sequence: over the whole interval for N = 32 (the systems
    with at least 100 accepted steps) and over the first 240 / 150 attempts
    for N = 64 / 157 (at least 100 accepted steps each)."""
    rx, k, y0, sc = dense_case(N, ns)
    tend = W.MASS_ACTION_TEND
    tr = oracle_trace(api.SIE, O.mass_action(N, rx, k), tend, 1, sc, y0, ns, 1e-10, 8e-5, api.CTRL_ADAPTIVE, 1,
                      max_attempts=cap)
    accepted = tr["accept"].sum(axis=1)
    keep = accepted >= 100
    assert keep.sum() >= min(ns, 48)
    net = api.Network.mass_action(N, rx, k)
    y, t = replay(net, api.SIE, api.MODE_FAST, tend, 1, sc, y0, ns, 1e-10, 8e-5, api.CTRL_ADAPTIVE, tr)
    Y = y.reshape(ns, N)[keep]
    Yr = tr["y"].reshape(ns, N)[keep]
    floor = 1e-6 * np.abs(Yr.astype(np.float64)).sum(axis=1, keepdims=True)
    dev = rel_dev(Y, Yr, floor)
    frac = float(np.mean(dev <= REL_BOUND))
    print(f"dense N={N} fast replay: {int(keep.sum())} of {ns} systems, attempts mean {tr['count'].mean():.0f}, "
          f"accepted min {accepted[keep].min()}; rel dev max {dev.max():.3e}, median {np.median(dev):.3e}, within "
          f"1e-5: {frac:.5f}; own-test decisions differing {decision_mismatch(t, tr):.2e}", flush=True)
    assert np.all(np.isfinite(y))
    assert frac >= 0.999
    assert dev.max() <= 5e-5


def test_trace_rejected_where_unsupported():
    """Kernels without a trace path say so instead of ignoring it."""
    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(64, seed=3)
    net = api.Network.builtin(api.NET_HHE6, p)
    t = {"capacity": 4, "flag": np.zeros((64, 4), np.int32)}
    with pytest.raises(api.WindError, match="block kernel has no trace"):
        api.integrate(net, 0.0, 1.0, 1, nH, y0.copy(), 64, 1e-3, 1e-3, kernel=api.KERNEL_BLOCK, trace=t)
    with pytest.raises(api.WindError, match="replay needs h, accept and count"):
        api.integrate(net, 0.0, 1.0, 1, nH, y0.copy(), 64, 1e-3, 1e-3,
                      trace={"capacity": 4, "h": np.zeros((64, 4), np.float32)})
