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
  accept decisions - and is compared per species with it and with the same
  steps in float64 (the oracle's replay in double on the run's float time
  grid), as |y - y_ref| / max(|y_ref|, floor), the floor being the
  controller's absolute tolerance ABS for hhe6 and 1e-6 of the system's
  total abundance for mass action (the accuracy floor of the benchmark's
  criterion):
  - per attempt: every attempt starts from the oracle's state before it
    (wind_trace.y_start); the full-step and two-half-step states y1, y2 of
    every accepted attempt are within 1e-5 of the float oracle's, relative
    to the attempt's own magnitudes max(|y_oracle|, |y_start|,
    |y_oracle - y_start|, floor) (an attempt computes y_start + (its
    increment), so its rounding scales with both; for explicit RK2 the
    increment nearly cancels the state in places); where the two float
    results differ by more than 1e-5 (at most 1e-4 of the values: each
    carries its own float rounding, and the oracle's float result is itself
    up to ~8e-6 from the float64 attempt), the kernel is within 1e-5 of the
    float64 attempt; the kernel's own error test agrees with the oracle's on
    at least 99.9% of all attempts;
  - whole runs: the kernel carries its own state through the oracle's
    accepted-step sequence and ends as close to the float64 replay of that
    sequence as the float oracle's own run does (largest and 99th-percentile
    difference within 1.5x the oracle's; the float oracle itself ends up to
    ~1e-4 from the float64 replay on hhe6), and within 5e-5 of the float
    oracle.  Explicit RK2 on this stiff network runs at its stability limit,
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


def replay(net, method, mode, tend, nfix, c, y0, n, abs_, rel, controller, tr, record=False, stride=0, flags=0):
    cap = tr["capacity"]
    y = y0.copy()
    steps = np.zeros(n, np.int32)
    t = {"capacity": cap, "h": tr["h"], "accept": tr["accept"], "count": tr["count"]}
    if record:
        t["y1"] = np.zeros((n, cap, net.N), np.float32)
        t["y2"] = np.zeros((n, cap, net.N), np.float32)
    t["flag"] = np.full((n, cap), -1, np.int32)
    api.integrate(net, 0.0, tend, nfix, c, y, n, abs_, rel, method=method, mode=mode, controller=controller,
                  steps_out=steps, trace=t, const_stride=stride, flags=flags)
    assert np.array_equal(steps, 3 * tr["count"])
    return y, t


start_states = O.start_states


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
def attempt_parity(method, onet, c, n, tr, t, ystart, floor, label):
    """Per-attempt parity of the recorded y1 / y2 of every accepted attempt,
    against the float32 oracle and against the same attempt in float64 (the
    oracle's replay in double from the same start state on the same time
    grid): every species is within REL_BOUND of the float oracle, or, where
    the two float results differ by more (both carry float rounding: the
    oracle's own float result differs from the float64 attempt by up to
    ~1e-5), the kernel's result is within REL_BOUND of the float64 attempt;
    at most 1e-4 of the values take that second branch.  Scale: the
    attempt's own magnitudes max(|y_oracle|, |y_start|, |y_oracle - y_start|,
    floor).  The kernel's own error test agrees with the oracle's on at
    least 99.9% of all attempts.  Returns the statistics."""
    K = tr["capacity"]
    live = np.arange(K)[None, :] < tr["count"][:, None]
    acc = tr["accept"].astype(bool) & live
    ex = O.replay_batch(method, onet, 0.0, c, None, n, tr, ystart=ystart, controller=api.CTRL_ADAPTIVE,
                        threads=THREADS)
    st = {"worst": 0.0, "worst_f64": 0.0, "oracle_f64": 0.0, "over": 0, "values": 0}
    for k, e in zip(("y1", "y2"), ex):
        ref = tr[k][acc].astype(np.float64)
        got = t[k][acc].astype(np.float64)
        y64 = e[acc]
        ys = ystart[acc].astype(np.float64)
        assert np.all(np.isfinite(ref)) and np.all(np.isfinite(got)), k
        fl = floor if np.ndim(floor) == 0 else np.asarray(floor)[acc]
        scale = np.maximum.reduce([np.abs(ref), np.abs(ys), np.abs(ref - ys), np.broadcast_to(fl, ref.shape)])
        dev = np.abs(got - ref) / scale
        dev64 = np.abs(got - y64) / scale
        over = dev > REL_BOUND
        st["worst"] = max(st["worst"], float(dev.max()))
        st["worst_f64"] = max(st["worst_f64"], float(dev64.max()))
        st["oracle_f64"] = max(st["oracle_f64"], float((np.abs(ref - y64) / scale).max()))
        st["over"] += int(over.sum())
        st["values"] += over.size
        bad = over & (dev64 > REL_BOUND)
        assert not bad.any(), (label, k, np.argwhere(bad)[:4], dev[bad][:4], dev64[bad][:4])
    st["agree"] = 1.0 - decision_mismatch(t, tr)
    print(f"{label}: {int(tr['count'].sum())} attempts of {n} systems ({int(acc.sum())} accepted); vs float oracle "
          f"max {st['worst']:.3e} ({st['over']} of {st['values']} values above {REL_BOUND:g}); vs float64 attempt "
          f"max {st['worst_f64']:.3e} (float oracle vs float64: max {st['oracle_f64']:.3e}); own-test decisions "
          f"agreeing {st['agree']:.6f}", flush=True)
    assert st["over"] <= 1e-4 * st["values"]
    assert st["agree"] >= 0.999
    return st


def traced_call(net, method, tend, c, y0, n, abs_, rel, tr, ystart, flags=0):
    K = tr["capacity"]
    t = {"capacity": K, "h": tr["h"], "accept": tr["accept"], "count": tr["count"], "y_start": ystart,
         "y1": np.zeros((n, K, net.N), np.float32), "y2": np.zeros((n, K, net.N), np.float32),
         "flag": np.full((n, K), -1, np.int32)}
    steps = np.zeros(n, np.int32)
    api.integrate(net, 0.0, tend, 1, c, y0.copy(), n, abs_, rel, method=method, mode=api.MODE_FAST,
                  controller=api.CTRL_ADAPTIVE, steps_out=steps, trace=t, flags=flags)
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
    attempt_parity(method, O.hhe6(p), c, n, tr, t, ystart, tol, f"hhe6 fast per-attempt {('SIE', 'RK2')[method]}")


# 49 <= N <= 160 runs on the register-resident solver by default (tile
# counts per dimension 4 .. 10; N = 80, 100, 129 cover 5, 7 and 9, odd
# counts leaving some of a wave's tile slots unused) and on the LDS-resident
# one with FLAG_DENSE_LDS; N = 32 on the LDS-resident one.
DENSE_ATTEMPT_CASES = [(32, 48, 400, 0), (32, 48, 400, api.FLAG_DENSE_LDS), (60, 16, 200, 0), (64, 32, 240, 0), (80, 16, 200, 0), (100, 12, 160, 0),
                       (129, 8, 150, 0), (150, 8, 150, 0), (157, 8, 150, 0), (157, 8, 150, api.FLAG_DENSE_LDS)]

# Networks with more reactions than the default 4 N: at N = 150, R = 1000
# the register solver's LDS footprint leaves room for three workgroups per
# CU, not four, so it runs its three-per-CU build at ten tiles; at N = 60,
# R = 1000 likewise three, not six, at four tiles; at N = 96,
# R = 3600 the Jacobian has 8,343 non-zero elements, more than the register
# solver's compact element buffer holds (8,191), so another kernel runs.
WIDE_R = {60: 1000, 96: 3600, 150: 1000}


def dense_case(N, ns):
    if N == 157:
        rx, k = W.mass_action_network(157, seed=W.DENSE_NET_SEED)
    elif N in WIDE_R:
        rx, k = W.mass_action_network(N, seed=1000 + N, R=WIDE_R[N])
    else:
        rx, k = W.mixed_network(N)
    y0, sc = W.mass_action_systems(ns, N, seed=3000 + N, scale_range=W.MIXED_SCALE_RANGE)
    return rx, k, y0, sc


def test_fast_dense_beyond_compact_buffer():
    """A network whose Jacobian has more non-zero elements than the register
    solver's compact buffer holds (N = 96, R = 3600: 8,343 > 8,191) runs
    another fast kernel: every system finishes, and its final state is
    within 1e-3 of the float oracle's run under the same controller and
    tolerances (relative to max(|y|, 1e-6 of the system's total abundance);
    the runs differ by rounding only)."""
    N, ns, tend = 96, 8, 1.0e5
    rx, k, y0, sc = dense_case(N, ns)
    net = api.Network.mass_action(N, rx, k)
    y = y0.copy()
    st = np.zeros(ns, np.int32)
    api.integrate(net, 0.0, tend, 1, sc, y, ns, 1e-10, 8e-5, mode=api.MODE_FAST, controller=api.CTRL_ADAPTIVE,
                  max_attempts=100000, status_out=st)
    yo = y0.copy()
    O.gold_integrate(api.SIE, O.mass_action(N, rx, k), 0.0, tend, 1, sc, yo, ns, 1e-10, 8e-5, const_stride=1,
                     threads=ns, controller=api.CTRL_ADAPTIVE)
    assert (st == api.STATUS_OK).all()
    a, b = y.reshape(ns, N).astype(np.float64), yo.reshape(ns, N).astype(np.float64)
    floor = 1e-6 * np.abs(b).sum(axis=1, keepdims=True)
    rel = np.abs(a - b) / np.maximum(np.abs(b), floor)
    print(f"N={N} R=3600 fast: max rel diff vs float oracle {rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-3


@pytest.mark.parametrize("N,ns,cap,flags", DENSE_ATTEMPT_CASES)
def test_fast_dense_attempts(N, ns, cap, flags):
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
    t = traced_call(net, api.SIE, tend, sc, y0, ns, 1e-10, 8e-5, tr, ystart, flags)
    floor = 1e-6 * np.abs(ystart.astype(np.float64)).sum(axis=2)  # [ns, K]
    floor = np.broadcast_to(floor[:, :, None], ystart.shape)
    attempt_parity(api.SIE, O.mass_action(N, rx, k), sc, ns, tr, t, ystart, floor, f"dense N={N} flags {flags} fast per-attempt")


# -------------------------------------------------------------- whole runs
def whole_run_parity(method, onet, c, y0, n, N, tr, y, floor, label, t):
    """The kernel's run through the oracle's step sequence against the float
    oracle's run and against the same sequence replayed in float64 on the
    same time grid: the kernel is as close to the float64 replay as the
    float oracle's own run is (max and 99th percentile of the per-species
    relative difference within 1.5x the oracle's, plus 1e-7), and within
    5e-5 of the float oracle."""
    yo = tr["y"].reshape(n, N).astype(np.float64)
    yg = y.reshape(n, N).astype(np.float64)
    ex = O.replay_batch(method, onet, 0.0, c, y0, n, tr, controller=api.CTRL_ADAPTIVE, threads=THREADS)
    fl = np.broadcast_to(floor, yo.shape)
    dev = np.abs(yg - yo) / np.maximum(np.abs(yo), fl)
    e_g = np.abs(yg - ex) / np.maximum(np.abs(ex), fl)
    e_o = np.abs(yo - ex) / np.maximum(np.abs(ex), fl)
    q = lambda e: (float(e.max()), float(np.quantile(e, 0.99)), float(np.median(e)))
    print(f"{label}: {n} systems, attempts mean {tr['count'].mean():.0f} max {tr['count'].max()}; vs float oracle "
          f"max/p99/median {q(dev)[0]:.2e}/{q(dev)[1]:.2e}/{q(dev)[2]:.2e}; vs float64 replay: kernel "
          f"{q(e_g)[0]:.2e}/{q(e_g)[1]:.2e}/{q(e_g)[2]:.2e}, float oracle {q(e_o)[0]:.2e}/{q(e_o)[1]:.2e}/"
          f"{q(e_o)[2]:.2e}; own-test decisions differing {decision_mismatch(t, tr):.2e}", flush=True)
    assert np.all(np.isfinite(yg))
    assert q(e_g)[0] <= 1.5 * q(e_o)[0] + 1e-7
    assert q(e_g)[1] <= 1.5 * q(e_o)[1] + 1e-7
    assert dev.max() <= 5e-5, np.unravel_index(np.argmax(dev), dev.shape)


def test_fast_lane_replay_hhe6_sie():
    """The benchmarked fast lane kernel (SIE, tol 2e-6) carries its own state
    through the oracle's accepted-step sequence over 1 Myr."""
    p, y0, c, tend = hhe6_sample()
    n, tol = len(c), 2e-6
    tr = oracle_trace(api.SIE, O.hhe6(p), tend, 1, c, y0, n, tol, tol, api.CTRL_ADAPTIVE, 1)
    net = api.Network.builtin(api.NET_HHE6, p)
    y, t = replay(net, api.SIE, api.MODE_FAST, tend, 1, c, y0, n, tol, tol, api.CTRL_ADAPTIVE, tr)
    whole_run_parity(api.SIE, O.hhe6(p), c, y0, n, 6, tr, y, tol, "hhe6 fast replay SIE", t)


DENSE_CASES = [(32, 96, 0, 0), (32, 96, 0, api.FLAG_DENSE_LDS), (64, 32, 240, 0), (100, 12, 240, 0), (157, 8, 150, 0),
               (157, 8, 150, api.FLAG_DENSE_LDS)]


@pytest.mark.parametrize("N,ns,cap,flags", DENSE_CASES)
def test_fast_dense_replay(N, ns, cap, flags):
    """The dense LDS/MFMA solver carries its own state through the oracle's
    accepted-step sequence: over the whole interval for N = 32 (the systems
    with at least 100 accepted steps) and over the first 240 / 150 attempts
    for N = 64, 100 / 157 (at least 100 accepted steps each); floor 1e-6 of
    the system's total abundance."""
    rx, k, y0, sc = dense_case(N, ns)
    tend = W.MASS_ACTION_TEND
    onet = O.mass_action(N, rx, k)
    tr = oracle_trace(api.SIE, onet, tend, 1, sc, y0, ns, 1e-10, 8e-5, api.CTRL_ADAPTIVE, 1, max_attempts=cap)
    accepted = tr["accept"].sum(axis=1)
    keep = accepted >= 100
    assert keep.sum() >= min(ns, 48)
    net = api.Network.mass_action(N, rx, k)
    y, t = replay(net, api.SIE, api.MODE_FAST, tend, 1, sc, y0, ns, 1e-10, 8e-5, api.CTRL_ADAPTIVE, tr,
                  flags=flags)
    sub = {key: tr[key][keep] for key in ("h", "accept", "count", "flag")}
    sub.update(capacity=tr["capacity"], y=tr["y"].reshape(ns, N)[keep].reshape(-1))
    Yr = sub["y"].reshape(-1, N).astype(np.float64)
    floor = 1e-6 * np.abs(Yr).sum(axis=1, keepdims=True)
    nk = int(keep.sum())
    whole_run_parity(api.SIE, onet, sc.reshape(ns, -1)[keep].reshape(-1), y0.reshape(ns, N)[keep].reshape(-1), nk, N,
                     sub, y.reshape(ns, N)[keep], floor, f"dense N={N} flags {flags} fast replay ({nk} of {ns} systems)",
                     {"flag": t["flag"][keep]})


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
