"""The oracle's lockstep SIE (the semantics of the reference's
cudaIntegrateSIE / errorLoop, cuda/SIE_host/harness_kernel.cu:248-425, with
its error kernel cuda/SIE_host/vector_kernels.cu:45-70) pinned by what
those semantics fix, with the sub-steps taken by the reference's own
c_baseline take_step compiled from source (oracle/_ref/sie_gold.so):

* tolerances that never fail: every attempt is accepted and the step doubles
  after each, so the state is the chain y <- take_step(t, t + dt, 2, y),
  dt <- 2 dt (clamped to the interval's end), and the return value is
  3 * attempts * Nsystems;
* ABS = REL = 0: every attempt fails, the loop gives up after 9 consecutive
  failures, the state is unchanged and the return value is 27 * Nsystems;
* in between, a separate statement of the control loop (written here in
  numpy from the reference text: strict '>' tests, the relative test gated
  on y1 > ABS and y2 > ABS without absolute values, any failing system
  halving the shared step) driving the reference's take_step gives the
  oracle's states and return value bit for bit.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "c_baseline_kats.npz"))
N = 5


def systems(n):
    return G["sys_y"][:n].reshape(-1).astype(np.float32).copy(), G["sys_c"][:n].reshape(-1).astype(np.float32).copy()


def ref_take_step2(R, tnow, tb, y, c):
    f = np.zeros(N, np.float32)
    J = np.zeros(N * N, np.float32)
    Inv = np.zeros(N * N, np.float32)
    R.take_step(C.c_float(tnow), C.c_float(tb), 2, y.ctypes.data, c.ctypes.data, f.ctypes.data, J.ctypes.data,
                Inv.ctypes.data, N)


def ref_take_step(R, tnow, tb, nsub, y, c):
    f = np.zeros(N, np.float32)
    J = np.zeros(N * N, np.float32)
    Inv = np.zeros(N * N, np.float32)
    R.take_step(C.c_float(tnow), C.c_float(tb), nsub, y.ctypes.data, c.ctypes.data, f.ctypes.data, J.ctypes.data,
                Inv.ctypes.data, N)


def lockstep_statement(R, tnow, tend, nfix, Y, Cs, ns, ABS, REL):
    """errorLoop (harness_kernel.cu:248-425) restated with numpy float32
    control and the reference's take_step; returns (state, nsteps * ns)."""
    f32 = np.float32
    tnow, tend = f32(tnow), f32(tend)
    ABS, REL = f32(ABS), f32(REL)
    Y = Y.copy()
    unsolved = 0
    nsteps = 0
    timestep = f32(f32(tend - tnow) / f32(nfix))
    while tnow < tend and unsolved < 9:
        timestep = f32(min(float(timestep), float(f32(tend - tnow))))
        nsteps += 3
        tb = f32(tnow + timestep)
        Y1 = Y.copy()
        Y2 = Y.copy()
        for s in range(ns):
            c = np.ascontiguousarray(Cs[10 * s:10 * s + 10])
            y1 = np.ascontiguousarray(Y1[N * s:N * s + N])
            y2 = np.ascontiguousarray(Y2[N * s:N * s + N])
            ref_take_step(R, tnow, tb, 1, y1, c)
            ref_take_step(R, tnow, tb, 2, y2, c)
            Y1[N * s:N * s + N] = y1
            Y2[N * s:N * s + N] = y2
        d = (Y1 - Y2).astype(np.float32)
        flag = np.any(np.abs(d) > ABS)
        rel = np.abs(d.astype(np.float64) / (Y2.astype(np.float64) + 1e-12))
        flag = flag or np.any((rel > np.float64(REL)) & (Y1 > ABS) & (Y2 > ABS))
        if flag:
            unsolved += 1
            timestep = f32(timestep / f32(2))
        else:
            unsolved = 0
            Y = Y2
            tnow = tb
            timestep = f32(timestep * f32(2))
    return Y, nsteps * ns


@pytest.fixture(scope="module")
def R():
    if not O.ref_available():
        pytest.skip("reference c_baseline not built (oracle/_ref)")
    return O.ref(O.SIE)


def test_always_accept_is_doubling_chain(R):
    ns = 8
    Y0, Cs = systems(ns)
    tend, nfix = 50.0, 64
    Y = Y0.copy()
    ret = O.lockstep(O.katz96(), 0.0, tend, nfix, Cs, Y, ns, 1e30, 1e30, 10)
    # the chain, system by system
    attempts = 0
    f32 = np.float32
    want = Y0.copy()
    for s in range(ns):
        y = np.ascontiguousarray(Y0[N * s:N * s + N])
        c = np.ascontiguousarray(Cs[10 * s:10 * s + 10])
        t, dt, k = f32(0.0), f32(f32(tend) / f32(nfix)), 0
        while t < f32(tend):
            dt = f32(min(float(dt), float(f32(f32(tend) - t))))
            tb = f32(t + dt)
            ref_take_step2(R, t, tb, y, c)
            t = tb
            dt = f32(dt * f32(2))
            k += 1
        want[N * s:N * s + N] = y
        attempts = k
    assert np.array_equal(Y.view(np.uint32), want.view(np.uint32))
    assert ret == 3 * attempts * ns
    assert attempts == 7  # 64 steps of 50/64 doubling: 1+2+4+8+16+32 = 63, then the clamped last one


def test_zero_tolerance_gives_up_after_nine_failures(R):
    ns = 6
    Y0, Cs = systems(ns)
    Y = Y0.copy()
    ret = O.lockstep(O.katz96(), 0.0, 50.0, 1, Cs, Y, ns, 0.0, 0.0, 10)
    assert ret == 27 * ns
    assert np.array_equal(Y.view(np.uint32), Y0.view(np.uint32))


@pytest.mark.parametrize("tol", [5e-3, 1e-4])
def test_lockstep_matches_statement(R, tol):
    ns = 6
    Y0, Cs = systems(ns)
    Y = Y0.copy()
    ret = O.lockstep(O.katz96(), 0.0, 50.0, 1, Cs, Y, ns, tol, tol, 10)
    want, wret = lockstep_statement(R, 0.0, 50.0, 1, Y0, Cs, ns, tol, tol)
    assert ret == wret
    assert np.array_equal(Y.view(np.uint32), want.view(np.uint32))
