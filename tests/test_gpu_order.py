"""GPU: start order from per-system cost hints (wind_options.cost_hint),
through the C-ABI.

The lane kernels take systems from a queue; a cost hint only permutes that
queue (longest first), and each system's arithmetic does not depend on the
lane that runs it.  So final states, per-system step counts, status and
totals must be bit-identical with and without the hint, in both arithmetic
modes, and exact mode stays bit-identical to the oracle.
"""
import numpy as np
import pytest

import oracle_lib as O
from wind_amd import api
from wind_amd import networks as W

pytestmark = pytest.mark.gpu


def run(net, y0, c, nsys, tend, tol, mode, hint=None, device=False):
    steps = np.zeros(nsys, np.int32)
    status = np.zeros(nsys, np.int32)
    if device:
        import torch

        dev = torch.device("cuda:0")
        y = torch.from_numpy(y0).to(dev)
        cd = torch.from_numpy(c).to(dev)
        sd = torch.zeros(nsys, dtype=torch.int32, device=dev)
        std = torch.zeros(nsys, dtype=torch.int32, device=dev)
        hd = None if hint is None else torch.from_numpy(hint).to(dev)
        total, _ = api.integrate(net, 0.0, tend, 1, cd, y, nsys, tol, tol, method=api.SIE, mode=mode,
                                 kernel=api.KERNEL_LANE, device_pointers=True, steps_out=sd, status_out=std,
                                 controller=api.CTRL_ADAPTIVE, cost_hint=hd)
        torch.cuda.synchronize()
        return y.cpu().numpy(), sd.cpu().numpy(), std.cpu().numpy(), total
    y = y0.copy()
    total, _ = api.integrate(net, 0.0, tend, 1, c, y, nsys, tol, tol, method=api.SIE, mode=mode,
                             kernel=api.KERNEL_LANE, steps_out=steps, status_out=status,
                             controller=api.CTRL_ADAPTIVE, cost_hint=hint)
    return y, steps, status, total


def same(a, b, label):
    assert np.array_equal(a[2], b[2]), label
    assert np.array_equal(a[1], b[1]), label
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)), label
    assert a[3] == b[3], label


@pytest.mark.parametrize("device", [False, True])
def test_hhe6_fast_hinted_order_is_bit_identical(device):
    """2^18 hhe6 systems, fast mode, adaptive controller at tol 1e-5: the
    hint is the unhinted run's steps_out, and a random hint."""
    n = 1 << 18
    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(n, seed=7)
    net = api.Network.builtin(api.NET_HHE6, p)
    base = run(net, y0, nH, n, W.HHE6_TEND_YR, 1e-5, api.MODE_FAST, device=device)
    assert (base[2] == 0).all()
    same(run(net, y0, nH, n, W.HHE6_TEND_YR, 1e-5, api.MODE_FAST, hint=base[1].copy(), device=device), base,
         "steps_out hint")
    rnd = np.random.default_rng(8).integers(-5, 1 << 30, n, dtype=np.int64).astype(np.int32)
    same(run(net, y0, nH, n, W.HHE6_TEND_YR, 1e-5, api.MODE_FAST, hint=rnd, device=device), base, "random hint")


def test_exact_mode_hinted_matches_oracle():
    """Exact arithmetic with a hint still reproduces the oracle bit for bit."""
    n = 256
    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(n, seed=9)
    net = api.Network.builtin(api.NET_HHE6, p)
    hint = np.arange(n, dtype=np.int32)[::-1].copy()
    g = run(net, y0, nH, n, 10.0, 1e-3, api.MODE_EXACT, hint=hint)
    yr = y0.copy()
    total, steps, status = O.gold_integrate(api.SIE, O.hhe6(p), 0.0, 10.0, 1, nH, yr, n, 1e-3, 1e-3,
                                            const_stride=1, controller=api.CTRL_ADAPTIVE)
    assert np.array_equal(g[2], status)
    assert np.array_equal(g[1], steps)
    assert np.array_equal(g[0].view(np.uint32), yr.view(np.uint32))
    assert g[3] == total


def test_mass_action_dense_hinted_order_is_bit_identical():
    """The dense solver (mass action, N = 64, fast mode) takes systems from
    its queue in the hinted order; results are bit-identical."""
    b = [b for b in W.mixed_workload(1 << 14, seed=4) if b["N"] == 64][0]
    n = min(1024, len(b["scale"]))
    net = api.Network.mass_action(64, b["rx"], b["k"])
    out = []
    for hint in (None, "steps"):
        y = b["y0"].reshape(-1, 64)[:n].reshape(-1).copy()
        steps = np.zeros(n, np.int32)
        status = np.zeros(n, np.int32)
        h = None if hint is None else out[0][1][::-1].copy()
        total, _ = api.integrate(net, 0.0, 1e9, 1, b["scale"][:n].copy(), y, n, 1e-10, 8e-5, mode=api.MODE_FAST,
                                 controller=api.CTRL_ADAPTIVE, steps_out=steps, status_out=status,
                                 max_attempts=100000, cost_hint=h)
        out.append((y, steps, status, total))
    same(out[1], out[0], "dense N=64")


def test_legacy_symbol_repeated_calls_are_bit_identical():
    """WIND's output loop calls the drop-in symbol on one state array per
    output step; from the second call on, the shim starts systems longest
    first by the call before's sub-step counts (2^18 systems, above its
    threshold).  States and returned sub-step totals equal those of
    unhinted calls over the same intervals."""
    import ctypes as C
    import os

    n = 1 << 18
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    so = C.CDLL(os.path.join(root, "wind_amd", "lib", "hhe6", "sie.so"))
    fn = getattr(so, "_Z19cudaIntegrateSystemffiPfS_iiff")
    fn.argtypes = [C.c_float, C.c_float, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float]
    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(n, seed=10)
    net = api.Network.builtin(api.NET_HHE6, p)
    ys, yr = y0.copy(), y0.copy()
    for t0, t1 in ((0.0, 1e3), (1e3, 2e3), (2e3, 3e3)):
        got = fn(t0, t1, 1, nH.ctypes.data, ys.ctypes.data, n, 6, 1e-3, 1e-3)
        total, _ = api.integrate(net, t0, t1, 1, nH, yr, n, 1e-3, 1e-3, method=api.SIE, mode=api.MODE_EXACT)
        assert got == (total & 0xFFFFFFFF) - (1 << 32 if total & 0x80000000 else 0)
        assert np.array_equal(ys.view(np.uint32), yr.view(np.uint32)), (t0, t1)


def test_generated_network_hinted_order_is_bit_identical():
    """A generated network (plugin lane kernel, N = 5, fast mode) follows the
    hinted order with bit-identical results."""
    import os

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    plugin = os.path.join(root, "wind_amd", "lib", "gen", "mixed5", "libwind_net.so")
    b = [b for b in W.mixed_workload(1 << 14, seed=4) if b["N"] == 5][0]
    n = min(4096, len(b["scale"]))
    net = api.Network.plugin(plugin)
    out = []
    for k in range(2):
        y = b["y0"].reshape(-1, 5)[:n].reshape(-1).copy()
        steps = np.zeros(n, np.int32)
        status = np.zeros(n, np.int32)
        h = None if k == 0 else out[0][1].copy()
        total, _ = api.integrate(net, 0.0, 1e9, 1, b["scale"][:n].copy(), y, n, 1e-10, 8e-5, mode=api.MODE_FAST,
                                 controller=api.CTRL_ADAPTIVE, steps_out=steps, status_out=status,
                                 max_attempts=100000, cost_hint=h)
        out.append((y, steps, status, total))
    same(out[1], out[0], "generated N=5")
