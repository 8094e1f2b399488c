"""wind_integrate_noalloc (the allocation-free entry point) against
wind_integrate on the same device inputs: bit-identical final states,
per-system sub-step counts, status and sub-step totals.  Each case makes two
calls in a row, the second with fewer systems: the scratch array keeps the
first call's counters and start order, so the second call shows that the
entry point clears what must start at zero."""
import ctypes as C

import numpy as np
import pytest

from wind_amd import api
from wind_amd import networks as W

pytestmark = pytest.mark.gpu


def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def call(fn, net, c_dev, y_dev, nsys, tend, nfix, tol, method, mode, controller, hint=None, stride=0):
    """One integration through the C-ABI entry point `fn` on device arrays;
    returns (total sub-steps, steps[nsys], status[nsys])."""
    torch = torch_cuda()
    lib = api.load()
    steps = torch.zeros(nsys, dtype=torch.int32, device="cuda")
    status = torch.full((nsys,), -1, dtype=torch.int32, device="cuda")
    o = api.WindOptions()
    lib.wind_options_init(C.byref(o))
    o.method, o.mode, o.controller, o.const_stride = method, mode, controller, stride
    o.device_pointers = 1
    total = C.c_int64(-1)
    o.total_steps = C.cast(C.byref(total), C.c_void_p)
    o.steps_out = steps.data_ptr()
    o.status_out = status.data_ptr()
    o.cost_hint = None if hint is None else hint.data_ptr()
    torch.cuda.synchronize()
    rc = fn(net.handle, 0.0, tend, nfix, c_dev.data_ptr(), y_dev.data_ptr(), nsys, net.N, tol, tol, C.byref(o))
    assert rc == 0, lib.wind_last_error().decode()
    torch.cuda.synchronize()
    return total.value, steps.cpu().numpy(), status.cpu().numpy()


def compare(net, y0, c, sizes, tend, nfix, tol, method, mode, controller, hinted=False, stride=0):
    torch = torch_cuda()
    lib = api.load()
    for n in sizes:
        print(f"noalloc {net.name} n={n} method={method} mode={mode} hinted={hinted}", flush=True)
        yy = np.ascontiguousarray(y0[: n * net.N])
        cc = np.ascontiguousarray(c[: (n - 1) * (stride or net.nconst) + net.nconst])
        c_dev = torch.from_numpy(cc).cuda()
        y_ref = torch.from_numpy(yy).cuda()
        y_new = y_ref.clone()
        hint = None
        if hinted:
            hint = torch.from_numpy(np.random.default_rng(n).integers(1, 5000, n).astype(np.int32)).cuda()
        ref = call(lib.wind_integrate, net, c_dev, y_ref, n, tend, nfix, tol, method, mode, controller, hint, stride)
        new = call(lib.wind_integrate_noalloc, net, c_dev, y_new, n, tend, nfix, tol, method, mode, controller, hint,
                   stride)
        assert new[0] == ref[0] and ref[0] > 0
        assert np.array_equal(new[1], ref[1])
        assert np.array_equal(new[2], ref[2])
        a = y_new.cpu().numpy().view(np.uint32)
        b = y_ref.cpu().numpy().view(np.uint32)
        assert np.array_equal(a, b), np.where(a != b)[0][:5]


@pytest.mark.parametrize("method", [api.SIE, api.RK2])
def test_noalloc_hhe6_fast_start_order(method):
    """Fast lane kernel; the first call is large enough (2^17 + 1000
    systems) to start in the order of an initial-state work estimate, built
    in the scratch array; the second is small and starts in index order."""
    n1, n2 = (1 << 17) + 1000, 3000
    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(n1, seed=21)
    net = api.Network.builtin(api.NET_HHE6, p)
    tend = 1.0e4 if method == api.SIE else 1.0e2
    compare(net, y0.reshape(-1), nH, [n1, n2], tend, 1, 1e-4, method, api.MODE_FAST, api.CTRL_ADAPTIVE)


def test_noalloc_cost_hint_order():
    """Caller cost hints (device array): the start order is sorted into the
    scratch array; results equal wind_integrate's with the same hints."""
    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(20000, seed=22)
    net = api.Network.builtin(api.NET_HHE6, p)
    compare(net, y0.reshape(-1), nH, [20000, 700], 1.0e4, 1, 1e-4, api.SIE, api.MODE_FAST, api.CTRL_ADAPTIVE,
            hinted=True)


def test_noalloc_exact_reference_controller():
    """Exact mode, the reference controller (NR_test workload shape)."""
    y0, c, _ = W.nr_test_systems(1024, seed=23)
    net = api.Network.builtin(api.NET_NR_TEST)
    compare(net, y0.reshape(-1), c.reshape(-1), [1024, 100], 1.0, 1000, 1e-3, api.SIE, api.MODE_EXACT,
            api.CTRL_REFERENCE, stride=4)


@pytest.mark.parametrize("N", [10, 32, 64])
def test_noalloc_mass_action_dense(N):
    """Mass-action networks in fast mode (the dense LDS/MFMA solver; at
    N = 64 the register-resident one; work queue in the scratch array)."""
    rx, k = W.mass_action_network(N, seed=N)
    y0, sc = W.mass_action_systems(600, N, seed=40 + N, scale_range=(0.5, 2.0))
    net = api.Network.mass_action(N, rx, k)
    compare(net, y0.reshape(-1), sc, [600, 90], 1.0e7, 1, 1e-4, api.SIE, api.MODE_FAST, api.CTRL_ADAPTIVE)


def test_noalloc_plugin_network():
    """A generated network (plugin lane kernel)."""
    net = api.Network.plugin("mixed10")
    rx, k = W.mass_action_network(10, seed=10)
    y0, sc = W.mass_action_systems(4000, 10, seed=51, scale_range=(0.5, 2.0))
    compare(net, y0.reshape(-1), sc, [4000, 500], 1.0e7, 1, 1e-4, api.SIE, api.MODE_FAST, api.CTRL_ADAPTIVE)
