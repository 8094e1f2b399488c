"""GPU parity of networks written by the code generator (plugin libraries).

Exact mode must reproduce the oracle bit for bit: for the two-body mixed
networks against the oracle's table-driven mass-action network (the
generator writes that path's operation order), for the general network
against its generated oracle statement.  Also: the per-network drop-in
library (sie.so), lockstep SIE, fast-mode accuracy against FP64 truth, and
generated networks inside a mixed batch.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
from wind_amd import api
from wind_amd import networks as W

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "wind_amd", "lib", "gen")


def same(a, b):
    a = np.asarray(a, np.float32).reshape(-1)
    b = np.asarray(b, np.float32).reshape(-1)
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


def mixed_inputs(N, nsys, seed):
    y0, sc = W.mass_action_systems(nsys, N, seed=seed, scale_range=W.MIXED_SCALE_RANGE)
    return y0, sc


@pytest.mark.parametrize("N", list(W.GENERATED_MIXED))
@pytest.mark.parametrize("method", [api.SIE, api.RK2])
@pytest.mark.parametrize("kernel", [api.KERNEL_LANE, api.KERNEL_BLOCK])
@pytest.mark.parametrize("controller", [api.CTRL_REFERENCE, api.CTRL_ADAPTIVE])
def test_generated_exact_matches_table_oracle(N, method, kernel, controller):
    ns = 96
    y0, sc = mixed_inputs(N, ns, seed=31 + N)
    tol = 1e-4 if controller == api.CTRL_ADAPTIVE else 1e-3
    tend = 1e7
    rx, k = W.mixed_network(N)
    yr = y0.copy()
    tr, sr, str_ = O.gold_integrate(method, O.mass_action(N, rx, k), 0.0, tend, 1, sc, yr, ns, tol, tol,
                                    const_stride=1, max_attempts=20000, controller=controller)
    y = y0.copy()
    steps = np.zeros(ns, np.int32)
    status = np.zeros(ns, np.int32)
    net = api.Network.plugin(f"mixed{N}")
    assert net.N == N and net.nconst == 1
    tot, _ = api.integrate(net, 0.0, tend, 1, sc, y, ns, tol, tol, method=method, mode=api.MODE_EXACT, kernel=kernel,
                           steps_out=steps, status_out=status, max_attempts=20000, controller=controller)
    print(f"N {N}: substeps gpu {tot} oracle {tr}")
    assert np.array_equal(status, str_)
    assert np.array_equal(steps, sr)
    assert same(y, yr)


def test_general_network_exact_matches_generated_oracle():
    ns = 128
    rng = np.random.default_rng(9)
    y0 = rng.uniform(0.1, 1.0, (ns, 4)).astype(np.float32).reshape(-1)
    c = rng.uniform(0.5, 2.0, (ns, 2)).astype(np.float32).reshape(-1)
    for method in (api.SIE, api.RK2):
        yr = y0.copy()
        tr, sr, str_ = O.gold_integrate(method, O.generated("demo_decay"), 0.0, 50.0, 10, c, yr, ns, 1e-4, 1e-4,
                                        const_stride=2, max_attempts=20000)
        y = y0.copy()
        steps = np.zeros(ns, np.int32)
        status = np.zeros(ns, np.int32)
        api.integrate(api.Network.plugin("demo_decay"), 0.0, 50.0, 10, c, y, ns, 1e-4, 1e-4, method=method,
                      mode=api.MODE_EXACT, steps_out=steps, status_out=status, max_attempts=20000)
        assert np.array_equal(status, str_) and np.array_equal(steps, sr), method
        assert same(y, yr), method


def test_generated_legacy_library():
    """The generated directory's sie.so, called exactly as WIND's harness does."""
    N, ns = 10, 64
    y0, sc = mixed_inputs(N, ns, seed=77)
    rx, k = W.mixed_network(N)
    yr = y0.copy()
    tr, _, _ = O.gold_integrate(O.SIE, O.mass_action(N, rx, k), 0.0, 1e6, 1, sc, yr, ns, 1e-3, 1e-3,
                                const_stride=1, max_attempts=20000)
    lib = C.CDLL(os.path.join(GEN, f"mixed{N}", "sie.so"))
    fn = getattr(lib, "_Z19cudaIntegrateSystemffiPfS_iiff")
    fn.restype = C.c_int
    y = y0.copy()
    c = sc.copy()
    n = fn(C.c_float(0.0), C.c_float(1e6), C.c_int(1), c.ctypes.data_as(C.POINTER(C.c_float)),
           y.ctypes.data_as(C.POINTER(C.c_float)), C.c_int(ns), C.c_int(N), C.c_float(1e-3), C.c_float(1e-3))
    assert n == tr
    assert same(y, yr)


def test_generated_lockstep_matches_oracle():
    N, ns = 5, 256
    y0, sc = mixed_inputs(N, ns, seed=3)
    rx, k = W.mixed_network(N)
    yr = y0.copy()
    r_or = O.lockstep(O.mass_action(N, rx, k), 0.0, 1e5, 10, sc, yr, ns, 1e-3, 1e-3, 1)
    y = y0.copy()
    r_gpu = api.integrate_lockstep(api.Network.plugin(f"mixed{N}"), 0.0, 1e5, 10, sc, y, ns, 1e-3, 1e-3)
    assert r_gpu == r_or
    assert same(y, yr)


@pytest.mark.parametrize("N", list(W.GENERATED_MIXED))
def test_generated_fast_accuracy_vs_fp64_truth(N):
    T = np.load(os.path.join(ROOT, "tests", "golden", "truth_mass_action.npz"))
    b = [b for b in W.mixed_workload(int(T["mixed_nsys"]), seed=int(T["mixed_seed"])) if b["N"] == N][0]
    sel = T[f"mixed_{N}_idx"]
    y = b["y0"].reshape(-1, N)[sel].reshape(-1).copy()
    c = b["scale"][sel].copy()
    st = np.zeros(len(sel), np.int32)
    api.integrate(api.Network.plugin(f"mixed{N}"), 0.0, float(T["tend"]), 1, c, y, len(sel), 1e-10, 1e-4,
                  mode=api.MODE_FAST, controller=api.CTRL_ADAPTIVE, status_out=st, max_attempts=100000)
    yt = T[f"mixed_{N}_y"]
    tot = np.abs(yt).sum(axis=1, keepdims=True)
    e = np.abs(y.reshape(-1, N) - yt) / np.maximum(np.abs(yt), 1e-6 * tot)
    e = np.where(np.isnan(e), np.inf, e).max(axis=1)
    print(f"N {N}: frac<=1e-3 {np.mean(e <= 1e-3):.3f} median {np.median(e):.2e} max {e.max():.2e}")
    assert (st == 0).all()
    assert np.mean(e <= 1e-3) >= 0.9


def test_generated_networks_in_a_batch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    dev = torch.device("cuda:0")
    bk = []
    for name, N, seed in (("mixed5", 5, 1), ("mixed10", 10, 2)):
        y0, sc = mixed_inputs(N, 3000, seed)
        for net in (api.Network.plugin(name), api.Network.mass_action(N, *W.mixed_network(N))):
            t0 = torch.from_numpy(y0).to(dev)
            bk.append(dict(net=net, nsys=3000, y0=t0, equations=t0.clone(), constants=torch.from_numpy(sc).to(dev)))
    alone = []
    for b in bk:
        y = b["y0"].clone()
        api.integrate(b["net"], 0.0, 1e9, 1, b["constants"], y, b["nsys"], 1e-10, 1e-4, mode=api.MODE_FAST,
                      controller=api.CTRL_ADAPTIVE, device_pointers=True, max_attempts=100000)
        alone.append(y.cpu().numpy())
    api.integrate_batch(bk, 0.0, 1e9, 1, 1e-10, 1e-4, max_attempts=100000)
    torch.cuda.synchronize()
    for b, ya in zip(bk, alone):
        assert np.array_equal(b["equations"].cpu().numpy(), ya), b["net"].name


def test_tiled_network_reproduces_untiled_per_tile():
    """Identical tiles make identical error tests, so the tiled system takes
    the untiled system's steps: each tile's final state equals the untiled
    result bit for bit, with the same sub-step counts; and the tiled network
    equals its generated oracle statement."""
    ns = 96
    rng = np.random.default_rng(4)
    y1 = rng.uniform(0.1, 1.0, (ns, 4)).astype(np.float32)
    c = rng.uniform(0.5, 2.0, (ns, 2)).astype(np.float32).reshape(-1)
    y3 = np.tile(y1, (1, 3)).reshape(-1).copy()
    s1 = np.zeros(ns, np.int32)
    s3 = np.zeros(ns, np.int32)
    ya = y1.reshape(-1).copy()
    api.integrate(api.Network.plugin("demo_decay"), 0.0, 50.0, 10, c, ya, ns, 1e-4, 1e-4, mode=api.MODE_EXACT,
                  steps_out=s1, max_attempts=20000)
    yb = y3.copy()
    api.integrate(api.Network.plugin("demo_decay_t3"), 0.0, 50.0, 10, c, yb, ns, 1e-4, 1e-4, mode=api.MODE_EXACT,
                  steps_out=s3, max_attempts=20000)
    assert np.array_equal(s1, s3)
    assert same(yb, np.tile(ya.reshape(ns, 4), (1, 3)))
    yr = y3.copy()
    O.gold_integrate(api.SIE, O.generated("demo_decay_t3"), 0.0, 50.0, 10, c, yr, ns, 1e-4, 1e-4, const_stride=2,
                     max_attempts=20000)
    assert same(yb, yr)


# ------------------------------------------------- networks from expressions
G = np.load(os.path.join(ROOT, "tests", "golden", "c_baseline_kats.npz"))


@pytest.mark.parametrize("method", [api.SIE, api.RK2])
@pytest.mark.parametrize("controller", [api.CTRL_REFERENCE, api.CTRL_ADAPTIVE])
@pytest.mark.parametrize("kernel", [api.KERNEL_LANE, api.KERNEL_BLOCK])
def test_generated_katz96_equals_builtin(method, controller, kernel):
    """Katz96 generated from its expressions integrates to the compiled-in
    network's bits (exact mode), which the oracle and the reference's
    c_baseline pin."""
    ys, cs = G["sys_y"], G["sys_c"]
    ns = len(ys)
    y0 = ys.reshape(-1).astype(np.float32).copy()
    c = cs.reshape(-1).astype(np.float32).copy()
    out = []
    for net in (api.Network.plugin("katz96_gen"), api.Network.builtin(api.NET_KATZ96)):
        y = y0.copy()
        steps = np.zeros(ns, np.int32)
        api.integrate(net, 0.0, 50.0, 1, c, y, ns, 1e-4, 1e-4, method=method, mode=api.MODE_EXACT, kernel=kernel,
                      steps_out=steps, controller=controller, max_attempts=100000)
        out.append((y, steps))
    assert same(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("method", [api.SIE, api.RK2])
def test_generated_nr_test_equals_builtin(method):
    y0, c, _ = W.nr_test_systems(1024, seed=9)
    out = []
    for net in (api.Network.plugin("nr_test_gen"), api.Network.builtin(api.NET_NR_TEST)):
        y = y0.copy()
        steps = np.zeros(1024, np.int32)
        api.integrate(net, 0.0, 1.0, 1000, c, y, 1024, 1e-3, 1e-3, method=method, steps_out=steps)
        out.append((y, steps))
    assert same(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("name,T,sel", [("katz96_t4", 4, slice(0, 12)), (f"katz96_t{W.KATZ96_NTILE}",
                                                                        W.KATZ96_NTILE, slice(0, 2))])
@pytest.mark.parametrize("method", [api.SIE, api.RK2])
def test_tiled_katz96_exact_matches_oracle(name, T, sel, method):
    """Katz96 tiled as in the reference's Ntile scaling study (N = 5 T,
    T = 4 and 50): the one-workgroup-per-system kernel (matrices in LDS at
    N = 20, in the device scratch at N = 250) reproduces the oracle's
    c_baseline integration of the generated network bit for bit, and every
    tile ends equal to the untiled network's result."""
    ys, cs = G["sys_y"][sel], G["sys_c"][sel]
    ns = len(ys)
    y0 = np.concatenate([np.tile(y.astype(np.float32), T) for y in ys])
    c = cs.reshape(-1).astype(np.float32).copy()
    onet = O.generated(name)
    yr = y0.copy()
    tr, sr, str_ = O.gold_integrate(method, onet, 0.0, 50.0, 1, c, yr, ns, 5e-3, 5e-3, const_stride=10,
                                    threads=ns) if method == api.SIE else \
        O.gold_integrate(method, onet, 0.0, 50.0, 1, c, yr, ns, 5e-3, 5e-3, const_stride=10)
    net = api.Network.plugin(name)
    assert net.N == 5 * T
    y = y0.copy()
    steps = np.zeros(ns, np.int32)
    tot, _ = api.integrate(net, 0.0, 50.0, 1, c, y, ns, 5e-3, 5e-3, method=method, mode=api.MODE_EXACT,
                           steps_out=steps)
    print(f"{name} m{method}: substeps gpu {tot} oracle {tr}", flush=True)
    assert np.array_equal(steps, sr)
    assert same(y, yr)
    # every tile equals the untiled Katz96 (tiles are independent copies)
    y1 = ys.reshape(-1).astype(np.float32).copy()
    s1 = np.zeros(ns, np.int32)
    api.integrate(api.Network.builtin(api.NET_KATZ96), 0.0, 50.0, 1, c, y1, ns, 5e-3, 5e-3, method=method,
                  mode=api.MODE_EXACT, steps_out=s1)
    # values (a zero's sign may differ: the zero blocks add signed zeros)
    assert np.array_equal(y.reshape(ns, T, 5), np.broadcast_to(y1.reshape(ns, 1, 5), (ns, T, 5)))


def test_tiled_katz96_fast_mode():
    """Fast mode at N = 250 (LU in the device scratch): the tiles stay
    identical to each other and end within 1e-3 of the exact-mode result."""
    T = W.KATZ96_NTILE
    ys, cs = G["sys_y"][:16], G["sys_c"][:16]
    ns = len(ys)
    y0 = np.concatenate([np.tile(y.astype(np.float32), T) for y in ys])
    c = cs.reshape(-1).astype(np.float32).copy()
    net = api.Network.plugin(f"katz96_t{T}")
    yf, ye = y0.copy(), y0.copy()
    api.integrate(net, 0.0, 50.0, 1, c, yf, ns, 1e-4, 1e-4, mode=api.MODE_FAST, controller=api.CTRL_ADAPTIVE)
    api.integrate(net, 0.0, 50.0, 1, c, ye, ns, 1e-4, 1e-4, mode=api.MODE_EXACT, controller=api.CTRL_ADAPTIVE)
    F = yf.reshape(ns, T, 5)
    assert np.array_equal(F, np.broadcast_to(F[:, :1], F.shape))
    E = ye.reshape(ns, T, 5).astype(np.float64)
    err = np.abs(F - E) / np.maximum(np.abs(E), 1e-3)
    assert err.max() <= 1e-3, err.max()


def test_tiled_katz96_drop_in_symbol():
    """The tiled network's drop-in sie.so (its NUM_CONST = 10) gives the
    oracle's bits through the reference's entry point."""
    T = 4
    ys, cs = G["sys_y"][:8], G["sys_c"][:8]
    ns = len(ys)
    y0 = np.concatenate([np.tile(y.astype(np.float32), T) for y in ys])
    c = cs.reshape(-1).astype(np.float32).copy()
    so = C.CDLL(os.path.join(GEN, "katz96_t4", "sie.so"))
    fn = getattr(so, "_Z19cudaIntegrateSystemffiPfS_iiff")
    fn.argtypes = [C.c_float, C.c_float, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float]
    y = y0.copy()
    k = fn(0.0, 50.0, 1, c.ctypes.data, y.ctypes.data, ns, 5 * T, 5e-3, 5e-3)
    yr = y0.copy()
    kr, _, _ = O.gold_integrate(O.SIE, O.generated("katz96_t4"), 0.0, 50.0, 1, c, yr, ns, 5e-3, 5e-3,
                                const_stride=10)
    assert k == kr and same(y, yr)
