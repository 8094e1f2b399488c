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
