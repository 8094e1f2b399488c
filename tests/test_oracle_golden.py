"""The oracle against known-answer vectors produced by the reference's own
c_baseline (tests/golden/c_baseline_kats.npz, written by make_golden.py from
oracle/_ref).  Bit-exact comparisons throughout."""
import hashlib
import os

import numpy as np
import pytest

import oracle_lib as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "c_baseline_kats.npz"))


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("N", [2, 5, 6, 10, 32, 64, 157])
def test_gje_invert(N):
    L = O.lib()
    keys = [k[:-3] for k in G.files if k.startswith(f"gje_N{N}_") and k.endswith("_in")]
    assert keys
    for key in keys:
        M = G[key + "_in"].copy()
        Inv = np.zeros_like(M)
        L.oracle_gje_invert(M.ctypes.data, Inv.ctypes.data, N)
        if key + "_inv" in G.files:
            assert same_bits(Inv, G[key + "_inv"]), key
            assert same_bits(M, G[key + "_mout"]), key
        else:
            assert hashlib.sha256(Inv.tobytes()).hexdigest() == str(G[key + "_inv_sha"]), key
            assert hashlib.sha256(M.tobytes()).hexdigest() == str(G[key + "_mout_sha"]), key


@pytest.mark.parametrize("N", [2, 5, 6, 32])
def test_matvec(N):
    v = G[f"matvec_N{N}_v"].copy()
    M = G[f"matvec_N{N}_M"]
    O.lib().oracle_matvec(M.ctypes.data, v.ctypes.data, N)
    assert same_bits(v, G[f"matvec_N{N}_out"])


@pytest.mark.parametrize("method,name", [(O.SIE, "sie"), (O.RK2, "rk2")])
def test_check_error_table(method, name):
    y1, y2 = G["check_y1"], G["check_y2"]
    ABS, REL = (float(x) for x in G["check_abs_rel"])
    flags = []
    for i in range(len(y1)):
        a, b = y1[i:i + 1], y2[i:i + 1]
        flags.append(O.lib().oracle_check_error(method, a.ctypes.data, b.ctypes.data, 1, ABS, REL))
    assert np.array_equal(np.array(flags), G[f"check_{name}_flags"])
    assert O.lib().oracle_check_error(method, y1.ctypes.data, y2.ctypes.data, len(y1), ABS, REL) == int(
        G[f"check_{name}_all"])
    # the table exercises both outcomes
    assert 0 < G[f"check_{name}_flags"].sum() < len(y1)


def test_katz96_network():
    net = O.katz96()
    for y, c, f, J in zip(G["katz_y"], G["katz_c"], G["katz_f"], G["katz_J"]):
        fo = np.zeros(5, np.float32)
        Jo = np.zeros(25, np.float32)
        O.lib().oracle_rhs(net.ref, 0.0, y.ctypes.data, c.ctypes.data, fo.ctypes.data)
        O.lib().oracle_jac(net.ref, y.ctypes.data, c.ctypes.data, Jo.ctypes.data)
        assert same_bits(fo, f)
        assert same_bits(Jo, J)


@pytest.mark.parametrize("method,name", [(O.SIE, "sie"), (O.RK2, "rk2")])
@pytest.mark.parametrize("nsub", [1, 2])
def test_take_step(method, name, nsub):
    net = O.katz96()
    out = G["take_y"].copy()
    cs = G["take_c"]
    hs = G["take_h"]
    for i in range(len(out)):
        row, c = out[i], cs[i]
        O.lib().oracle_take_step(method, net.ref, 0.0, float(hs[i]), nsub, row.ctypes.data, c.ctypes.data, 5)
    assert same_bits(out, G[f"take_{name}_n{nsub}"])


@pytest.mark.parametrize("tend", [10, 200])
@pytest.mark.parametrize("method,name", [(O.SIE, "sie"), (O.RK2, "rk2")])
def test_fixture_integration(tend, method, name):
    """The reference's smoke test (cuda/debug/debug.c with inputKatz96_1_1_1.h), NUM_CONST = 2."""
    from golden.make_golden import KATZ_FIXTURE_C, KATZ_FIXTURE_EQ

    eq = KATZ_FIXTURE_EQ.copy()
    n, _, st = O.gold_integrate(method, O.katz96(), 0.0, float(tend), 1, KATZ_FIXTURE_C, eq, 4, 5e-3, 5e-3,
                                const_stride=2)
    assert same_bits(eq, G[f"fixture_t{tend}_{name}_eq"])
    assert n == int(G[f"fixture_t{tend}_{name}_n"])
    assert (st == 0).all()


@pytest.mark.parametrize("tol", [5e-3, 1e-4, 1e-6])
@pytest.mark.parametrize("method,name", [(O.SIE, "sie"), (O.RK2, "rk2")])
def test_per_system_integration(tol, method, name):
    ys, cs = G["sys_y"], G["sys_c"]
    eq = ys.reshape(-1).copy()
    _, steps, st = O.gold_integrate(method, O.katz96(), 0.0, 50.0, 1, cs.reshape(-1).copy(), eq, len(ys), tol, tol,
                                    const_stride=10)
    assert same_bits(eq.reshape(ys.shape), G[f"sys_tol{tol:g}_{name}_eq"])
    assert np.array_equal(steps, G[f"sys_tol{tol:g}_{name}_steps"])
    assert (st == 0).all()
