"""Network code generator (wind_amd/codegen.py), without a GPU.

* The generated oracle statement of a two-body network computes the same
  bits as the oracle's table-driven mass-action network (float and double):
  the generator writes the table path's operation order.
* A general network (one- and two-body reactions, two constants) against an
  independent float64 evaluation of its description, and its Jacobian
  against finite differences of its RHS.
* Operation counts, the plugin descriptor, the per-network shims' exported
  names, and rejection of malformed descriptions.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from wind_amd import codegen
from wind_amd import networks as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "wind_amd", "lib", "gen")


def _eval(net, y, c, dtype):
    L = O.lib()
    N = net.N
    f = np.zeros(N, dtype)
    J = np.zeros(N * N, dtype)
    y = np.ascontiguousarray(y, dtype)
    c = np.ascontiguousarray(c, dtype)
    if dtype == np.float32:
        L.oracle_rhs(net.ref, C.c_float(0.0), y.ctypes.data, c.ctypes.data, f.ctypes.data)
        L.oracle_jac(net.ref, y.ctypes.data, c.ctypes.data, J.ctypes.data)
    else:
        L.oracle_rhs_f64(net.ref, C.c_double(0.0), y.ctypes.data, c.ctypes.data, f.ctypes.data)
        L.oracle_jac_f64(net.ref, y.ctypes.data, c.ctypes.data, J.ctypes.data)
    return f, J


@pytest.mark.parametrize("N", list(W.GENERATED_MIXED))
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_generated_equals_table(N, dtype):
    rx, k = W.mixed_network(N)
    gen = O.generated(f"mixed{N}")
    tab = O.mass_action(N, rx, k)
    rng = np.random.default_rng(N)
    for trial in range(20):
        y = np.exp(rng.uniform(np.log(1e-12), 0.0, N)).astype(np.float32)
        c = np.array([np.exp(rng.uniform(np.log(0.1), np.log(10.0)))], np.float32)
        fg, Jg = _eval(gen, y, c, dtype)
        ft, Jt = _eval(tab, y, c, dtype)
        assert np.array_equal(fg.view(np.uint8), ft.view(np.uint8)), trial
        assert np.array_equal(Jg.view(np.uint8), Jt.view(np.uint8)), trial


def _reference_eval(spec, y, c):
    """float64 RHS and Jacobian straight from the description."""
    N = spec["species"]
    f = np.zeros(N)
    J = np.zeros((N, N))  # J[i, j] = d f_i / d y_j
    for rx in spec["reactions"]:
        kc = float(np.float32(rx["k"])) * (c[rx["const"]] if rx["const"] is not None else 1.0)
        reac = rx["reactants"]
        w = kc * np.prod([y[s] for s in reac])
        dw = np.zeros(N)
        if len(reac) == 2:
            a, b = reac
            dw[a] += kc * y[b]
            dw[b] += kc * y[a]
        else:
            dw[reac[0]] += kc
        for s in reac:
            f[s] -= w
            J[s] -= dw
        for s in rx["products"]:
            f[s] += w
            J[s] += dw
    return f, J


def test_general_network_against_description():
    spec = [s for s in W.generated_specs() if s["name"] == "demo_decay"][0]
    gen = O.generated("demo_decay")
    rng = np.random.default_rng(5)
    for trial in range(20):
        y = rng.uniform(0.01, 2.0, 4)
        c = rng.uniform(0.5, 2.0, 2)
        f, Jc = _eval(gen, y, c, np.float64)
        fr, Jr = _reference_eval(spec, y, c)
        assert np.allclose(f, fr, rtol=1e-13, atol=1e-15)
        J = Jc.reshape(4, 4).T  # stored column-major
        assert np.allclose(J, Jr, rtol=1e-13, atol=1e-15)
        # finite differences of the generated RHS
        eps = 1e-6
        for j in range(4):
            e = np.zeros(4)
            e[j] = eps
            fp, _ = _eval(gen, y + e, c, np.float64)
            fm, _ = _eval(gen, y - e, c, np.float64)
            assert np.allclose((fp - fm) / (2 * eps), J[:, j], rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("name", ["mixed5", "mixed10", "demo_decay"])
def test_plugin_descriptor_and_counts(name):
    spec = [s for s in W.generated_specs() if s["name"] == name][0]
    g = codegen.generate(spec)
    meta = json.load(open(os.path.join(GEN, name, "network.json")))
    assert meta["flops_rhs"] == g["flops_rhs"] and meta["flops_jac"] == g["flops_jac"]
    if name.startswith("mixed"):
        R = len(spec["reactions"])
        assert g["flops_rhs"] == 7 * R and g["flops_jac"] == 11 * R

    class Desc(C.Structure):
        _fields_ = [("abi", C.c_int), ("N", C.c_int), ("nconst", C.c_int), ("nreactions", C.c_int),
                    ("name", C.c_char_p), ("flops_rhs", C.c_longlong), ("flops_jac", C.c_longlong),
                    ("launch", C.c_void_p), ("lockstep_attempt", C.c_void_p)]

    lib = C.CDLL(os.path.join(GEN, name, "libwind_net.so"))
    lib.wind_plugin_get.restype = C.POINTER(Desc)
    d = lib.wind_plugin_get().contents
    assert d.abi == 4 and d.N == g["N"] and d.nconst == g["nconst"] and d.nreactions == g["R"]
    assert d.name.decode() == name and d.flops_rhs == g["flops_rhs"] and d.flops_jac == g["flops_jac"]
    assert d.launch and d.lockstep_attempt


@pytest.mark.parametrize("name", ["mixed5", "demo_decay"])
def test_generated_shims_export_reference_symbols(name):
    import subprocess

    def syms(path):
        out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
        return {line.split()[-1] for line in out.splitlines() if line.strip()}

    d = os.path.join(GEN, name)
    assert "_Z19cudaIntegrateSystemffiPfS_iiff" in syms(os.path.join(d, "sie.so"))
    assert "_Z19cudaIntegrateSystemffiPfS_iiff" in syms(os.path.join(d, "rk2.so"))
    host = syms(os.path.join(d, "sie_host.so"))
    assert "_Z16cudaIntegrateSIEffiPfS_iiff" in host and "_Z26initializeCublasExternallyv" in host


@pytest.mark.parametrize("bad,msg", [
    ({"name": "x-y", "species": 2, "reactions": []}, "identifier"),
    ({"name": "x", "species": 2, "reactions": [{"reactants": [0, 1, 1], "products": [], "k": 1.0}]}, "reactants"),
    ({"name": "x", "species": 2, "reactions": [{"reactants": [2], "products": [], "k": 1.0}]}, "out of range"),
    ({"name": "x", "species": 2, "nconst": 1, "reactions": [{"reactants": [0], "products": [1], "k": 1.0,
                                                             "const": 1}]}, "constant"),
    ({"name": "x", "species": 2, "reactions": [{"reactants": [0], "products": [1], "k": float("inf")}]},
     "non-finite"),
])
def test_rejects_malformed(bad, msg):
    with pytest.raises(ValueError, match=msg):
        codegen.generate(bad)


def test_float_literals_round_trip():
    for v in (1.0, 3.0e-2, 1e-18, 7.5e-7, 123456.789, 2.0**-126):
        lit = codegen._lit(v)
        assert np.float32(float(lit[:-1])) == np.float32(v), lit


def test_tiled_network_is_block_diagonal():
    """demo_decay tiled three times: each tile's RHS and Jacobian block equal
    the untiled network's on that tile's state, bit for bit, and the
    Jacobian is zero off the diagonal blocks."""
    one, three = O.generated("demo_decay"), O.generated("demo_decay_t3")
    assert three.N == 3 * one.N
    rng = np.random.default_rng(11)
    N = one.N
    for trial in range(10):
        y = rng.uniform(0.01, 2.0, 3 * N).astype(np.float32)
        c = rng.uniform(0.5, 2.0, 2).astype(np.float32)
        for dtype in (np.float32, np.float64):
            f3, J3 = _eval(three, y, c, dtype)
            J3 = J3.reshape(3 * N, 3 * N)  # J3[col, row]
            for t in range(3):
                f1, J1 = _eval(one, y[t * N:(t + 1) * N], c, dtype)
                assert np.array_equal(f3[t * N:(t + 1) * N].view(np.uint8), f1.view(np.uint8))
                blk = J3[t * N:(t + 1) * N, t * N:(t + 1) * N].reshape(-1)
                assert np.array_equal(blk.view(np.uint8), J1.view(np.uint8))
            mask = np.kron(np.eye(3), np.ones((N, N))) == 0
            assert not J3[mask].any()


def test_tile_expansion():
    spec = {"name": "x", "species": 2, "ntile": 2, "nconst": 1,
            "reactions": [{"reactants": [0, 1], "products": [1], "k": 1.0, "const": 0}]}
    t = codegen.tile(spec)
    assert t["species"] == 4 and [r["reactants"] for r in t["reactions"]] == [[0, 1], [2, 3]]
    assert [r["products"] for r in t["reactions"]] == [[1], [3]]
    assert codegen.generate(spec)["N"] == 4


# ------------------------------------------------------- expression networks
G = np.load(os.path.join(os.path.dirname(__file__), "golden", "c_baseline_kats.npz"))


def test_generated_katz96_equals_reference_golden():
    """Katz96 generated from its expressions (networks.katz96_spec) computes
    the reference's compiled calculate_dydt / calculate_jacobian bits
    (golden vectors from cuda/ode_system/ode_gold.c built from source) and
    the oracle's Katz96, in float."""
    gen = O.generated("katz96_gen")
    ref = O.katz96()
    for y, c, f, J in zip(G["katz_y"], G["katz_c"], G["katz_f"], G["katz_J"]):
        fg, Jg = _eval(gen, y, c, np.float32)
        assert np.array_equal(fg.view(np.uint32), np.asarray(f, np.float32).view(np.uint32))
        assert np.array_equal(Jg.view(np.uint32), np.asarray(J, np.float32).view(np.uint32))
        fo, Jo = _eval(ref, y, c, np.float32)
        assert np.array_equal(fg.view(np.uint32), fo.view(np.uint32))
        assert np.array_equal(Jg.view(np.uint32), Jo.view(np.uint32))


def test_generated_nr_test_equals_oracle():
    gen = O.generated("nr_test_gen")
    ref = O.nr_test()
    y0, c, _ = W.nr_test_systems(16, seed=2)
    rng = np.random.default_rng(3)
    for s in range(16):
        y = rng.uniform(-2.0, 2.0, 2).astype(np.float32)
        cs = c[4 * s:4 * s + 4]
        for dtype in (np.float32, np.float64):
            fg, Jg = _eval(gen, y, cs, dtype)
            fo, Jo = _eval(ref, y, cs, dtype)
            assert np.array_equal(fg.view(np.uint8), fo.view(np.uint8))
            assert np.array_equal(Jg.view(np.uint8), Jo.view(np.uint8))


@pytest.mark.parametrize("name,T", [("katz96_t4", 4), ("katz96_t50", W.KATZ96_NTILE)])
def test_tiled_katz96(name, T):
    """Katz96 tiled T times as the reference tiles it (ode_base.py:86-88,
    215-242): with every tile holding the same state (the reference tiles
    its initial conditions, np.tile), each tile's RHS and Jacobian block are
    the untiled network's bit for bit, the Jacobian is block diagonal, and
    the electron density is the first tile's (the reference's
    dconstants_string is inserted once, ahead of the per-tile code)."""
    one, tiled = O.generated("katz96_gen"), O.generated(name)
    assert tiled.N == 5 * T
    for y, c in zip(G["katz_y"][:6], G["katz_c"][:6]):
        yt = np.tile(np.asarray(y, np.float32), T)
        f1, J1 = _eval(one, y, c, np.float32)
        fT, JT = _eval(tiled, yt, c, np.float32)
        JT = JT.reshape(5 * T, 5 * T)  # JT[col, row]
        for t in range(T):
            assert np.array_equal(fT[5 * t:5 * t + 5].view(np.uint32), f1.view(np.uint32))
            blk = JT[5 * t:5 * t + 5, 5 * t:5 * t + 5].reshape(-1)
            assert np.array_equal(blk.view(np.uint32), J1.view(np.uint32))
        assert not JT[np.kron(np.eye(T), np.ones((5, 5))) == 0].any()
    # a different state in the last tile: its rates still use the first tile's ne
    y = np.asarray(G["katz_y"][0], np.float32)
    c = G["katz_c"][0]
    yt = np.tile(y, T)
    yt[-5:] = np.asarray(G["katz_y"][1], np.float32)
    fT, _ = _eval(tiled, yt, c, np.float32)
    ne0 = np.float32(np.float64(np.float32(y[1] + y[3])) + np.float64(y[4]) * 2.0)
    y1 = yt[-5:]
    want = np.float32(np.float32(np.float32(c[2] * ne0) * y1[1]) -
                      np.float32(np.float32(np.float32(c[0] * ne0) + c[1]) * y1[0]))
    assert fT[-5] == want


def test_expression_network_counts_and_checks():
    g = codegen.generate(W.katz96_spec())
    assert g["N"] == 5 and g["nconst"] == 10
    # ne: 3 operations; RHS 6 + 6 + 7 + 14 + 6
    assert g["flops_rhs"] == 42
    g50 = codegen.generate(W.katz96_spec("k", 50))
    assert g50["N"] == 250 and g50["flops_rhs"] == 39 * 50 + 3
    bad = dict(W.katz96_spec("bad"))
    bad["rhs"] = list(bad["rhs"][:-1]) + ["y[5] * c[0]"]
    with pytest.raises(ValueError, match="out of range"):
        codegen.generate(bad)
    bad["rhs"] = list(bad["rhs"][:-1]) + ["nf * y[0]"]
    with pytest.raises(ValueError, match="unknown name"):
        codegen.generate(bad)
    bad["rhs"] = list(bad["rhs"][:-1]) + ["J[0] * y[0]"]
    with pytest.raises(ValueError, match="only readable in Jacobian"):
        codegen.generate(bad)
    bad["rhs"] = list(bad["rhs"][:-1]) + ["(y[0] * c[0]"]
    with pytest.raises(ValueError, match="parentheses"):
        codegen.generate(bad)
