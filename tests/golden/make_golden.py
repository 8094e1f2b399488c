"""Writes tests/golden/c_baseline_kats.npz: known-answer vectors produced by
the reference's own CPU c_baseline (agurvich/WIND cuda/c_baseline + the shipped
Katz96 cuda/ode_system/ode_gold.c), compiled unmodified from source by
oracle/ref.mk into oracle/_ref/.

    make -f oracle/ref.mk && python tests/golden/make_golden.py

Each case stores its inputs and the reference's outputs (or, for large
matrices, the SHA-256 of the output bytes).  Every integration case is first
run through the oracle with its termination guards to make sure the
reference call returns.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

OUT = os.path.join(HERE, "c_baseline_kats.npz")

# cuda/debug/inputKatz96_1_1_1.h:1-10 (inputs of the reference's smoke test)
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from wind_amd.networks import KATZ96_FIXTURE_C as KATZ_FIXTURE_C  # noqa: E402
from wind_amd.networks import KATZ96_FIXTURE_Y as KATZ_FIXTURE_EQ  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def katz_random_state(rng, n):
    """Katz96-like states: H and He fractions with He/H ~ 0.08, rates like the fixture's."""
    y = np.zeros((n, 5), np.float32)
    fh = rng.uniform(0, 1, n)
    y[:, 0] = 1 - fh
    y[:, 1] = fh
    he = rng.dirichlet([1, 1, 1], n) * 0.0993
    y[:, 2:5] = he
    c = np.exp(rng.uniform(np.log(1e-7), np.log(1.0), (n, 10))).astype(np.float32)
    c[:, [0, 3, 5]] *= rng.uniform(0, 1, (n, 3)) < 0.5  # some collisional rates off, like the fixture
    return y.astype(np.float32), c.astype(np.float32)


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref not built: make -f oracle/ref.mk")
    rs, rr = O.ref(O.SIE), O.ref(O.RK2)
    rng = np.random.default_rng(12345)
    g = {}

    # --- gjeInvertMatrix (gje_gold.c:136-144) ---------------------------------
    for N in (2, 5, 6, 10, 32, 64, 157):
        for t in range(3 if N <= 64 else 1):
            J = rng.standard_normal((N, N)).astype(np.float32) * np.float32(rng.uniform(0.1, 3.0))
            h = np.float32(rng.uniform(0.01, 0.3))
            M = (np.eye(N, dtype=np.float32) - h * J).astype(np.float32)
            if t == 2:  # diagonally dominant variant
                M += np.eye(N, dtype=np.float32) * np.float32(N)
            Min = M.copy()
            Inv = np.zeros_like(M)
            rs.gjeInvertMatrix(M.ctypes.data, Inv.ctypes.data, N)
            key = f"gje_N{N}_{t}"
            g[key + "_in"] = Min
            if N <= 32:
                g[key + "_inv"] = Inv
                g[key + "_mout"] = M
            else:
                g[key + "_inv_sha"] = np.array(sha(Inv))
                g[key + "_mout_sha"] = np.array(sha(M))

    # --- matrixVectorMult (matrix_ops.c:6-27) --------------------------------
    for N in (2, 5, 6, 32):
        M = rng.standard_normal((N, N)).astype(np.float32)
        v = rng.standard_normal(N).astype(np.float32)
        vout = v.copy()
        rs.matrixVectorMult(M.ctypes.data, vout.ctypes.data, N)
        g[f"matvec_N{N}_M"], g[f"matvec_N{N}_v"], g[f"matvec_N{N}_out"] = M, v, vout

    # --- checkError (common_gold.c:8-46): boundary table ---------------------
    ABS, REL = np.float32(1e-3), np.float32(1e-2)
    pairs = [
        (0.5, 0.5), (0.5, 0.5 + 1e-3), (0.5, 0.5 + 9.9e-4), (0.5, 0.5 + 1.1e-3),
        (1.0, 1.0 + 1e-3), (0.2, 0.2 + 0.002), (0.2, 0.2 + 0.0021),
        (1e-4, 2e-4), (1e-3, 1.009e-3), (2e-3, 2.00002e-3), (0.0, 0.0), (-0.5, -0.5049),
        (-0.5, -0.506), (0.1, 0.101), (0.1, 0.1011), (1e-12, -1e-12), (0.05, 0.0505), (5.0, 5.00099),
        (0.0011, 0.00111), (0.0011, 0.0011112),
    ]
    y1 = np.array([p[0] for p in pairs], np.float32)
    y2 = np.array([p[1] for p in pairs], np.float32)
    for name, L in (("sie", rs), ("rk2", rr)):
        flags = np.array([L.checkError(y1[i:i + 1].ctypes.data, y2[i:i + 1].ctypes.data, 1, ABS, REL)
                          for i in range(len(pairs))], np.int32)
        g[f"check_{name}_flags"] = flags
        g[f"check_{name}_all"] = np.array(L.checkError(y1.ctypes.data, y2.ctypes.data, len(y1), ABS, REL))
    g["check_y1"], g["check_y2"], g["check_abs_rel"] = y1, y2, np.array([ABS, REL], np.float32)

    # --- Katz96 network (ode_gold.c:3-93) ------------------------------------
    ys, cs = katz_random_state(rng, 20)
    F = np.zeros((20, 5), np.float32)
    Jm = np.zeros((20, 25), np.float32)
    for i in range(20):
        rs.calculate_dydt(np.float32(0), ys[i].ctypes.data, cs[i].ctypes.data, F[i].ctypes.data, 5)
        rs.calculate_jacobian(ys[i].ctypes.data, cs[i].ctypes.data, Jm[i].ctypes.data)
    g["katz_y"], g["katz_c"], g["katz_f"], g["katz_J"] = ys, cs, F, Jm

    # --- take_step (sie_gold.c:38-79, rk2_gold.c:18-46) -----------------------
    ys, cs = katz_random_state(rng, 10)
    hs = np.exp(rng.uniform(np.log(1e-2), np.log(10.0), 10)).astype(np.float32)
    for name, L in (("sie", rs), ("rk2", rr)):
        for nsub in (1, 2):
            out = ys.copy()
            dydt = np.zeros(5, np.float32)
            Jb = np.zeros(25, np.float32)
            Ib = np.zeros(25, np.float32)
            for i in range(10):
                L.take_step(np.float32(0), hs[i], nsub, out[i].ctypes.data, cs[i].ctypes.data,
                            dydt.ctypes.data, Jb.ctypes.data, Ib.ctypes.data, 5)
            g[f"take_{name}_n{nsub}"] = out
    g["take_y"], g["take_c"], g["take_h"] = ys, cs, hs

    # --- goldIntegrateSystem on the reference's debug fixture -----------------
    # NUM_CONST = 2 as shipped: system s reads constants[2s : 2s+10].
    for tend in (10.0, 200.0):
        for name, L, meth in (("sie", rs, O.SIE), ("rk2", rr, O.RK2)):
            eq = KATZ_FIXTURE_EQ.copy()
            chk = eq.copy()
            _, _, st = O.gold_integrate(meth, O.katz96(), 0.0, tend, 1, KATZ_FIXTURE_C, chk, 4, 5e-3, 5e-3,
                                        const_stride=2, max_attempts=1 << 22)
            assert (st == 0).all(), "oracle guard fired; reference call would not return"
            n = L.goldIntegrateSystem(0.0, tend, 1, KATZ_FIXTURE_C.ctypes.data, eq.ctypes.data, 4, 5,
                                      5e-3, 5e-3)
            g[f"fixture_t{int(tend)}_{name}_eq"] = eq
            g[f"fixture_t{int(tend)}_{name}_n"] = np.array(n)

    # --- integrateSystem per system (10 constants each) -----------------------
    # Keep systems for which every (method, tolerance) run returns: explicit
    # RK2 on a stiff system can force-accept unstable steps indefinitely.
    cand_y, cand_c = katz_random_state(rng, 96)
    keep = []
    for s in range(96):
        ok = True
        for tol in (5e-3, 1e-4, 1e-6):
            for meth in (O.SIE, O.RK2):
                chk = cand_y[s].copy()
                _, _, st = O.gold_integrate(meth, O.katz96(), 0.0, 50.0, 1, cand_c[s], chk, 1, tol, tol,
                                            const_stride=10, max_attempts=1 << 20)
                ok = ok and st[0] == 0
        if ok:
            keep.append(s)
    ys, cs = cand_y[keep[:48]], cand_c[keep[:48]]
    ns = len(ys)
    for tol in (5e-3, 1e-4, 1e-6):
        for name, L, meth in (("sie", rs, O.SIE), ("rk2", rr, O.RK2)):
            out = ys.copy()
            steps = np.zeros(ns, np.int32)
            Jb = np.zeros(25, np.float32)
            Ib = np.zeros(25, np.float32)
            tend = np.float32(50.0)
            for s in range(ns):
                steps[s] = L.integrateSystem(np.float32(0), tend, tend / np.float32(1),
                                             out[s].ctypes.data, cs[s].ctypes.data, Jb.ctypes.data,
                                             Ib.ctypes.data, 5, tol, tol)
            g[f"sys_tol{tol:g}_{name}_eq"] = out
            g[f"sys_tol{tol:g}_{name}_steps"] = steps
    g["sys_y"], g["sys_c"] = ys, cs

    np.savez_compressed(OUT, **g)
    print("wrote", OUT, os.path.getsize(OUT), "bytes,", len(g), "arrays")


if __name__ == "__main__":
    main()
