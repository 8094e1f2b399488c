"""Writes FP64 truth solutions for sampled systems of the benchmark workloads:

    tests/golden/truth_hhe6.npz   hhe6 workload ( 2^20 systems, seed 1, 1 Myr): 512 systems,
                                  every 2048th, integrated with scipy's Radau IIA
                                  (order 5) at rtol 1e-10, atol 1e-14 on the float64
                                  network statement in tests/hhe6_f64.py.
    tests/golden/truth_hhe6_tail.npz
                                  the same workload's accuracy tail: 8192 systems, 6144 drawn
                                  stratified over (log T, log n_H) (24 from each cell of a 16 x 16
                                  grid) and the 2048 with the most sub-steps in the oracle's
                                  float32 run at the benchmark's tolerance (adaptive controller,
                                  abs = rel = 2e-6); same Radau settings.
    tests/golden/truth_mass_action.npz
                                  mixed-N workload (2^18 systems, seed 4, t in [0, 1e9]):
                                  32 evenly spaced systems of each N bucket; and the
                                  dense N = 157 workload (65536 systems): 32 systems,
                                  every 2048th.  Radau IIA at rtol 1e-10, atol 1e-20
                                  with the analytic Jacobian of tests/ma_f64.py.

    python tests/golden/make_truth.py [--procs 8] [--only hhe6|hhe6_tail|mass_action]

The benchmark compares its final states for the same systems against these
arrays (final-state relative error vs FP64).
"""
from __future__ import annotations

import argparse
import os
import sys

# one BLAS thread per worker process (the pool supplies the parallelism)
for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ.setdefault(_v, "1")
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from scipy.integrate import solve_ivp  # noqa: E402

import hhe6_f64 as H  # noqa: E402
import ma_f64  # noqa: E402
from wind_amd import networks as W  # noqa: E402

HHE6_NSYS = 1 << 20
HHE6_SEED = 1
HHE6_STRIDE = 2048


def _solve(args):
    y0, nH, p, tend = args
    sol = solve_ivp(H.rhs, (0.0, tend), y0, method="Radau", rtol=1e-10, atol=1e-14, args=(nH, p))
    return sol.status, sol.y[:, -1]


def hhe6_truth(procs):
    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(HHE6_NSYS, seed=HHE6_SEED)
    Y0 = y0.reshape(-1, 6).astype(np.float64)
    idx = np.arange(0, HHE6_NSYS, HHE6_STRIDE)
    jobs = [(Y0[i], float(nH[i]), p.astype(np.float64), W.HHE6_TEND_YR) for i in idx]
    with Pool(procs) as pool:
        res = pool.map(_solve, jobs, chunksize=4)
    status = np.array([r[0] for r in res], np.int32)
    yT = np.array([r[1] for r in res])
    out = os.path.join(HERE, "truth_hhe6.npz")
    np.savez_compressed(out, idx=idx, y=yT, status=status, tend=W.HHE6_TEND_YR, seed=HHE6_SEED,
                        nsys=HHE6_NSYS, params=p)
    print("wrote", out, "failed solves:", int((status != 0).sum()))


TAIL_GRID, TAIL_PER_CELL, TAIL_LONGEST, TAIL_TOL = 16, 24, 2048, 2e-6


def hhe6_tail_selection(y0, nH, steps, seed=11):
    """Indices of the tail sample: TAIL_PER_CELL systems from each cell of a
    TAIL_GRID x TAIL_GRID grid over (log10 T, log10 n_H) (all of a cell's
    systems if it has fewer), then the TAIL_LONGEST systems with the most
    sub-steps not already chosen.  T is recovered from the internal energy,
    e = 1.5 mu T / 1e4."""
    rng = np.random.default_rng(seed)
    Y = y0.reshape(-1, 6).astype(np.float64)
    xe = Y[:, 1] + Y[:, 3] + 2.0 * Y[:, 4]
    mu = Y[:, 0:5].sum(axis=1) + xe
    T = Y[:, 5] * 1.0e4 / (1.5 * mu)
    bt = np.clip(((np.log10(T) - 2.0) / 6.0 * TAIL_GRID).astype(int), 0, TAIL_GRID - 1)
    bn = np.clip(((np.log10(nH) + 4.0) / 8.0 * TAIL_GRID).astype(int), 0, TAIL_GRID - 1)
    cell = bt * TAIL_GRID + bn
    chosen = []
    for cidx in range(TAIL_GRID * TAIL_GRID):
        members = np.nonzero(cell == cidx)[0]
        if len(members):
            chosen.append(rng.choice(members, size=min(TAIL_PER_CELL, len(members)), replace=False))
    strat = np.unique(np.concatenate(chosen))
    order = np.argsort(-steps, kind="stable")
    longest = [i for i in order[: TAIL_LONGEST + len(strat)] if i not in set(strat.tolist())][:TAIL_LONGEST]
    return np.sort(strat), np.array(longest, np.int64)


def hhe6_tail_truth(procs):
    import oracle_lib as O

    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(HHE6_NSYS, seed=HHE6_SEED)
    y = y0.copy()
    _, steps, _ = O.gold_integrate(O.SIE, O.hhe6(p), 0.0, W.HHE6_TEND_YR, 1, nH, y, HHE6_NSYS, TAIL_TOL, TAIL_TOL,
                                   const_stride=1, threads=procs, controller=1)
    strat, longest = hhe6_tail_selection(y0, nH, steps)
    idx = np.concatenate([strat, longest])
    Y0 = y0.reshape(-1, 6).astype(np.float64)
    jobs = [(Y0[i], float(nH[i]), p.astype(np.float64), W.HHE6_TEND_YR) for i in idx]
    with Pool(procs) as pool:
        res = pool.map(_solve, jobs, chunksize=8)
    status = np.array([r[0] for r in res], np.int32)
    yT = np.array([r[1] for r in res])
    out = os.path.join(HERE, "truth_hhe6_tail.npz")
    np.savez_compressed(out, idx=idx, n_stratified=len(strat), y=yT, status=status, tend=W.HHE6_TEND_YR,
                        seed=HHE6_SEED, nsys=HHE6_NSYS, params=p, oracle_substeps=steps[idx])
    print("wrote", out, "systems:", len(idx), "failed solves:", int((status != 0).sum()))


MIXED_NSYS = 1 << 18
MIXED_SEED = 4
PER_BUCKET = 32


def _solve_ma(args):
    N, rx, k, scale, y0, tend = args
    m = ma_f64.MassAction64(N, rx, k, scale)
    sol = solve_ivp(m.rhs, (0.0, tend), y0, method="Radau", rtol=1e-10, atol=1e-20, jac=m.jac)
    return sol.status, sol.y[:, -1]


def mass_action_truth(procs):
    jobs, keys = [], []
    for b in W.mixed_workload(MIXED_NSYS, seed=MIXED_SEED):
        n = len(b["scale"])
        sel = np.linspace(0, n - 1, PER_BUCKET).astype(np.int64)
        Y0 = b["y0"].reshape(n, b["N"]).astype(np.float64)
        for s in sel:
            jobs.append((b["N"], b["rx"], b["k"], float(b["scale"][s]), Y0[s], W.MASS_ACTION_TEND))
        keys.append((f"mixed_{b['N']}", sel))
    rx, k, y0, scale = W.dense_workload(W.DENSE_NSYS)
    sel = np.arange(0, W.DENSE_NSYS, 2048)
    Y0 = y0.reshape(-1, W.DENSE_N).astype(np.float64)
    for s in sel:
        jobs.append((W.DENSE_N, rx, k, float(scale[s]), Y0[s], W.MASS_ACTION_TEND))
    keys.append(("dense_157", sel))
    # longest solves first
    order = sorted(range(len(jobs)), key=lambda i: -jobs[i][0])
    with Pool(procs) as pool:
        res = pool.map(_solve_ma, [jobs[i] for i in order], chunksize=1)
    out_y = [None] * len(jobs)
    status = np.zeros(len(jobs), np.int32)
    for j, i in enumerate(order):
        status[i], out_y[i] = res[j]
    arrays, pos = {}, 0
    for name, sel in keys:
        arrays[name + "_idx"] = sel
        arrays[name + "_y"] = np.array(out_y[pos:pos + len(sel)])
        arrays[name + "_status"] = status[pos:pos + len(sel)]
        pos += len(sel)
    out = os.path.join(HERE, "truth_mass_action.npz")
    np.savez_compressed(out, tend=W.MASS_ACTION_TEND, mixed_nsys=MIXED_NSYS, mixed_seed=MIXED_SEED,
                        dense_nsys=W.DENSE_NSYS, **arrays)
    print("wrote", out, "failed solves:", int((status != 0).sum()))


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=8)
    ap.add_argument("--only", choices=["hhe6", "hhe6_tail", "mass_action"])
    a = ap.parse_args()
    if a.only in (None, "hhe6"):
        hhe6_truth(a.procs)
    if a.only in (None, "hhe6_tail"):
        hhe6_tail_truth(a.procs)
    if a.only in (None, "mass_action"):
        mass_action_truth(a.procs)
