"""Writes FP64 truth solutions for sampled systems of the benchmark workloads:

    tests/golden/truth_hhe6.npz   hhe6 workload ( 2^20 systems, seed 1, 1 Myr): 512 systems,
                                  every 2048th, integrated with scipy's Radau IIA
                                  (order 5) at rtol 1e-10, atol 1e-14 on the float64
                                  network statement in tests/hhe6_f64.py.
    tests/golden/truth_mass_action.npz
                                  mixed-N workload (2^18 systems, seed 4, t in [0, 1e9]):
                                  32 evenly spaced systems of each N bucket; and the
                                  dense N = 157 workload (65536 systems): 32 systems,
                                  every 2048th.  Radau IIA at rtol 1e-10, atol 1e-20
                                  with the analytic Jacobian of tests/ma_f64.py.

    python tests/golden/make_truth.py [--procs 8] [--only hhe6|mass_action]

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
    ap.add_argument("--only", choices=["hhe6", "mass_action"])
    a = ap.parse_args()
    if a.only != "mass_action":
        hhe6_truth(a.procs)
    if a.only != "hhe6":
        mass_action_truth(a.procs)
