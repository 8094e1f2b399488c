"""The drop-in install as WIND finds it: the shims symlinked (or copied) into
a fake <WIND>/cuda/lib, WIND's gold CPU libraries beside them, and the
library lookups of python/wind_harness.py:18-54 (loadCLibraries) run from
<WIND>/python in a fresh process with no LD_LIBRARY_PATH.

CPU: the five lookups resolve, initializeCublasExternally() loads the core
library from the build tree (through the symlink's target or, for a copy,
the build-time path) and sets up the network.  GPU: the three GPU entry
points called through that layout give the oracle's bits."""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "wind_amd", "lib")

# Restates the lookups of loadCLibraries (python/wind_harness.py:18-54): the
# directory is the parent of the working directory, and the five libraries
# are loaded in this order.  With "call" set it also runs each GPU entry
# point on the given inputs and prints the results as JSON.
HARNESS = r'''
import ctypes, json, os, sys
import numpy as np
curdir = os.path.split(os.getcwd())[0]
out = {}
c_obj = ctypes.CDLL(os.path.join(curdir, "cuda", "lib", "sie_host.so"))
cublas_init = getattr(c_obj, "_Z26initializeCublasExternallyv")
sie_host = getattr(c_obj, "_Z16cudaIntegrateSIEffiPfS_iiff")
c_obj = ctypes.CDLL(os.path.join(curdir, "cuda", "lib", "sie.so"))
sie = getattr(c_obj, "_Z19cudaIntegrateSystemffiPfS_iiff")
c_obj = ctypes.CDLL(os.path.join(curdir, "cuda", "lib", "rk2.so"))
rk2 = getattr(c_obj, "_Z19cudaIntegrateSystemffiPfS_iiff")
if os.path.exists(os.path.join(curdir, "cuda", "lib", "rk2_gold.so")):
    c_obj = ctypes.CDLL(os.path.join(curdir, "cuda", "lib", "rk2_gold.so"))
    getattr(c_obj, "goldIntegrateSystem")
    c_obj = ctypes.CDLL(os.path.join(curdir, "cuda", "lib", "sie_gold.so"))
    getattr(c_obj, "goldIntegrateSystem")
    out["gold"] = True
cublas_init()
maps = open("/proc/self/maps").read()
out["core"] = sorted({l.split()[-1] for l in maps.splitlines() if l.endswith("libwind_amd.so")})
if len(sys.argv) > 1:
    spec = json.load(open(sys.argv[1]))
    c = np.array(spec["c"], np.float32)
    P = ctypes.POINTER(ctypes.c_float)
    for name, fn in (("sie", sie), ("rk2", rk2), ("sie_host", sie_host)):
        y = np.array(spec["y"], np.float32)
        n = fn(ctypes.c_float(0.0), ctypes.c_float(spec["tend"]), ctypes.c_int(spec["nfix"]), c.ctypes.data_as(P),
               y.ctypes.data_as(P), ctypes.c_int(spec["nsys"]), ctypes.c_int(spec["N"]), ctypes.c_float(spec["tol"]),
               ctypes.c_float(spec["tol"]))
        out[name] = {"n": int(n), "y": y.view(np.uint32).tolist()}
print("RESULT " + json.dumps(out))
'''


def make_layout(tmp, network, copy=False):
    wind = os.path.join(str(tmp), "WIND")
    lib = os.path.join(wind, "cuda", "lib")
    os.makedirs(lib)
    os.makedirs(os.path.join(wind, "python"))
    for so in ("sie.so", "rk2.so", "sie_host.so"):
        src = os.path.join(LIB, network, so)
        (shutil.copy if copy else os.symlink)(src, os.path.join(lib, so))
    if O.ref_available():
        for so in ("sie_gold.so", "rk2_gold.so"):
            shutil.copy(os.path.join(O.REF_DIR, so), os.path.join(lib, so))
    with open(os.path.join(wind, "python", "harness_lookup.py"), "w") as f:
        f.write(HARNESS)
    return wind


def run_harness(wind, *args):
    env = {k: v for k, v in os.environ.items() if k not in ("LD_LIBRARY_PATH", "WIND_AMD_LIBRARY", "LD_PRELOAD")}
    r = subprocess.run([sys.executable, "harness_lookup.py", *args], cwd=os.path.join(wind, "python"), env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT ")][-1]
    return json.loads(line[len("RESULT "):])


@pytest.mark.parametrize("copy", [False, True])
def test_harness_lookups_resolve(tmp_path, copy):
    wind = make_layout(tmp_path, "hhe6", copy=copy)
    out = run_harness(wind)
    assert out["core"] == [os.path.realpath(os.path.join(LIB, "libwind_amd.so"))]
    if O.ref_available():
        assert out.get("gold")


def test_generated_network_shims_resolve(tmp_path):
    """A generated network's shims find the core library two levels up and
    the network plugin beside their real path."""
    wind = make_layout(tmp_path, os.path.join("gen", "demo_decay"))
    out = run_harness(wind)
    assert out["core"] == [os.path.realpath(os.path.join(LIB, "libwind_amd.so"))]


@pytest.mark.gpu
@pytest.mark.parametrize("copy", [False, True])
def test_harness_calls_match_oracle(tmp_path, copy):
    """sie.so, rk2.so and sie_host.so of the NR_test network called through
    the WIND layout: the oracle's bits and sub-step totals."""
    from wind_amd import networks as W

    ns = 128
    y0, c, _ = W.nr_test_systems(ns, seed=31)
    wind = make_layout(tmp_path, "nr_test", copy=copy)
    spec = os.path.join(str(tmp_path), "spec.json")
    with open(spec, "w") as f:
        json.dump({"y": y0.tolist(), "c": c.tolist(), "nsys": ns, "N": 2, "tend": 1.0, "nfix": 1000, "tol": 1e-3},
                  f)
    out = run_harness(wind, spec)
    for name, method in (("sie", O.SIE), ("rk2", O.RK2)):
        yr = y0.copy()
        tot, _, _ = O.gold_integrate(method, O.nr_test(), 0.0, 1.0, 1000, c, yr, ns, 1e-3, 1e-3, const_stride=4)
        assert out[name]["n"] == tot, name
        assert np.array_equal(np.array(out[name]["y"], np.uint32), yr.view(np.uint32)), name
    yr = y0.copy()
    want = O.lockstep(O.nr_test(), 0.0, 1.0, 1000, c, yr, ns, 1e-3, 1e-3, 4)
    assert out["sie_host"]["n"] == want
    assert np.array_equal(np.array(out["sie_host"]["y"], np.uint32), yr.view(np.uint32))
