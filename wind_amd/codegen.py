"""Network code generator: a reaction-list description -> a compiled network.

Replaces the reference's Precompiler (python/ode_systems/ode_base.py:16-123),
which splices a network's RHS and Jacobian text into the CUDA and C sources
and rebuilds the libraries.  Here a network description becomes:

  (i)   a HIP header with the network as straight-line code (the RHS and the
        column-major Jacobian, unrolled for the network's N), compiled into a
        plugin library wind_amd/lib/gen/<name>/libwind_net.so that the core
        library loads at run time (wind_network_create_plugin) and integrates
        with its lane-per-system kernels;
  (ii)  the same functions as templated C++ for the oracle (float and double),
        built into oracle/_build/gen/<name>.so for tests;
  (iii) the per-evaluation operation counts (RHS and Jacobian) for the
        roofline accounting;
  (iv)  per-network drop-in libraries sie.so / rk2.so / sie_host.so next to
        the plugin, exporting the reference's mangled entry points.

Description (JSON or dict):

    {"name": "net", "species": N | [names], "nconst": C, "ntile": T,
     "reactions": [{"reactants": [a] | [a, b], "products": [c, ...],
                    "k": float, "const": j | null}, ...]}

"ntile" (default 1) repeats the network T times block-diagonally in one
system of T*N equations, every copy reading the same constants: the
reference's Ntile tiling (ode_base.py:215-242, the reindexed per-tile code
blocks of ode_base.py:86-88).  Tile t's species are t*N .. t*N + N-1 and its
reactions follow tile t-1's.

Reaction rate w = ((k * c[j]) * y_a) * y_b (the factor c[j] only when "const"
is given, y_b only for two reactants).  Each reactant loses w and each
product gains w, once per appearance, in the listed order; reactions are
applied in list order.  That is the operation order of the library's
table-driven mass-action network, so a two-body network generated here and
the same table integrate to identical bits in exact mode.

    python -m wind_amd.codegen spec.json
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "wind_amd")
CSRC = os.path.join(PKG, "csrc")
GEN = os.path.join(PKG, "lib", "gen")
ORACLE_GEN = os.path.join(ROOT, "oracle", "_build", "gen")


def _lit(v: float) -> str:
    """float32 literal that round-trips exactly."""
    x = np.float32(v)
    if not np.isfinite(x):
        raise ValueError(f"non-finite rate coefficient {v}")
    t = format(float(x), ".9g")
    if not any(ch in t for ch in ".e"):
        t += ".0"
    return t + "f"


def tile(spec) -> dict:
    """The description with its "ntile" copies expanded into one network."""
    T = int(spec.get("ntile", 1))
    if T < 1:
        raise ValueError("ntile must be >= 1")
    if T == 1:
        return spec
    N = spec["species"] if isinstance(spec["species"], int) else len(spec["species"])
    out = dict(spec)
    out["ntile"] = 1
    out["species"] = N * T
    out["reactions"] = [
        {**rx, "reactants": [int(s) + t * N for s in rx["reactants"]],
         "products": [int(s) + t * N for s in rx.get("products", [])]}
        for t in range(T) for rx in spec["reactions"]]
    return out


def _check(spec):
    name = spec["name"]
    if not name.isidentifier():
        raise ValueError(f"network name must be an identifier: {name!r}")
    N = spec["species"] if isinstance(spec["species"], int) else len(spec["species"])
    C = int(spec.get("nconst", 1))
    for r, rx in enumerate(spec["reactions"]):
        if len(rx["reactants"]) not in (1, 2):
            raise ValueError(f"reaction {r}: one or two reactants")
        for s in list(rx["reactants"]) + list(rx.get("products", [])):
            if not 0 <= int(s) < N:
                raise ValueError(f"reaction {r}: species {s} out of range")
        j = rx.get("const")
        if j is not None and not 0 <= int(j) < C:
            raise ValueError(f"reaction {r}: constant {j} out of range")
    return name, N, C


def _rate_factor(rx, cexpr):
    k = _lit(rx["k"])
    j = rx.get("const")
    return f"({k} * {cexpr(j)})" if j is not None else k


def _body(spec, N, real, cexpr):
    """RHS and Jacobian statements (lists of C lines) and their op counts."""
    rhs = [f"f[{i}] = {real}(0.0f);" for i in range(N)]
    jac = []
    fr = fj = 0
    for rx in spec["reactions"]:
        reac = [int(s) for s in rx["reactants"]]
        prod = [int(s) for s in rx.get("products", [])]
        ks = _rate_factor(rx, cexpr)
        has_c = rx.get("const") is not None
        targets = [(s, "-") for s in reac] + [(s, "+") for s in prod]
        if real != "":
            ks = ks.replace(_lit(rx["k"]), f"{real}({_lit(rx['k'])})")
        if len(reac) == 2:
            a, b = reac
            w = f"(({ks}) * y[{a}]) * y[{b}]" if has_c else f"({ks} * y[{a}]) * y[{b}]"
            fr += 2 + has_c
        else:
            a, b = reac[0], None
            w = f"({ks}) * y[{a}]" if has_c else f"{ks} * y[{a}]"
            fr += 1 + has_c
        stm = [f"const auto w = {w};"] + [f"f[{t}] = f[{t}] {sg} w;" for t, sg in targets]
        fr += len(targets)
        rhs.append("{ " + " ".join(stm) + " }")
        js = [f"const auto ks = {ks};"] if has_c else [f"const auto ks = {ks};"]
        fj += has_c
        if b is not None:
            js += [f"const auto da = ks * y[{b}];", f"const auto db = ks * y[{a}];"]
            fj += 2
            for t, sg in targets:
                js += [f"J[{a * N + t}] = J[{a * N + t}] {sg} da;", f"J[{b * N + t}] = J[{b * N + t}] {sg} db;"]
                fj += 2
        else:
            js += ["const auto da = ks;"]
            for t, sg in targets:
                js += [f"J[{a * N + t}] = J[{a * N + t}] {sg} da;"]
                fj += 1
        jac.append("{ " + " ".join(js) + " }")
    return rhs, jac, fr, fj


def generate(spec) -> dict:
    """Returns {"header", "plugin", "oracle", "N", "nconst", "R", "flops_rhs", "flops_jac"}."""
    _check(spec)
    spec = tile(spec)
    name, N, C = _check(spec)
    R = len(spec["reactions"])
    rhs, jac, fr, fj = _body(spec, N, "", lambda j: f"c[{j}]")
    ind = "        "
    header = f"""// Generated by wind_amd/codegen.py from network "{name}": {N} species,
// {R} reactions, {C} constants per system.  RHS {fr} operations, Jacobian
// {fj} (column-major, J[j*N + i] = d f_i / d y_j, accumulated onto zeros).
#pragma once
#include "wind_device.h"

namespace wind {{
namespace gen_{name} {{

struct Net {{
    static constexpr int N = {N};
    static constexpr int NCONST = {C};
    struct Params {{}};
    template <class M>
    __device__ static inline void rhs(const Params &, const float *c, const float *y, float *f) {{
{chr(10).join(ind + l for l in rhs)}
    }}
    template <class M>
    __device__ static inline void jac(const Params &, const float *c, const float *y, float *J) {{
{chr(10).join(ind + l for l in jac)}
    }}
}};

}}  // namespace gen_{name}
}}  // namespace wind
"""
    plugin = f"""// Generated by wind_amd/codegen.py: plugin library of network "{name}".
#include "gen_{name}.h"
#include "plugin_impl.h"

WIND_DEFINE_PLUGIN(wind::gen_{name}::Net, "{name}", {R}, {fr}, {fj})
"""
    orhs, ojac, _, _ = _body(spec, N, "R", lambda j: f"c[{j}]")
    oracle = f"""// Generated by wind_amd/codegen.py: oracle (test) statement of network
// "{name}" in float and double, same operation order as the HIP header.
#include <stdint.h>

namespace {{
template <class R>
void rhs(const R *y, const R *c, R *f) {{
{chr(10).join("    " + l for l in orhs)}
}}
template <class R>
void jac(const R *y, const R *c, R *J) {{
{chr(10).join("    " + l for l in ojac)}
}}
}}  // namespace

extern "C" {{
int gen_N(void) {{ return {N}; }}
int gen_nconst(void) {{ return {C}; }}
void gen_rhs_f32(const float *y, const float *c, float *f) {{ rhs<float>(y, c, f); }}
void gen_jac_f32(const float *y, const float *c, float *J) {{ jac<float>(y, c, J); }}
void gen_rhs_f64(const double *y, const double *c, double *f) {{ rhs<double>(y, c, f); }}
void gen_jac_f64(const double *y, const double *c, double *J) {{ jac<double>(y, c, J); }}
}}
"""
    return {"name": name, "header": header, "plugin": plugin, "oracle": oracle, "N": N, "nconst": C, "R": R,
            "flops_rhs": fr, "flops_jac": fj}


def mass_action_spec(name: str, N: int, rx, k, nconst: int = 1) -> dict:
    """Description of a table-driven mass-action network (reaction r:
    rx[4r] + rx[4r+1] -> rx[4r+2] + rx[4r+3], rate k[r] * c[0] * y_a * y_b)."""
    rx = np.asarray(rx, np.int64).reshape(-1, 4)
    k = np.asarray(k, np.float32)
    return {"name": name, "species": N, "nconst": nconst,
            "reactions": [{"reactants": [int(r[0]), int(r[1])], "products": [int(r[2]), int(r[3])],
                           "k": float(kk), "const": 0} for r, kk in zip(rx, k)]}


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")


def _write(path, text):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)


def build(spec, oracle: bool = True, verbose: bool = False) -> str:
    """Generates and compiles a network; returns the plugin library's path."""
    from wind_amd.build import ARCH, HIP_FLAGS, HIPCC, shim_location_flags

    g = generate(spec)
    name = g["name"]
    d = os.path.join(GEN, name)
    src = os.path.join(d, "src")
    _write(os.path.join(src, f"gen_{name}.h"), g["header"])
    _write(os.path.join(src, f"plugin_{name}.hip"), g["plugin"])
    with open(os.path.join(d, "network.json"), "w") as f:
        json.dump({"name": name, "N": g["N"], "nconst": g["nconst"], "R": g["R"], "flops_rhs": g["flops_rhs"],
                   "flops_jac": g["flops_jac"]}, f)
    plugin = os.path.join(d, "libwind_net.so")
    psrc = os.path.join(src, f"plugin_{name}.hip")
    part1, part2 = os.path.join(src, "part1.o"), os.path.join(src, "part2.o")
    # part 1: exact mode, block and lockstep kernels, the descriptor (no FMA
    # contraction); part 2: fast-mode lane kernels (FMA contraction).
    compile_jobs = [
        [HIPCC, *HIP_FLAGS, "-DWIND_PLUGIN_PART=1", f"-I{CSRC}", f"-I{src}", "-c", psrc, "-o", part1],
        [HIPCC, *HIP_FLAGS, "-ffp-contract=fast", "-DWIND_PLUGIN_PART=2", f"-I{CSRC}", f"-I{src}", "-c", psrc,
         "-o", part2],
    ]
    jobs = []
    for kind, lib in ((1, "sie.so"), (2, "rk2.so"), (3, "sie_host.so")):
        jobs.append(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-DWIND_SHIM_KIND={kind}",
                     "-DWIND_SHIM_NETWORK=4", f"-DWIND_SHIM_NUM_CONST={g['nconst']}",
                     *shim_location_flags("../../libwind_amd.so", os.path.join(d, "libwind_net.so")),
                     os.path.join(CSRC, "legacy_shim.cpp"), "-o", os.path.join(d, lib), "-ldl"])
    if oracle and os.path.isdir(os.path.join(ROOT, "oracle")):
        osrc = os.path.join(ORACLE_GEN, f"{name}.cpp")
        _write(osrc, g["oracle"])
        jobs.append(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", osrc,
                     "-o", os.path.join(ORACLE_GEN, f"{name}.so")])
    with ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(_run, compile_jobs + jobs))
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", part1, part2, "-o", plugin + ".tmp"])
    os.replace(plugin + ".tmp", plugin)
    if verbose:
        print("built network", name, "->", plugin)
    return plugin


if __name__ == "__main__":
    sys.path.insert(0, ROOT)
    if len(sys.argv) != 2:
        print(__doc__)
        sys.exit(2)
    with open(sys.argv[1]) as fh:
        build(json.load(fh), verbose=True)
