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

A description may instead give the network as expressions, the way the
reference's network classes hand their RHS and Jacobian text to its
precompiler ("temps", "rhs", "jac_temps", "jac"; the format is described
above _check_expr_network below).  Katz96 and NR_test are generated from
their reference expressions this way (networks.katz96_spec, nr_test_spec).
Networks with N <= 16 get the lane-per-system kernels; larger ones (up to
N = 1023, the reference's fused-kernel limit) a one-workgroup-per-system
kernel with the matrices in LDS or in device scratch; the generated code
loops over tiles, so its size does not grow with ntile.

    python -m wind_amd.codegen spec.json
"""
from __future__ import annotations

import json
import os
import re
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
    """The description with its "ntile" copies expanded into one network
    (reaction-list networks; the generated code itself loops over tiles)."""
    T = int(spec.get("ntile", 1))
    if T < 1:
        raise ValueError("ntile must be >= 1")
    if T == 1:
        return spec
    N = _nspecies(spec)
    out = dict(spec)
    out["ntile"] = 1
    out["species"] = N * T
    out["reactions"] = [
        {**rx, "reactants": [int(s) + t * N for s in rx["reactants"]],
         "products": [int(s) + t * N for s in rx.get("products", [])]}
        for t in range(T) for rx in spec["reactions"]]
    return out


def _nspecies(spec) -> int:
    return spec["species"] if isinstance(spec["species"], int) else len(spec["species"])


def _check(spec):
    name = spec["name"]
    if not name.isidentifier():
        raise ValueError(f"network name must be an identifier: {name!r}")
    N = _nspecies(spec)
    C = int(spec.get("nconst", 1))
    if int(spec.get("ntile", 1)) < 1:
        raise ValueError("ntile must be >= 1")
    if "rhs" in spec:
        _check_expr_network(spec, N, C)
        return name, N, C
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


# ------------------------------------------------------- expression networks
# An expression network gives the RHS and Jacobian as C expressions, as the
# reference's network classes hand its precompiler (e.g.
# python/ode_systems/katz96.py:331-448, NR_test.py:106-134):
#
#   "temps":     [[name, expr], ...]  float temporaries, evaluated before the
#                RHS from the system's first tile (the reference's
#                dconstants_string, inserted once ahead of the per-tile code)
#   "rhs":       [expr_0, ..., expr_{N-1}]     f_i = expr_i
#   "jac_temps": [[name, expr], ...]  as "temps", ahead of the Jacobian
#                (default: "temps"; the reference's jconstants_string)
#   "jac":       [[k, expr], ...]     J[k] = expr in this order, k = j*N + i
#                the column-major index of d f_i / d y_j (entries not listed
#                are zero); expressions may read J entries written before.
#
# Expressions use y[i], c[k], J[k], the temporaries, float or double
# literals (C promotion rules apply, as in the reference's sources),
# + - * / and parentheses.  The reference's names equations[], constants[],
# Jacobian[] and dydt[] are accepted for y[], c[], J[].  With "ntile" T the
# per-tile code is reindexed like the reference's tiling
# (ode_base.py:86-88, reindex): tile t's y[i] is y[t*N + i], its J[k] the
# block-diagonal entry, and the temporaries keep the first tile's indices.

_TOKEN = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][-+]?\d+)?f?|\d*\.\d+(?:[eE][-+]?\d+)?f?|\d+(?:[eE][-+]?\d+)?f?)"
                    r"|([A-Za-z_]\w*)|(\[|\]|\(|\)|\+|-|\*|/))")
_ALIASES = {"equations": "y", "constants": "c", "Jacobian": "J", "jacobian": "J"}


def _tokens(expr):
    pos, out = 0, []
    expr = expr.strip()
    while pos < len(expr):
        m = _TOKEN.match(expr, pos)
        if not m or m.end() == pos:
            raise ValueError(f"cannot parse expression at {expr[pos:pos + 20]!r}")
        num, ident, op = m.groups()
        out.append(("num", num) if num else ("id", _ALIASES.get(ident, ident)) if ident else ("op", op))
        pos = m.end()
    return out


def _check_tokens(toks, where, N, C, names, allow_J):
    depth = 0
    for i, (kind, v) in enumerate(toks):
        if kind == "id":
            if v in ("y", "c", "J"):
                if v == "J" and not allow_J:
                    raise ValueError(f"{where}: J[] is only readable in Jacobian entries")
                if i + 3 >= len(toks) + 1 or toks[i + 1] != ("op", "[") or toks[i + 2][0] != "num" or \
                        toks[i + 3] != ("op", "]"):
                    raise ValueError(f"{where}: {v} must be indexed by an integer constant")
                k = int(toks[i + 2][1])
                lim = {"y": N, "c": C, "J": N * N}[v]
                if not 0 <= k < lim:
                    raise ValueError(f"{where}: {v}[{k}] out of range")
            elif v not in names:
                raise ValueError(f"{where}: unknown name {v!r}")
        elif v == "(":
            depth += 1
        elif v == ")":
            depth -= 1
            if depth < 0:
                raise ValueError(f"{where}: unbalanced parentheses")
    if depth:
        raise ValueError(f"{where}: unbalanced parentheses")


def _check_expr_network(spec, N, C):
    names = set()
    for name, expr in spec.get("temps", []):
        _check_tokens(_tokens(expr), f"temp {name}", N, C, names, False)
        names.add(name)
    if len(spec["rhs"]) != N:
        raise ValueError(f"rhs: {len(spec['rhs'])} expressions for {N} species")
    for i, e in enumerate(spec["rhs"]):
        _check_tokens(_tokens(e), f"rhs[{i}]", N, C, names, False)
    jnames = set()
    for name, expr in spec.get("jac_temps", spec.get("temps", [])):
        _check_tokens(_tokens(expr), f"jac temp {name}", N, C, jnames, False)
        jnames.add(name)
    for k, e in spec.get("jac", []):
        if not 0 <= int(k) < N * N:
            raise ValueError(f"jac entry {k} out of range")
        _check_tokens(_tokens(e), f"jac[{k}]", N, C, jnames, True)


def _emit(toks, NB, S, tile_scope):
    """C text of a token list; y[i] and J[k] reindexed for the tile scope."""
    out, i = [], 0
    while i < len(toks):
        kind, v = toks[i]
        if kind == "id" and v in ("y", "J", "c"):
            k = int(toks[i + 2][1])
            if v == "y":
                out.append(f"yt[{k}]" if tile_scope else f"y[{k}]")
            elif v == "c":
                out.append(f"c[{k}]")
            else:
                out.append(f"Jt[{(k // NB) * S + k % NB}]")
            i += 4
            continue
        out.append(v)
        i += 1
    return " ".join(out)


def _count_ops(toks):
    """Binary arithmetic operations (a leading or post-operator '-' is a sign)."""
    n, prev = 0, None
    for kind, v in toks:
        if kind == "op" and v in "+-*/" and prev is not None and not (prev[0] == "op" and prev[1] in "+-*/("):
            n += 1
        prev = (kind, v)
    return n


def _expr_body(spec, N, S, real):
    """Per-tile RHS and Jacobian lines of an expression network, and op counts
    per tile.  `real` is the declared type of temporaries."""
    temps = [f"const {real} {name} = {_emit(_tokens(e), N, S, False)};" for name, e in spec.get("temps", [])]
    jtemps = [f"const {real} {name} = {_emit(_tokens(e), N, S, False)};"
              for name, e in spec.get("jac_temps", spec.get("temps", []))]
    fr = sum(_count_ops(_tokens(e)) for e in spec["rhs"])
    fr_t = sum(_count_ops(_tokens(e)) for _, e in spec.get("temps", []))
    fj = sum(_count_ops(_tokens(e)) for _, e in spec.get("jac", []))
    fj_t = sum(_count_ops(_tokens(e)) for _, e in spec.get("jac_temps", spec.get("temps", [])))
    rhs = [f"ft[{i}] = {_emit(_tokens(e), N, S, True)};" for i, e in enumerate(spec["rhs"])]
    jac = [f"Jt[{(int(k) // N) * S + int(k) % N}] = {_emit(_tokens(e), N, S, True)};" for k, e in spec.get("jac", [])]
    return temps, rhs, jtemps, jac, (fr, fr_t), (fj, fj_t)


# ---------------------------------------------------- reaction-list networks
def _rate_factor(rx, cexpr):
    k = _lit(rx["k"])
    j = rx.get("const")
    return f"({k} * {cexpr(j)})" if j is not None else k


def _reaction_body(spec, N, S, real):
    """Per-tile RHS and Jacobian statements of a reaction-list network and
    their op counts per tile."""
    cast = "" if real == "float" else real
    rhs = [f"ft[{i}] = {cast}(0.0f);" for i in range(N)]
    jac = []
    fr = fj = 0
    cexpr = lambda j: f"c[{j}]"  # noqa: E731
    for rx in spec["reactions"]:
        reac = [int(s) for s in rx["reactants"]]
        prod = [int(s) for s in rx.get("products", [])]
        ks = _rate_factor(rx, cexpr)
        has_c = rx.get("const") is not None
        targets = [(s, "-") for s in reac] + [(s, "+") for s in prod]
        if cast:
            ks = ks.replace(_lit(rx["k"]), f"{cast}({_lit(rx['k'])})")
        if len(reac) == 2:
            a, b = reac
            w = f"(({ks}) * yt[{a}]) * yt[{b}]" if has_c else f"({ks} * yt[{a}]) * yt[{b}]"
            fr += 2 + has_c
        else:
            a, b = reac[0], None
            w = f"({ks}) * yt[{a}]" if has_c else f"{ks} * yt[{a}]"
            fr += 1 + has_c
        stm = [f"const auto w = {w};"] + [f"ft[{t}] = ft[{t}] {sg} w;" for t, sg in targets]
        fr += len(targets)
        rhs.append("{ " + " ".join(stm) + " }")
        js = [f"const auto ks = {ks};"]
        fj += has_c
        if b is not None:
            js += [f"const auto da = ks * yt[{b}];", f"const auto db = ks * yt[{a}];"]
            fj += 2
            for t, sg in targets:
                js += [f"Jt[{a * S + t}] = Jt[{a * S + t}] {sg} da;", f"Jt[{b * S + t}] = Jt[{b * S + t}] {sg} db;"]
                fj += 2
        else:
            js += ["const auto da = ks;"]
            for t, sg in targets:
                js += [f"Jt[{a * S + t}] = Jt[{a * S + t}] {sg} da;"]
                fj += 1
        jac.append("{ " + " ".join(js) + " }")
    return [], rhs, [], jac, (fr, 0), (fj, 0)


# Generated networks up to this N are also integrated by the lane-per-system
# kernels (whole system in registers, fully unrolled); larger ones by the
# one-workgroup-per-system kernel only, whose code loops over the tiles so
# that its size does not grow with N.
LANE_MAX_N = 16


def generate(spec) -> dict:
    """Returns {"header", "plugin", "oracle", "N", "nconst", "R", "flops_rhs", "flops_jac"}
    (flop counts per evaluation of the whole system)."""
    name, NB, C = _check(spec)
    T = int(spec.get("ntile", 1))
    N = NB * T
    expr = "rhs" in spec
    R = 0 if expr else len(spec["reactions"]) * T
    body = _expr_body if expr else _reaction_body
    temps, rhs, jtemps, jac, (fr, fr_t), (fj, fj_t) = body(spec, NB, N, "float")
    flops_rhs, flops_jac = fr * T + fr_t, fj * T + fj_t
    unroll = "#pragma unroll" if N <= LANE_MAX_N else "#pragma unroll 1"
    ind = "        "
    header = f"""// Generated by wind_amd/codegen.py from network "{name}": {NB} species x {T} tile(s)
// = {N} equations, {C} constants per system.  RHS {flops_rhs} operations, Jacobian
// {flops_jac} (column-major, J[j*N + i] = d f_i / d y_j, written onto zeros).
#pragma once
#include "wind_device.h"

namespace wind {{
namespace gen_{name} {{

struct Net {{
    static constexpr int NB = {NB};      // equations per tile
    static constexpr int NTILE = {T};
    static constexpr int N = {N};
    static constexpr int NCONST = {C};
    struct Params {{}};
    // tile t of the RHS / Jacobian (the tiles are independent)
    template <class M>
    __device__ static inline void rhs_tile(const Params &, const float *c, const float *y, float *f, int t) {{
{chr(10).join(ind + l for l in temps)}
        const float *yt = y + t * NB;
        float *ft = f + t * NB;
{chr(10).join(ind + l for l in rhs)}
    }}
    template <class M>
    __device__ static inline void jac_tile(const Params &, const float *c, const float *y, float *J, int t) {{
{chr(10).join(ind + l for l in jtemps)}
        const float *yt = y + t * NB;
        float *Jt = J + (size_t)t * NB * (N + 1);
{chr(10).join(ind + l for l in jac)}
    }}
    template <class M>
    __device__ static inline void rhs(const Params &p, const float *c, const float *y, float *f) {{
{unroll}
        for (int t = 0; t < NTILE; t++) rhs_tile<M>(p, c, y, f, t);
    }}
    template <class M>
    __device__ static inline void jac(const Params &p, const float *c, const float *y, float *J) {{
{unroll}
        for (int t = 0; t < NTILE; t++) jac_tile<M>(p, c, y, J, t);
    }}
}};

}}  // namespace gen_{name}
}}  // namespace wind
"""
    plugin = f"""// Generated by wind_amd/codegen.py: plugin library of network "{name}".
#include "gen_{name}.h"
#include "plugin_impl.h"

WIND_DEFINE_PLUGIN(wind::gen_{name}::Net, "{name}", {R}, {flops_rhs}, {flops_jac})
"""
    otemps, orhs, ojtemps, ojac, _, _ = body(spec, NB, N, "R")
    oracle = f"""// Generated by wind_amd/codegen.py: oracle (test) statement of network
// "{name}" in float and double, same operation order as the HIP header.
#include <stddef.h>
#include <stdint.h>

namespace {{
constexpr int NB = {NB}, NTILE = {T}, N = {N};
template <class R>
void rhs(const R *y, const R *c, R *f) {{
    for (int t = 0; t < NTILE; t++) {{
{chr(10).join(ind + l for l in otemps)}
        const R *yt = y + t * NB;
        R *ft = f + t * NB;
{chr(10).join(ind + l for l in orhs)}
    }}
}}
template <class R>
void jac(const R *y, const R *c, R *J) {{
    for (int t = 0; t < NTILE; t++) {{
{chr(10).join(ind + l for l in ojtemps)}
        const R *yt = y + t * NB;
        R *Jt = J + (size_t)t * NB * (N + 1);
{chr(10).join(ind + l for l in ojac)}
    }}
}}
}}  // namespace

extern "C" {{
int gen_N(void) {{ return N; }}
int gen_nconst(void) {{ return {C}; }}
void gen_rhs_f32(const float *y, const float *c, float *f) {{ rhs<float>(y, c, f); }}
void gen_jac_f32(const float *y, const float *c, float *J) {{ jac<float>(y, c, J); }}
void gen_rhs_f64(const double *y, const double *c, double *f) {{ rhs<double>(y, c, f); }}
void gen_jac_f64(const double *y, const double *c, double *J) {{ jac<double>(y, c, J); }}
}}
"""
    return {"name": name, "header": header, "plugin": plugin, "oracle": oracle, "N": N, "nconst": C, "R": R,
            "flops_rhs": flops_rhs, "flops_jac": flops_jac}


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
