"""Seeded synthetic networks and initial conditions for the benchmark workloads.

Networks:
  nr_test      NR eq. 16.6.1 stiff linear pair (python/ode_systems/NR_test.py:34-38),
               per-system constants a * [998, 1998, -999, -1999].
  katz96       the reference's shipped H/He network (cuda/ode_system/ode_gold.c:3-93).
  hhe6         H/He ionisation + internal energy, rates A T^b exp(-C/T) (see
               wind_amd/csrc/wind_device.h, NetHHe6, for the equations).
  mass_action  N species, R = 4N two-body reactions a + b -> c + d.

All generators take an explicit seed (numpy PCG64) and return float32 arrays
laid out as the reference ABI expects: equations[s*N + i], constants[s*C + k].
"""
from __future__ import annotations

import numpy as np

HHE6_DEFAULT_SEED = 20190101
MASS_ACTION_K_RANGE = (1e-18, 1e-6)
MASS_ACTION_Y0_RANGE = (1e-12, 1.0)
# Integration interval of the mass-action workloads, in the network's time unit.
MASS_ACTION_TEND = 1.0e9
HHE6_TEND_YR = 1.0e6
HE_OVER_H = 0.08


def _logu(rng, lo, hi, size=None):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size))


# --------------------------------------------------------------------- NR_test
def nr_test_systems(nsys: int, seed: int = 0):
    """NR_test workload: y0 = (1, 0), constants a*[998, 1998, -999, -1999], a ~ U[0.9, 1.1]."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.9, 1.1, nsys).astype(np.float32)
    base = np.array([998.0, 1998.0, -999.0, -1999.0], dtype=np.float32)
    constants = (a[:, None] * base[None, :]).astype(np.float32)
    y0 = np.tile(np.array([1.0, 0.0], dtype=np.float32), (nsys, 1))
    return y0.reshape(-1).copy(), constants.reshape(-1).copy(), a


def nr_test_exact(a, t):
    """Closed form (NR_test.py:72-77 with the per-system scale a)."""
    a = np.asarray(a, dtype=np.float64)
    y1 = 2.0 * np.exp(-a * t) - np.exp(-1000.0 * a * t)
    y2 = -np.exp(-a * t) + np.exp(-1000.0 * a * t)
    return np.stack([y1, y2], axis=-1)


# --------------------------------------------------------------------- hhe6
def hhe6_params(seed: int = HHE6_DEFAULT_SEED) -> np.ndarray:
    """Network parameters [lnA[7], b[7], C[7], Gamma[3]] (hhe6 workload)."""
    rng = np.random.default_rng(seed)
    A = _logu(rng, 1e-13, 1e-9, 7)
    b = rng.uniform(-0.8, 0.5, 7)
    C = rng.uniform(0.0, 6.0e5, 7)
    G = _logu(rng, 1e-16, 1e-12, 3)
    return np.concatenate([np.log(A), b, C, G]).astype(np.float32)


def hhe6_systems(nsys: int, seed: int = 1):
    """Per-system initial state and constants for the hhe6 workload.

    T ~ logU[1e2, 1e8] K, n_H ~ logU[1e-4, 1e4] cm^-3, He/H = 0.08.  The
    hydrogen ionised fraction is U[0, 1]; helium is split over He0/He+/He++
    by a flat Dirichlet draw.  e = 1.5 mu T / 1e4 with mu the particles per H.
    """
    rng = np.random.default_rng(seed)
    T = _logu(rng, 1e2, 1e8, nsys)
    nH = _logu(rng, 1e-4, 1e4, nsys)
    fH = rng.uniform(0.0, 1.0, nsys)
    he = rng.dirichlet([1.0, 1.0, 1.0], nsys) * HE_OVER_H
    y = np.zeros((nsys, 6), dtype=np.float64)
    y[:, 0] = 1.0 - fH
    y[:, 1] = fH
    y[:, 2:5] = he
    xe = y[:, 1] + y[:, 3] + 2.0 * y[:, 4]
    mu = y[:, 0:5].sum(axis=1) + xe
    y[:, 5] = 1.5 * mu * T / 1.0e4
    return y.astype(np.float32).reshape(-1).copy(), nH.astype(np.float32).copy()


# --------------------------------------------------------------- mass action
def mass_action_network(N: int, seed: int, R: int | None = None):
    """Random two-body network: R = 4N reactions, species uniform, k ~ logU[1e-18, 1e-6]."""
    rng = np.random.default_rng(seed)
    R = 4 * N if R is None else R
    rx = rng.integers(0, N, size=(R, 4)).astype(np.int32)
    k = _logu(rng, *MASS_ACTION_K_RANGE, R).astype(np.float32)
    return rx.reshape(-1).copy(), k


def mass_action_systems(nsys: int, N: int, seed: int, scale_range=(1.0, 1.0)):
    """y0 ~ logU[1e-12, 1] per species; constants[s] = rate scale ~ logU(scale_range)."""
    rng = np.random.default_rng(seed)
    y0 = _logu(rng, *MASS_ACTION_Y0_RANGE, (nsys, N)).astype(np.float32)
    lo, hi = scale_range
    scale = (_logu(rng, lo, hi, nsys) if hi > lo else np.full(nsys, lo)).astype(np.float32)
    return y0.reshape(-1).copy(), scale


# ------------------------------------------------------------------- mixed N
MIXED_SIZES = (5, 10, 32, 64, 157)
MIXED_WEIGHTS = (0.40, 0.25, 0.15, 0.12, 0.08)


def mixed_population(nsys: int, seed: int):
    """Mixed-N population: N drawn from MIXED_SIZES with MIXED_WEIGHTS; returns the N of each system."""
    rng = np.random.default_rng(seed)
    return rng.choice(np.array(MIXED_SIZES), size=nsys, p=np.array(MIXED_WEIGHTS))


MIXED_SCALE_RANGE = (0.1, 10.0)


def mixed_workload(nsys: int, seed: int = 4):
    """Mixed-N workload: buckets of random mass-action networks, one network per N in
    MIXED_SIZES (seeded by N), systems assigned to N with MIXED_WEIGHTS, per-
    system rate scales logU(MIXED_SCALE_RANGE) so step counts vary widely.
    Returns a list of dicts {N, rx, k, y0, scale, index} (index = positions
    of the bucket's systems in the mixed population)."""
    Ns = mixed_population(nsys, seed)
    out = []
    for N in MIXED_SIZES:
        idx = np.nonzero(Ns == N)[0]
        rx, k = mixed_network(N)
        y0, sc = mass_action_systems(len(idx), N, seed=2000 + N + seed, scale_range=MIXED_SCALE_RANGE)
        out.append({"N": N, "rx": rx, "k": k, "y0": y0, "scale": sc, "index": idx})
    return out


# ------------------------------------------------------------ dense large N
DENSE_N = 157
DENSE_NSYS = 65536
DENSE_NET_SEED = 157
DENSE_SYS_SEED = 1157


def dense_workload(nsys: int = DENSE_NSYS):
    """One random mass-action network with N = 157 species and R = 4N
    reactions; y0 ~ logU[1e-12, 1], unit rate scales.  Returns (rx, k, y0, scale)."""
    rx, k = mass_action_network(DENSE_N, seed=DENSE_NET_SEED)
    y0, scale = mass_action_systems(nsys, DENSE_N, seed=DENSE_SYS_SEED)
    return rx, k, y0, scale


# ------------------------------------------------------- generated networks
GENERATED_MIXED = (5, 10)  # mixed-workload networks also built as generated plugins


def mixed_network(N: int):
    """The mixed workload's network for size N (rx, k)."""
    return mass_action_network(N, seed=1000 + N)


def katz96_spec(name: str = "katz96_gen", ntile: int = 1) -> dict:
    """The Katz96 H/He network as an expression description for the code
    generator: the RHS and Jacobian of cuda/ode_system/ode_gold.c:3-93 (the
    same expressions the reference's python/ode_systems/katz96.py:331-448
    hands its precompiler), with the electron density ne as a temporary of
    the first tile and the Jacobian entries in the reference's order,
    J[13] = J[12] included."""
    return {
        "name": name, "species": ["H0", "H+", "He0", "He+", "He++"], "nconst": 10, "ntile": ntile,
        "temps": [["ne", "y[1]+y[3]+y[4]*2.0"]],
        "rhs": [
            "c[2]*ne*y[1] - (c[0]*ne + c[1])*y[0]",
            "-c[2]*ne*y[1] + (c[0]*ne + c[1])*y[0]",
            "(c[7]+c[8])*ne*y[3] - (c[3]*ne+c[4])*y[2]",
            "c[9]*ne*y[4] + (c[3]*ne+c[4])*y[2] - (c[7]+c[8])*ne*y[3] - (c[5]*ne+c[6])*y[3]",
            "(c[5]*ne+c[6])*y[3] - c[9]*ne*y[4]",
        ],
        "jac": [[0, "-(c[0]*ne + c[1])"], [1, "-J[0]"], [6, "-c[2]*ne"], [5, "-J[6]"],
                [12, "-(c[3]*ne+c[4])"], [13, "J[12]"], [19, "c[5]*ne+c[6]"], [17, "(c[7]+c[8])*ne"],
                [18, "-J[17] - J[19]"], [24, "-c[9]*ne"], [23, "-J[24]"]],
    }


def nr_test_spec(name: str = "nr_test_gen", ntile: int = 1) -> dict:
    """NR_test (Numerical Recipes eq. 16.6.1) as the expressions of
    python/ode_systems/NR_test.py:106-134."""
    return {
        "name": name, "species": 2, "nconst": 4, "ntile": ntile,
        "rhs": ["c[0]*y[0] + c[1]*y[1]", "c[2]*y[0] + c[3]*y[1]"],
        "jac": [[0, "c[0]"], [1, "c[2]"], [2, "c[1]"], [3, "c[3]"]],
    }


# Katz96 tiled as in the reference's Ntile scaling study (NtileScaling.ipynb,
# neqn_test.sh): 50 tiles = 250 equations.
KATZ96_NTILE = 50
# Katz96 tilings of the Ntile-scaling benchmark (python/ntbks/NtileScaling.ipynb
# runs Ntile in 1 .. 50, N = 5 .. 250): the ones the build generates.
KATZ96_SCALING_NTILES = (1, 5, 10, 20, 50)

# The reference's Katz96 smoke-test inputs (cuda/debug/inputKatz96_1_1_1.h:
# 4 systems, 5 species, 10 rate constants per system in 1/yr; t in [0, 200],
# n_integration_steps = 1, abs = rel = 5e-3).
KATZ96_FIXTURE_Y = np.array([0, 1.000e+00, 0, 9.929e-02, 0] * 4, np.float32)
KATZ96_FIXTURE_C = np.array(
    [0, 1.386e-03, 4.187e-02, 0, 1.165e-04, 0, 5.355e-07, 2.534e-02, 0, 1.675e-01,
     0, 1.386e-03, 5.271e-02, 0, 1.165e-04, 0, 5.355e-07, 3.190e-02, 0, 2.108e-01,
     0, 1.386e-03, 3.563e-02, 0, 1.165e-04, 0, 5.355e-07, 2.189e-02, 0, 1.425e-01,
     0, 1.386e-03, 4.485e-02, 0, 1.165e-04, 0, 5.355e-07, 2.756e-02, 0, 1.794e-01], np.float32)


def katz96_scaling_systems(ntile: int, nsystem_tile: int = 20):
    """Inputs of the Ntile-scaling runs (Katz96_neqntile.<T>_nsystemtile.20_
    fixed.1_abs.5e3_rel.5e3, python/ntbks/NtileScaling.ipynb): the fixture's
    4 systems repeated nsystem_tile times, each system's 5 species repeated
    ntile times (the tiled network's block-diagonal copies, ode_base.py:215-242),
    10 constants per system shared by its tiles.  Returns (y0, constants)."""
    ys = np.tile(KATZ96_FIXTURE_Y.reshape(4, 5), (nsystem_tile, ntile))
    cs = np.tile(KATZ96_FIXTURE_C.reshape(4, 10), (nsystem_tile, 1))
    return ys.reshape(-1).copy(), cs.reshape(-1).copy()


def generated_specs():
    """Descriptions of the networks that the build generates: the mixed
    workload's N = 5 and N = 10 networks (as "mixed5", "mixed10"), a small
    network with one- and two-body reactions and two constants per system
    ("demo_decay"), that network tiled three times ("demo_decay_t3"), and the
    reference's networks from their expressions: Katz96 ("katz96_gen"), Katz96
    tiled 4, 5, 10, 20 and 50 times ("katz96_t4", N = 20, ... "katz96_t50",
    N = 250; the tilings of the Ntile-scaling benchmark) and
    NR_test ("nr_test_gen"), and a small system with a zero leading pivot
    ("pivot_demo", "pivot_demo_t10")."""
    from wind_amd import codegen

    specs = []
    for N in GENERATED_MIXED:
        rx, k = mixed_network(N)
        specs.append(codegen.mass_action_spec(f"mixed{N}", N, rx, k))
    demo = {
        "name": "demo_decay", "species": 4, "nconst": 2,
        "reactions": [
            {"reactants": [0], "products": [1], "k": 1.0, "const": 0},
            {"reactants": [1], "products": [2], "k": 3.0e-2, "const": None},
            {"reactants": [1, 1], "products": [3], "k": 0.5, "const": 1},
            {"reactants": [2, 3], "products": [0, 1], "k": 2.0e-3, "const": 1},
            {"reactants": [3], "products": [], "k": 1.0e-4, "const": 0},
        ],
    }
    specs.append(demo)
    specs.append({**demo, "name": "demo_decay_t3", "ntile": 3})
    specs.append(katz96_spec())
    specs.append(katz96_spec("katz96_t4", 4))
    for T in KATZ96_SCALING_NTILES:
        if T not in (1, 4, KATZ96_NTILE):
            specs.append(katz96_spec(f"katz96_t{T}", T))
    specs.append(katz96_spec(f"katz96_t{KATZ96_NTILE}", KATZ96_NTILE))
    specs.append(nr_test_spec())
    specs.append(pivot_demo_spec())
    specs.append(pivot_demo_spec("pivot_demo_t10", 10))
    return specs


def pivot_demo_spec(name: str = "pivot_demo", ntile: int = 1) -> dict:
    """A two-species linear system whose Jacobian is [[c0, c1], [c1, -3]]:
    with c0 = 1 / h, the leading pivot of I - h J is zero (test of the
    pivoted LU solves).  Tiled 10 times it takes the one-workgroup-per-
    system kernel (N = 20)."""
    return {
        "name": name, "species": 2, "nconst": 2, "ntile": ntile,
        "rhs": ["c[0]*y[0] + c[1]*y[1]", "c[1]*y[0] - 3.0f*y[1]"],
        "jac": [[0, "c[0]"], [1, "c[1]"], [2, "c[1]"], [3, "-3.0f"]],
    }

