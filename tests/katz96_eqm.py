"""Katz96 rate coefficients and ionisation equilibrium, restated from
python/ode_systems/katz96.py (katz_constants :102-145, get_eqm_abundances
:29-58 and sub_get_densities :60-100, i.e. Katz, Weinberg & Hernquist 1996
eqs. 33-38), for the long-time attractor test of the Katz96 network.  Test
infrastructure only."""
import numpy as np

YR = 3.15e7  # the reference's s -> yr factor (katz96.py:145)


def rates(nH, T):
    """The ten rate coefficients in the reference's order (per second),
    collisional ones multiplied by nH."""
    sq = np.sqrt(T)
    damp = 1.0 + np.sqrt(T / 1e5)
    r = {
        "G_eH0": 5.85e-11 * sq / damp * np.exp(-157809.1 / T) * nH,
        "G_gH0": 4.4e-11,
        "a_Hp": 8.4e-11 / sq * (T / 1e3) ** -0.2 / (1 + (T / 1e6) ** 0.7) * nH,
        "G_eHe0": 2.38e-11 * sq / damp * np.exp(-285335.4 / T) * nH,
        "G_gHe0": 3.7e-12,
        "G_eHep": 5.68e-12 * sq / damp * np.exp(-631515.0 / T) * nH,
        "G_gHep": 1.7e-14,
        "a_Hep": 1.5e-10 * T ** -0.6353 * nH,
        "a_d": 1.9e-3 * T ** -1.5 * np.exp(-470000.0 / T) * (1 + 0.3 * np.exp(-94000.0 / T)) * nH,
        "a_Hepp": 3.36e-10 * T ** -0.5 * (T / 1e3) ** -0.2 / (1 + (T / 1e6) ** 0.7) * nH,
    }
    return r


def constants(nH, T):
    """Per-system constants of the Katz96 network (per year, float32)."""
    r = rates(nH, T)
    keys = ["G_eH0", "G_gH0", "a_Hp", "G_eHe0", "G_gHe0", "G_eHep", "G_gHep", "a_Hep", "a_d", "a_Hepp"]
    return (np.array([r[k] for k in keys]) * YR).astype(np.float32)


def equilibrium(nH, T, y_helium, iterations=1000):
    """Equilibrium abundances per H: [H0, H+, He0, He+, He++] (fixed-point
    iteration on n_e, as get_eqm_abundances does)."""
    r = rates(nH, T)
    for k in r:
        if k.startswith("a_") or k.startswith("G_e"):
            r[k] = r[k] / nH  # per-density rate coefficients
    d = np.zeros(5)
    ne = nH
    for _ in range(iterations):
        d[0] = nH * r["a_Hp"] / (r["a_Hp"] + r["G_eH0"] + r["G_gH0"] / ne)
        d[1] = nH - d[0]
        ion_he0 = r["G_eHe0"] + r["G_gHe0"] / ne
        ion_hep = r["G_eHep"] + r["G_gHep"] / ne
        d[3] = nH * y_helium / 4 / (1 + (r["a_Hep"] + r["a_d"]) / ion_he0 + ion_hep / r["a_Hepp"])
        d[2] = d[3] * (r["a_Hep"] + r["a_d"]) / ion_he0
        d[4] = d[3] * ion_hep / r["a_Hepp"]
        ne = d[1] + d[3] + 2.0 * d[4]
    return d / nH
