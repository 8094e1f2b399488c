"""Output-step driver (wind_integrate_outputs) and the ODE-cache writer.

CPU: argument errors before any device call; the cache layout round trip.
GPU: the driver equals the reference harness's loop (runIntegratorOutput,
python/ode_systems/ode_base.py:315-397) done call by call, and the oracle
run interval by interval, bit for bit in exact mode, with the same per-interval
sub-step counts and output times.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from wind_amd import api
from wind_amd import networks as W


def test_output_driver_argument_errors():
    net = api.Network.builtin(api.NET_NR_TEST)
    y0, c, _ = W.nr_test_systems(8, seed=1)
    with pytest.raises(api.WindError, match="n_output_steps"):
        api.integrate_outputs(net, 0.0, 1.0, 0, 10, c, y0.copy(), 8, 1e-3, 1e-3)
    with pytest.raises(api.WindError, match="n_integration_steps"):
        api.integrate_outputs(net, 0.0, 1.0, 4, -1, c, y0.copy(), 8, 1e-3, 1e-3)


def test_cache_writer_layout(tmp_path):
    eot = np.arange(12, dtype=np.float32).reshape(3, 4)
    p = api.write_ode_cache(str(tmp_path / "run.npz"), "SIE", [0.0, 0.5, 1.0], eot, [6, 9], [0.1, 0.2],
                            attrs={"Nsystems": 2, "Nequations_per_system": 2, "absolute": 1e-3})
    p = api.write_ode_cache(p, "RK2", [0.0, 0.5, 1.0], eot + 1, [3, 3], [0.1, 0.1])
    with np.load(p) as f:
        assert np.array_equal(f["SIE/equations_over_time"], eot)
        assert np.array_equal(f["RK2/equations_over_time"], eot + 1)
        assert list(f["SIE/nsteps"]) == [6, 9] and list(f["SIE/times"]) == [0.0, 0.5, 1.0]
        assert "walltimes" in "".join(f.files)
        assert json.loads(str(f["attrs"])) == {}  # the second call passed no attributes


def same(a, b):
    """Bitwise equality, NaN positions equal (the CPU's and GPU's default NaN
    encodings differ in sign)."""
    a = np.asarray(a, np.float32).reshape(-1)
    b = np.asarray(b, np.float32).reshape(-1)
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


def _reference_loop(net, onet, tnow, tend, n_out, nfix, c, y0, nsys, tol, method, stride, max_attempts):
    """The harness's loop: tcur += dt in double, each interval passed as float."""
    dt = (tend - tnow) / n_out
    tcur = tnow
    y_gpu, y_or = y0.copy(), y0.copy()
    eot_gpu, eot_or, times, nst_gpu, nst_or = [y0.copy()], [y0.copy()], [tcur], [], []
    for _ in range(n_out):
        t0, t1 = np.float32(tcur), np.float32(tcur + dt)
        tot, _ = api.integrate(net, float(t0), float(t1), nfix, c, y_gpu, nsys, tol, tol, method=method,
                               mode=api.MODE_EXACT, const_stride=stride, max_attempts=max_attempts)
        tor, _, _ = O.gold_integrate(method, onet, float(t0), float(t1), nfix, c, y_or, nsys, tol, tol,
                                     const_stride=stride, max_attempts=max_attempts)
        print(f"interval [{t0:.6g}, {t1:.6g}]: substeps gpu {tot} oracle {tor}", flush=True)
        tcur += dt
        times.append(tcur)
        eot_gpu.append(y_gpu.copy())
        eot_or.append(y_or.copy())
        nst_gpu.append(tot)
        nst_or.append(tor)
    return np.array(times), np.array(eot_gpu), np.array(eot_or), nst_gpu, nst_or


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["hhe6_sie", "nr_test_rk2"])
def test_output_driver_equals_harness_loop(case):
    cap = 20000
    if case == "hhe6_sie":
        nsys, method, tend, n_out, nfix, stride = 256, api.SIE, 1.0e4, 7, 3, 1
        y0, c = W.hhe6_systems(nsys, seed=21)
        p = W.hhe6_params()
        net, onet = api.Network.builtin(api.NET_HHE6, p), O.hhe6(p)
    else:
        nsys, method, tend, n_out, nfix, stride = 256, api.RK2, 1.0, 5, 200, 4
        y0, c, _ = W.nr_test_systems(nsys, seed=3)
        net, onet = api.Network.builtin(api.NET_NR_TEST), O.nr_test()
    times_ref, eot_gpu, eot_or, nst_gpu, nst_or = _reference_loop(net, onet, 0.0, tend, n_out, nfix, c, y0, nsys,
                                                                  1e-3, method, stride, cap)
    y = y0.copy()
    times, eot, nsteps, step_ms = api.integrate_outputs(net, 0.0, tend, n_out, nfix, c, y, nsys, 1e-3, 1e-3,
                                                        method=method, mode=api.MODE_EXACT, max_attempts=cap)
    print("driver substeps", list(nsteps), flush=True)
    assert np.array_equal(times, times_ref)
    assert np.array_equal(eot.view(np.uint32), eot_gpu.view(np.uint32))
    assert same(eot, eot_or)
    assert np.array_equal(y.view(np.uint32), eot[-1].view(np.uint32))
    assert list(nsteps) == nst_gpu == nst_or
    assert (step_ms >= 0).all()
