"""wind_integrate from several threads at once: each call takes its own
slot (stream, events, pooled buffers), so calls overlap on the device and
every result equals the same call made alone.  Calls whose matrices exceed
LDS share the matrix scratch and wait for each other there."""
import threading

import numpy as np
import pytest

from wind_amd import api
from wind_amd import networks as W

pytestmark = pytest.mark.gpu


def _jobs():
    p = W.hhe6_params()
    hh = api.Network.builtin(api.NET_HHE6, p)
    jobs = []
    for i in range(4):
        y0, nH = W.hhe6_systems(20000 + 1000 * i, seed=60 + i)
        jobs.append(dict(net=hh, c=nH, y0=y0, n=len(nH), tend=1.0e5, tol=1e-5, mode=api.MODE_FAST, stride=0))
    rx, k = W.mass_action_network(157, seed=W.DENSE_NET_SEED)
    ma = api.Network.mass_action(157, rx, k)
    y0, sc = W.mass_action_systems(6, 157, seed=61, scale_range=(0.5, 2.0))
    # exact mode at N = 157: the matrices live in the shared scratch
    jobs.append(dict(net=ma, c=sc, y0=y0, n=6, tend=1.0e5, tol=1e-3, mode=api.MODE_EXACT, stride=0))
    y0, sc = W.mass_action_systems(512, 157, seed=62, scale_range=(0.5, 2.0))
    jobs.append(dict(net=ma, c=sc, y0=y0, n=512, tend=1.0e7, tol=1e-4, mode=api.MODE_FAST, stride=0))
    return jobs


def _run(j, out, key):
    y = j["y0"].copy()
    steps = np.zeros(j["n"], np.int32)
    tot, _ = api.integrate(j["net"], 0.0, j["tend"], 1, j["c"], y, j["n"], j["tol"], j["tol"], mode=j["mode"],
                           controller=api.CTRL_ADAPTIVE, steps_out=steps)
    out[key] = (y, steps, tot)


def test_concurrent_calls_equal_calls_alone():
    jobs = _jobs()
    alone = {}
    for i, j in enumerate(jobs):
        _run(j, alone, i)
    for rep in range(2):
        together = {}
        threads = [threading.Thread(target=_run, args=(j, together, i)) for i, j in enumerate(jobs)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for i in range(len(jobs)):
            ya, sa, ta = alone[i]
            yt, st, tt = together[i]
            assert ta == tt and np.array_equal(sa, st), (rep, i)
            assert np.array_equal(ya.view(np.uint32), yt.view(np.uint32)), (rep, i)
