"""ctypes access to the oracle (oracle/_build/libwind_oracle.so) and, where it
was built, the reference's own c_baseline (oracle/_ref/{sie,rk2}_gold.so).
Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libwind_oracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

NET_KATZ96, NET_NR_TEST, NET_HHE6, NET_MASS_ACTION, NET_GENERATED = 0, 1, 2, 3, 4
GEN_DIR = os.path.join(ROOT, "oracle", "_build", "gen")
SIE, RK2 = 0, 1


class OracleNet(C.Structure):
    _fields_ = [("id", C.c_int), ("N", C.c_int), ("params", C.c_void_p), ("R", C.c_int),
                ("rx", C.c_void_p), ("k", C.c_void_p), ("gen_rhs_f32", C.c_void_p), ("gen_jac_f32", C.c_void_p),
                ("gen_rhs_f64", C.c_void_p), ("gen_jac_f64", C.c_void_p)]


class OracleTrace(C.Structure):
    _fields_ = [("capacity", C.c_int), ("count", C.c_int), ("tnow", C.c_void_p), ("h", C.c_void_p),
                ("flag", C.c_void_p), ("y1", C.c_void_p), ("y2", C.c_void_p), ("accept", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
        L = C.CDLL(ORACLE_SO)
        vp, i, f, d = C.c_void_p, C.c_int, C.c_float, C.c_double
        L.oracle_gold_integrate.argtypes = [i, vp, f, f, i, vp, i, vp, i, i, f, f, vp]
        L.oracle_gold_integrate_ex.argtypes = [i, vp, f, f, i, vp, i, vp, i, i, f, f, vp, C.c_longlong, vp, i]
        L.oracle_gold_integrate_mt.argtypes = [i, vp, f, f, i, vp, i, vp, i, i, f, f, vp, i, i]
        L.oracle_gold_integrate_f64.argtypes = [i, vp, d, d, i, vp, i, vp, i, i, d, d, vp, i]
        L.oracle_integrate_system_traced.argtypes = [i, vp, f, f, f, vp, vp, i, f, f, vp, C.c_longlong, vp, i]
        L.oracle_integrate_traced_batch.argtypes = [i, vp, f, f, i, vp, i, vp, i, i, f, f, i, vp, vp, vp, vp, vp,
                                                    vp, vp, vp, C.c_longlong, i, i]
        L.oracle_integrate_traced_batch.restype = None
        for n in ("oracle_replay_batch", "oracle_replay_batch_f64"):
            getattr(L, n).argtypes = [i, vp, f, vp, i, vp, i, i, i, vp, vp, vp, vp, vp, vp, i, i]
            getattr(L, n).restype = None
        L.oracle_gje_invert.argtypes = [vp, vp, i]
        L.oracle_gje_invert.restype = None
        L.oracle_matvec.argtypes = [vp, vp, i]
        L.oracle_matvec.restype = None
        L.oracle_check_error.argtypes = [i, vp, vp, i, f, f]
        L.oracle_take_step.argtypes = [i, vp, f, f, i, vp, vp, i]
        for n in ("oracle_rhs", "oracle_jac"):
            getattr(L, n).restype = None
        L.oracle_rhs.argtypes = [vp, f, vp, vp, vp]
        L.oracle_jac.argtypes = [vp, vp, vp, vp]
        L.oracle_rhs_f64.argtypes = [vp, d, vp, vp, vp]
        L.oracle_rhs_f64.restype = None
        L.oracle_jac_f64.argtypes = [vp, vp, vp, vp]
        L.oracle_jac_f64.restype = None
        L.oracle_lockstep.argtypes = [vp, f, f, i, vp, i, vp, i, i, f, f]
        L.oracle_lockstep.restype = C.c_longlong
        L.oracle_wexpf.argtypes = [f]
        L.oracle_wexpf.restype = f
        L.oracle_wlogf.argtypes = [f]
        L.oracle_wlogf.restype = f
        _lib = L
    return _lib


class Net:
    """Oracle network descriptor; keeps its arrays alive."""

    def __init__(self, ident, N, params=None, rx=None, k=None):
        self.params = None if params is None else np.ascontiguousarray(params, np.float32)
        self.rx = None if rx is None else np.ascontiguousarray(rx, np.int32).reshape(-1)
        self.k = None if k is None else np.ascontiguousarray(k, np.float32)
        self.N = N
        self.s = OracleNet(ident, N, None if self.params is None else self.params.ctypes.data,
                           0 if self.k is None else len(self.k),
                           None if self.rx is None else self.rx.ctypes.data,
                           None if self.k is None else self.k.ctypes.data)

    @property
    def ref(self):
        return C.byref(self.s)


def katz96():
    return Net(NET_KATZ96, 5)


def nr_test():
    return Net(NET_NR_TEST, 2)


def hhe6(params):
    return Net(NET_HHE6, 6, params=params)


def mass_action(N, rx, k):
    return Net(NET_MASS_ACTION, N, rx=rx, k=k)


def generated(name):
    """A network written by wind_amd/codegen.py (its oracle statement,
    oracle/_build/gen/<name>.so)."""
    g = C.CDLL(os.path.join(GEN_DIR, f"{name}.so"))
    net = Net(NET_GENERATED, g.gen_N())
    net.gen = g  # keep the library loaded
    net.nconst = g.gen_nconst()
    for fld in ("gen_rhs_f32", "gen_jac_f32", "gen_rhs_f64", "gen_jac_f64"):
        setattr(net.s, fld, C.cast(getattr(g, fld), C.c_void_p).value)
    return net


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def gold_integrate(method, net, tnow, tend, nfix, constants, equations, nsys, ABS, REL,
                   const_stride=None, max_attempts=0, threads=1, controller=0):
    """Returns (total, steps[nsys], status[nsys]); `equations` updated in place."""
    L = lib()
    stride = len(constants) // max(nsys, 1) if const_stride is None else const_stride
    steps = np.zeros(nsys, np.int32)
    status = np.zeros(nsys, np.int32)
    if threads > 1:
        tot = L.oracle_gold_integrate_mt(method, net.ref, tnow, tend, nfix, constants.ctypes.data, stride,
                                         equations.ctypes.data, nsys, net.N, ABS, REL, steps.ctypes.data,
                                         threads, controller)
    else:
        tot = L.oracle_gold_integrate_ex(method, net.ref, tnow, tend, nfix, constants.ctypes.data, stride,
                                         equations.ctypes.data, nsys, net.N, ABS, REL, steps.ctypes.data,
                                         max_attempts, status.ctypes.data, controller)
    return tot, steps, status


def lockstep(net, tnow, tend, nfix, constants, equations, nsys, ABS, REL, const_stride):
    """Lockstep SIE (the reference's cudaIntegrateSIE semantics); returns 3*attempts*nsys."""
    return lib().oracle_lockstep(net.ref, tnow, tend, nfix, constants.ctypes.data, const_stride,
                                 equations.ctypes.data, nsys, net.N, ABS, REL)


def gold_integrate_f64(method, net, tnow, tend, nfix, constants, equations, nsys, ABS, REL, const_stride=None,
                       controller=0):
    L = lib()
    stride = len(constants) // max(nsys, 1) if const_stride is None else const_stride
    steps = np.zeros(nsys, np.int32)
    tot = L.oracle_gold_integrate_f64(method, net.ref, tnow, tend, nfix, constants.ctypes.data, stride,
                                      equations.ctypes.data, nsys, net.N, ABS, REL, steps.ctypes.data,
                                      controller)
    return tot, steps


def traced(method, net, tnow, tend, h, y, c, ABS, REL, capacity=100000, max_attempts=1 << 22, controller=0):
    """One system with its attempt trace; returns (substeps, y_final, trace, status)."""
    L = lib()
    N = net.N
    tr_t = np.zeros(capacity, np.float32)
    tr_h = np.zeros(capacity, np.float32)
    tr_f = np.zeros(capacity, np.int32)
    tr_a = np.zeros(capacity, np.int32)
    tr_y1 = np.zeros(capacity * N, np.float32)
    tr_y2 = np.zeros(capacity * N, np.float32)
    tr = OracleTrace(capacity, 0, tr_t.ctypes.data, tr_h.ctypes.data, tr_f.ctypes.data,
                     tr_y1.ctypes.data, tr_y2.ctypes.data, tr_a.ctypes.data)
    y = _f32(y).copy()
    c = _f32(c)
    st = C.c_int32(0)
    n = L.oracle_integrate_system_traced(method, net.ref, tnow, tend, h, y.ctypes.data, c.ctypes.data, N,
                                         ABS, REL, C.byref(tr), max_attempts, C.byref(st), controller)
    k = min(tr.count, capacity)
    return n, y, dict(tnow=tr_t[:k], h=tr_h[:k], flag=tr_f[:k], accept=tr_a[:k], y1=tr_y1[:k * N].reshape(k, N),
                      y2=tr_y2[:k * N].reshape(k, N)), st.value


def traced_batch(method, net, tnow, tend, nfix, constants, equations, nsys, ABS, REL, capacity, const_stride=None,
                 max_attempts=0, controller=0, threads=8, with_states=True):
    """Every system's attempt trace in the library's wind_trace layout.
    `equations` is updated in place.  Returns a dict with h, accept, flag
    [nsys, capacity], y1, y2 [nsys, capacity, N] (when with_states), count,
    steps, status [nsys]; count[s] < 0 marks a system whose -count[s]
    attempts did not fit in the capacity."""
    L = lib()
    N = net.N
    stride = len(constants) // max(nsys, 1) if const_stride is None else const_stride
    h = np.zeros((nsys, capacity), np.float32)
    acc = np.zeros((nsys, capacity), np.int32)
    flag = np.zeros((nsys, capacity), np.int32)
    y1 = np.zeros((nsys, capacity, N), np.float32) if with_states else None
    y2 = np.zeros((nsys, capacity, N), np.float32) if with_states else None
    count = np.zeros(nsys, np.int32)
    steps = np.zeros(nsys, np.int32)
    status = np.zeros(nsys, np.int32)
    c = _f32(constants)
    L.oracle_integrate_traced_batch(method, net.ref, tnow, tend, nfix, c.ctypes.data, stride, equations.ctypes.data,
                                    nsys, N, ABS, REL, capacity, h.ctypes.data, acc.ctypes.data, flag.ctypes.data,
                                    None if y1 is None else y1.ctypes.data, None if y2 is None else y2.ctypes.data,
                                    count.ctypes.data, steps.ctypes.data, status.ctypes.data, max_attempts,
                                    controller, threads)
    return dict(h=h, accept=acc, flag=flag, y1=y1, y2=y2, count=count, steps=steps, status=status)


def start_states(tr, y0, n, N, extrap):
    """The recorded run's state before each attempt [n, capacity, N]: the
    initial state, then the value kept at the last accepted attempt
    (2 y2 - y1 in float32 when extrap, else y2: the reference controller's
    SIE)."""
    cap = int(tr["capacity"])
    acc = tr["accept"].astype(bool) & (np.arange(cap)[None, :] < np.abs(tr["count"])[:, None])
    kept = (np.float32(2) * tr["y2"] - tr["y1"]).astype(np.float32) if extrap else tr["y2"]
    idx = np.where(acc, np.arange(cap)[None, :], -1)
    last = np.maximum.accumulate(idx, axis=1)
    prev = np.concatenate([np.full((n, 1), -1), last[:, :-1]], axis=1)
    ys = np.take_along_axis(kept, np.maximum(prev, 0)[:, :, None], axis=1)
    init = np.broadcast_to(np.asarray(y0, np.float32).reshape(n, 1, N), ys.shape)
    return np.ascontiguousarray(np.where(prev[:, :, None] >= 0, ys, init), dtype=np.float32)


def replay_batch(method, net, tnow, constants, equations, nsys, tr, ystart=None, controller=0, threads=8,
                 f64=True):
    """Replays a trace (dict with h, accept, count and capacity, as
    traced_batch returns) on the recorded run's float time grid, in float64
    (f64) or float32.  Whole run (ystart None): returns the final states
    [nsys, N].  Per attempt (ystart [nsys, capacity, N]): returns (y1, y2),
    each [nsys, capacity, N], every attempt started from its ystart."""
    L = lib()
    N = net.N
    cap = int(tr["capacity"])
    dt = np.float64 if f64 else np.float32
    stride = len(constants) // max(nsys, 1)
    c = _f32(constants)
    h = _f32(tr["h"])
    acc = np.ascontiguousarray(tr["accept"], np.int32)
    cnt = np.ascontiguousarray(np.abs(tr["count"]), np.int32)
    fn = L.oracle_replay_batch_f64 if f64 else L.oracle_replay_batch
    if ystart is None:
        y = np.ascontiguousarray(np.asarray(equations, dt).reshape(nsys, N)).copy()
        fn(method, net.ref, tnow, c.ctypes.data, stride, y.ctypes.data, nsys, N, cap, h.ctypes.data, acc.ctypes.data,
           cnt.ctypes.data, None, None, None, controller, threads)
        return y
    ys = _f32(ystart)
    assert ys.shape == (nsys, cap, N)
    y1 = np.zeros((nsys, cap, N), dt)
    y2 = np.zeros((nsys, cap, N), dt)
    fn(method, net.ref, tnow, c.ctypes.data, stride, None, nsys, N, cap, h.ctypes.data, acc.ctypes.data,
       cnt.ctypes.data, ys.ctypes.data, y1.ctypes.data, y2.ctypes.data, controller, threads)
    return y1, y2


# ----------------------------------------------------------------- reference
def ref_available():
    return all(os.path.exists(os.path.join(REF_DIR, n)) for n in ("sie_gold.so", "rk2_gold.so"))


_ref = {}


def ref(method):
    """The reference c_baseline built from source by oracle/ref.mk (SIE or RK2 library)."""
    if method not in _ref:
        L = C.CDLL(os.path.join(REF_DIR, "sie_gold.so" if method == SIE else "rk2_gold.so"))
        vp, i, f = C.c_void_p, C.c_int, C.c_float
        L.goldIntegrateSystem.argtypes = [f, f, i, vp, vp, i, i, f, f]
        L.integrateSystem.argtypes = [f, f, f, vp, vp, vp, vp, i, f, f]
        L.checkError.argtypes = [vp, vp, i, f, f]
        L.take_step.argtypes = [f, f, i, vp, vp, vp, vp, vp, i]
        L.calculate_dydt.argtypes = [f, vp, vp, vp, i]
        L.calculate_dydt.restype = None
        L.calculate_jacobian.argtypes = [vp, vp, vp]
        L.calculate_jacobian.restype = None
        if method == SIE:
            L.gjeInvertMatrix.argtypes = [vp, vp, i]
            L.gjeInvertMatrix.restype = None
            L.matrixVectorMult.argtypes = [vp, vp, i]
            L.matrixVectorMult.restype = None
        _ref[method] = L
    return _ref[method]
