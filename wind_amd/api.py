"""ctypes binding of the wind_amd C-ABI (include/wind.h).

The binding loads wind_amd/lib/libwind_amd.so and raises if it is missing:
there is no CPU fallback.  Arrays are numpy float32/int32 on the host, or
device pointers (integers) with ``device_pointers=True`` for state kept on
the GPU (e.g. torch tensors' ``data_ptr()``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libwind_amd.so")

SIE, RK2 = 0, 1
MODE_EXACT, MODE_FAST = 0, 1
KERNEL_AUTO, KERNEL_BLOCK, KERNEL_LANE = 0, 1, 2
CTRL_REFERENCE, CTRL_ADAPTIVE = 0, 1
SOLVER_AUTO, SOLVER_VALU, SOLVER_MFMA_F32, SOLVER_MFMA_BF16, SOLVER_MFMA_BF16X3 = 0, 1, 2, 3, 4
FLAG_GRAPH, FLAG_NO_AUTOTUNE, FLAG_PIVOT, FLAG_DENSE_LDS = 1, 2, 4, 8
NET_KATZ96, NET_NR_TEST, NET_HHE6, NET_MASS_ACTION = 0, 1, 2, 3
STATUS_OK, STATUS_STALLED, STATUS_EXHAUSTED = 0, 1, 2

EXPORTED = (
    "wind_network_create_builtin", "wind_network_create_mass_action", "wind_network_destroy",
    "wind_network_size", "wind_network_nconst", "wind_network_create_plugin", "wind_options_init",
    "wind_integrate", "wind_integrate_noalloc", "wind_integrate_lockstep", "wind_integrate_outputs",
    "wind_last_error", "wind_version", "wind_integrate_batch", "wind_server_create", "wind_server_destroy",
    "wind_server_block_create", "wind_server_block_read", "wind_server_block_destroy", "wind_server_submit",
    "wind_server_wait", "wind_server_inflight",
)


class WindOptions(C.Structure):
    _fields_ = [
        ("method", C.c_int), ("mode", C.c_int), ("kernel", C.c_int), ("const_stride", C.c_int),
        ("device_pointers", C.c_int), ("max_attempts", C.c_int64),
        ("steps_out", C.c_void_p), ("status_out", C.c_void_p), ("total_steps", C.c_void_p),
        ("stream", C.c_void_p), ("kernel_ms", C.c_void_p), ("controller", C.c_int),
        ("solver", C.c_int), ("ir_sweeps", C.c_int), ("flags", C.c_int), ("cost_hint", C.c_void_p),
        ("trace", C.c_void_p),
    ]


class WindTrace(C.Structure):
    _fields_ = [("capacity", C.c_int), ("h", C.c_void_p), ("accept", C.c_void_p), ("count", C.c_void_p),
                ("y_start", C.c_void_p), ("y1", C.c_void_p), ("y2", C.c_void_p), ("flag", C.c_void_p)]


class WindRequestDesc(C.Structure):
    _fields_ = [("net", C.c_void_p), ("constants", C.c_void_p), ("const_stride", C.c_int),
                ("equations", C.c_void_p), ("nsys", C.c_int), ("block", C.c_void_p), ("tnow", C.c_float),
                ("tend", C.c_float), ("n_integration_steps", C.c_int), ("ABSOLUTE", C.c_float),
                ("RELATIVE", C.c_float), ("method", C.c_int), ("mode", C.c_int), ("controller", C.c_int)]


class WindBucket(C.Structure):
    _fields_ = [("net", C.c_void_p), ("constants", C.c_void_p), ("equations", C.c_void_p), ("nsys", C.c_int),
                ("const_stride", C.c_int), ("steps_out", C.c_void_p), ("status_out", C.c_void_p),
                ("cost_hint", C.c_void_p)]


class WindError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    """Loads the native library (once)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise WindError(f"wind_amd native library not built: {path} (run python -m wind_amd.build)")
    # PyTorch ships its own HIP runtime under the same soname.  Loading torch
    # first makes this library bind to that one runtime, so device pointers
    # and streams from torch and from this library live in one HIP context.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    lib.wind_network_create_builtin.argtypes = [C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]
    lib.wind_network_create_mass_action.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                                    C.POINTER(C.c_void_p)]
    lib.wind_network_destroy.argtypes = [C.c_void_p]
    lib.wind_network_destroy.restype = None
    lib.wind_network_size.argtypes = [C.c_void_p]
    lib.wind_network_nconst.argtypes = [C.c_void_p]
    lib.wind_network_create_plugin.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    lib.wind_options_init.argtypes = [C.POINTER(WindOptions)]
    lib.wind_options_init.restype = None
    lib.wind_integrate.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_int, C.c_int, C.c_float, C.c_float, C.POINTER(WindOptions)]
    lib.wind_integrate_noalloc.argtypes = lib.wind_integrate.argtypes
    lib.wind_integrate_outputs.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int, C.c_void_p,
                                           C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float,
                                           C.POINTER(WindOptions), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.wind_integrate_lockstep.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int, C.c_void_p,
                                            C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float,
                                            C.c_int, C.POINTER(C.c_int64)]
    lib.wind_integrate_batch.argtypes = [C.POINTER(WindBucket), C.c_int, C.c_float, C.c_float, C.c_int,
                                         C.c_float, C.c_float, C.POINTER(WindOptions)]
    lib.wind_server_create.argtypes = [C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]
    lib.wind_server_destroy.argtypes = [C.c_void_p]
    lib.wind_server_destroy.restype = None
    lib.wind_server_block_create.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                             C.POINTER(C.c_void_p)]
    lib.wind_server_block_read.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.wind_server_block_destroy.argtypes = [C.c_void_p, C.c_void_p]
    lib.wind_server_block_destroy.restype = None
    lib.wind_server_submit.argtypes = [C.c_void_p, C.POINTER(WindRequestDesc), C.POINTER(C.c_int64)]
    lib.wind_server_wait.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    lib.wind_server_inflight.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_size_t)]
    lib.wind_last_error.restype = C.c_char_p
    lib.wind_version.restype = C.c_char_p
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise WindError(f"wind_amd error {rc}: {load().wind_last_error().decode()}")


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


class Network:
    """A network handle (compiled-in or mass-action)."""

    def __init__(self, handle, N, nconst, name):
        self.handle, self.N, self.nconst, self.name = handle, N, nconst, name

    @classmethod
    def builtin(cls, which, params=None):
        lib = load()
        h = C.c_void_p()
        p = None
        if params is not None:
            params = np.ascontiguousarray(params, dtype=np.float32)
            p = params.ctypes.data
        _check(lib.wind_network_create_builtin(which, p, 0 if params is None else len(params), C.byref(h)))
        N = lib.wind_network_size(h)
        nconst = {NET_KATZ96: 10, NET_NR_TEST: 4, NET_HHE6: 1}[which]
        name = {NET_KATZ96: "katz96", NET_NR_TEST: "nr_test", NET_HHE6: "hhe6"}[which]
        net = cls(h, N, nconst, name)
        net._keep = params
        return net

    @classmethod
    def mass_action(cls, N, rx, k):
        lib = load()
        rx = np.ascontiguousarray(rx, dtype=np.int32).reshape(-1)
        k = np.ascontiguousarray(k, dtype=np.float32)
        h = C.c_void_p()
        _check(lib.wind_network_create_mass_action(N, len(k), rx.ctypes.data, k.ctypes.data, C.byref(h)))
        return cls(h, N, 1, f"mass_action_{N}")

    @classmethod
    def plugin(cls, name_or_path):
        """A network generated by wind_amd/codegen.py: its name (looked up in
        wind_amd/lib/gen/<name>/) or the path of its libwind_net.so."""
        lib = load()
        path = name_or_path
        if not path.endswith(".so"):
            path = os.path.join(_HERE, "lib", "gen", name_or_path, "libwind_net.so")
        h = C.c_void_p()
        _check(lib.wind_network_create_plugin(path.encode(), C.byref(h)))
        return cls(h, lib.wind_network_size(h), lib.wind_network_nconst(h), os.path.basename(os.path.dirname(path)))

    def close(self):
        if self.handle:
            load().wind_network_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def integrate(net: Network, tnow, tend, n_integration_steps, constants, equations, nsys,
              absolute, relative, method=SIE, mode=MODE_EXACT, kernel=KERNEL_AUTO,
              const_stride=0, device_pointers=False, max_attempts=0, steps_out=None,
              status_out=None, stream=None, want_kernel_ms=False, controller=CTRL_REFERENCE,
              solver=SOLVER_AUTO, ir_sweeps=0, cost_hint=None, trace=None, flags=0):
    """Integrates `nsys` systems in place.  Returns (total_substeps, kernel_ms).
    `cost_hint`: optional int32 per-system cost estimates (e.g. a previous
    call's steps_out); lane kernels start systems in decreasing order of it.
    `trace`: optional dict for wind_trace: capacity, and arrays h, accept,
    count, y_start (replay) and y1, y2, flag (record), host or device like the
    state."""
    lib = load()
    o = WindOptions()
    lib.wind_options_init(C.byref(o))
    o.method, o.mode, o.kernel, o.const_stride = method, mode, kernel, const_stride
    o.device_pointers = 1 if device_pointers else 0
    o.controller = controller
    o.solver = solver
    o.ir_sweeps = ir_sweeps
    o.max_attempts = max_attempts
    o.flags = flags
    total = C.c_int64(0)
    kms = C.c_float(0.0)
    o.total_steps = C.cast(C.byref(total), C.c_void_p)
    if want_kernel_ms:
        o.kernel_ms = C.cast(C.byref(kms), C.c_void_p)
    o.steps_out = _ptr(steps_out)
    o.status_out = _ptr(status_out)
    o.cost_hint = _ptr(cost_hint)
    o.stream = stream
    tr = None
    if trace is not None:
        tr = WindTrace(int(trace["capacity"]), *(_ptr(trace.get(k)) for k in ("h", "accept", "count", "y_start",
                                                                             "y1", "y2", "flag")))
        o.trace = C.cast(C.byref(tr), C.c_void_p)
    if not device_pointers:
        for a, dt in ((constants, np.float32), (equations, np.float32)):
            if not (isinstance(a, np.ndarray) and a.dtype == dt and a.flags.c_contiguous):
                raise WindError("host arrays must be C-contiguous float32 numpy arrays")
    _check(lib.wind_integrate(net.handle, tnow, tend, n_integration_steps, _ptr(constants),
                              _ptr(equations), nsys, net.N, absolute, relative, C.byref(o)))
    return total.value, kms.value


def integrate_outputs(net: Network, tnow, tend, n_output_steps, n_integration_steps, constants, equations, nsys,
                      absolute, relative, method=SIE, mode=MODE_EXACT, kernel=KERNEL_AUTO, const_stride=0,
                      max_attempts=0, controller=CTRL_REFERENCE, solver=SOLVER_AUTO, ir_sweeps=0):
    """The reference harness's output loop (runIntegratorOutput) with the state kept on
    the device.  `equations` (host float32) receives the final state.  Returns
    (times[n_out + 1], equations_over_time[n_out + 1, nsys * N] float32,
    nsteps[n_out] int64, step_ms[n_out])."""
    lib = load()
    o = WindOptions()
    lib.wind_options_init(C.byref(o))
    o.method, o.mode, o.kernel, o.const_stride = method, mode, kernel, const_stride
    o.controller, o.solver, o.ir_sweeps, o.max_attempts = controller, solver, ir_sweeps, max_attempts
    for a in (constants, equations):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous):
            raise WindError("host arrays must be C-contiguous float32 numpy arrays")
    eot = np.empty((n_output_steps + 1, nsys * net.N), np.float32)
    times = np.empty(n_output_steps + 1, np.float64)
    nsteps = np.empty(n_output_steps, np.int64)
    step_ms = np.empty(n_output_steps, np.float64)
    _check(lib.wind_integrate_outputs(net.handle, tnow, tend, n_output_steps, n_integration_steps,
                                      constants.ctypes.data, equations.ctypes.data, nsys, net.N, absolute, relative,
                                      C.byref(o), eot.ctypes.data, times.ctypes.data, nsteps.ctypes.data,
                                      step_ms.ctypes.data))
    return times, eot, nsteps, step_ms


def write_ode_cache(path, integrator_name, times, equations_over_time, nsteps, walltimes, attrs=None):
    """Writes one integrator's results in the layout of the reference's ODE
    cache (ode_base.py:371-383 and dumpToODECache :295-309, read by odecache.py:90-158):
    a group named after the integrator with datasets equations_over_time,
    times, nsteps, walltimes, and the run's attributes on the file.  HDF5 when
    h5py is importable, else an .npz holding the same names ("<group>/<dataset>",
    and "attrs" as JSON)."""
    attrs = dict(attrs or {})
    try:
        import h5py  # noqa: F401
    except ImportError:
        h5py = None
    if h5py is not None and not path.endswith(".npz"):
        with h5py.File(path, "a") as f:
            if integrator_name in f:
                del f[integrator_name]
            g = f.create_group(integrator_name)
            g["equations_over_time"] = equations_over_time
            g["times"] = times
            g["nsteps"] = nsteps
            g["walltimes"] = walltimes
            for k, v in attrs.items():
                f.attrs[k] = v
        return path
    import json
    out = path if path.endswith(".npz") else path + ".npz"
    data = {}
    if os.path.exists(out):
        with np.load(out) as old:
            data = {k: old[k] for k in old.files if not k.startswith(integrator_name + "/")}
    data[f"{integrator_name}/equations_over_time"] = np.asarray(equations_over_time)
    data[f"{integrator_name}/times"] = np.asarray(times)
    data[f"{integrator_name}/nsteps"] = np.asarray(nsteps)
    data[f"{integrator_name}/walltimes"] = np.asarray(walltimes)
    data["attrs"] = np.array(json.dumps({k: (v.tolist() if hasattr(v, "tolist") else v) for k, v in attrs.items()}))
    np.savez(out, **data)
    return out


def integrate_lockstep(net: Network, tnow, tend, n_integration_steps, constants, equations, nsys,
                       absolute, relative, const_stride=0):
    lib = load()
    res = C.c_int64(0)
    _check(lib.wind_integrate_lockstep(net.handle, tnow, tend, n_integration_steps, _ptr(constants),
                                       _ptr(equations), nsys, net.N, absolute, relative, const_stride,
                                       C.byref(res)))
    return res.value


def integrate_batch(buckets, tnow, tend, n_integration_steps, absolute, relative, method=SIE, mode=MODE_FAST,
                    controller=CTRL_ADAPTIVE, solver=SOLVER_AUTO, ir_sweeps=0, graph=False, autotune=True,
                    stream=None, max_attempts=0):
    """Integrates a mixed batch.  `buckets` is a list of dicts with keys net (Network),
    constants, equations (device tensors or pointers), nsys, and optional const_stride,
    steps_out, status_out, cost_hint.  Returns (total_substeps, batch_ms)."""
    lib = load()
    arr = (WindBucket * len(buckets))()
    for i, b in enumerate(buckets):
        arr[i] = WindBucket(b["net"].handle, _ptr(b["constants"]), _ptr(b["equations"]), b["nsys"],
                            b.get("const_stride", 0), _ptr(b.get("steps_out")), _ptr(b.get("status_out")),
                            _ptr(b.get("cost_hint")))
    o = WindOptions()
    lib.wind_options_init(C.byref(o))
    o.method, o.mode, o.controller, o.solver, o.ir_sweeps = method, mode, controller, solver, ir_sweeps
    o.flags = (FLAG_GRAPH if graph else 0) | (0 if autotune else FLAG_NO_AUTOTUNE)
    o.device_pointers = 1
    o.max_attempts = max_attempts
    total = C.c_int64(0)
    kms = C.c_float(0.0)
    o.total_steps = C.cast(C.byref(total), C.c_void_p)
    o.kernel_ms = C.cast(C.byref(kms), C.c_void_p)
    o.stream = stream
    _check(lib.wind_integrate_batch(arr, len(buckets), tnow, tend, n_integration_steps, absolute, relative,
                                    C.byref(o)))
    return total.value, kms.value


class Server:
    """Continuous-batching request server (wind_server_* in include/wind.h)."""

    def __init__(self, hbm_budget_bytes, nslots=4):
        self.lib = load()
        h = C.c_void_p()
        _check(self.lib.wind_server_create(int(hbm_budget_bytes), nslots, C.byref(h)))
        self.handle = h
        self._keep = {}

    def block(self, net, constants, equations, nsys, const_stride=0):
        constants = np.ascontiguousarray(constants, np.float32)
        equations = np.ascontiguousarray(equations, np.float32)
        h = C.c_void_p()
        _check(self.lib.wind_server_block_create(self.handle, net.handle, constants.ctypes.data, const_stride,
                                                 equations.ctypes.data, nsys, C.byref(h)))
        return h

    def read_block(self, block, out):
        _check(self.lib.wind_server_block_read(self.handle, block, out.ctypes.data))

    def submit(self, tnow, tend, absolute, relative, net=None, constants=None, equations=None, nsys=0, block=None,
               n_integration_steps=1, method=SIE, mode=MODE_FAST, controller=CTRL_ADAPTIVE, const_stride=0):
        """Queues a request; `equations` (host float32) receives the result.  Returns the id."""
        d = WindRequestDesc(net.handle if net is not None else None,
                            None if constants is None else constants.ctypes.data, const_stride,
                            None if equations is None else equations.ctypes.data, nsys, block, tnow, tend,
                            n_integration_steps, absolute, relative, method, mode, controller)
        rid = C.c_int64(0)
        _check(self.lib.wind_server_submit(self.handle, C.byref(d), C.byref(rid)))
        self._keep[rid.value] = (constants, equations)
        return rid.value

    def wait(self, rid):
        """Returns (latency_us, total_substeps)."""
        lat = C.c_double(0.0)
        steps = C.c_int64(0)
        rc = self.lib.wind_server_wait(self.handle, rid, C.byref(lat), C.byref(steps))
        self._keep.pop(rid, None)
        _check(rc)
        return lat.value, steps.value

    def inflight(self):
        r, p, u = C.c_int(0), C.c_int(0), C.c_size_t(0)
        _check(self.lib.wind_server_inflight(self.handle, C.byref(r), C.byref(p), C.byref(u)))
        return r.value, p.value, u.value

    def close(self):
        if self.handle:
            self.lib.wind_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
