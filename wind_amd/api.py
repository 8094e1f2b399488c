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
NET_KATZ96, NET_NR_TEST, NET_HHE6, NET_MASS_ACTION = 0, 1, 2, 3
STATUS_OK, STATUS_STALLED, STATUS_EXHAUSTED = 0, 1, 2

EXPORTED = (
    "wind_network_create_builtin", "wind_network_create_mass_action", "wind_network_destroy",
    "wind_network_size", "wind_options_init", "wind_integrate", "wind_integrate_lockstep",
    "wind_last_error", "wind_version",
)


class WindOptions(C.Structure):
    _fields_ = [
        ("method", C.c_int), ("mode", C.c_int), ("kernel", C.c_int), ("const_stride", C.c_int),
        ("device_pointers", C.c_int), ("max_attempts", C.c_int64),
        ("steps_out", C.c_void_p), ("status_out", C.c_void_p), ("total_steps", C.c_void_p),
        ("stream", C.c_void_p), ("kernel_ms", C.c_void_p), ("controller", C.c_int),
    ]


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
    lib = C.CDLL(path)
    lib.wind_network_create_builtin.argtypes = [C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]
    lib.wind_network_create_mass_action.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                                    C.POINTER(C.c_void_p)]
    lib.wind_network_destroy.argtypes = [C.c_void_p]
    lib.wind_network_destroy.restype = None
    lib.wind_network_size.argtypes = [C.c_void_p]
    lib.wind_options_init.argtypes = [C.POINTER(WindOptions)]
    lib.wind_options_init.restype = None
    lib.wind_integrate.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_int, C.c_int, C.c_float, C.c_float, C.POINTER(WindOptions)]
    lib.wind_integrate_lockstep.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int, C.c_void_p,
                                            C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float,
                                            C.c_int, C.POINTER(C.c_int64)]
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
              status_out=None, stream=None, want_kernel_ms=False, controller=CTRL_REFERENCE):
    """Integrates `nsys` systems in place.  Returns (total_substeps, kernel_ms)."""
    lib = load()
    o = WindOptions()
    lib.wind_options_init(C.byref(o))
    o.method, o.mode, o.kernel, o.const_stride = method, mode, kernel, const_stride
    o.device_pointers = 1 if device_pointers else 0
    o.controller = controller
    o.max_attempts = max_attempts
    total = C.c_int64(0)
    kms = C.c_float(0.0)
    o.total_steps = C.cast(C.byref(total), C.c_void_p)
    if want_kernel_ms:
        o.kernel_ms = C.cast(C.byref(kms), C.c_void_p)
    o.steps_out = _ptr(steps_out)
    o.status_out = _ptr(status_out)
    o.stream = stream
    if not device_pointers:
        for a, dt in ((constants, np.float32), (equations, np.float32)):
            if not (isinstance(a, np.ndarray) and a.dtype == dt and a.flags.c_contiguous):
                raise WindError("host arrays must be C-contiguous float32 numpy arrays")
    _check(lib.wind_integrate(net.handle, tnow, tend, n_integration_steps, _ptr(constants),
                              _ptr(equations), nsys, net.N, absolute, relative, C.byref(o)))
    return total.value, kms.value


def integrate_lockstep(net: Network, tnow, tend, n_integration_steps, constants, equations, nsys,
                       absolute, relative, const_stride=0):
    lib = load()
    res = C.c_int64(0)
    _check(lib.wind_integrate_lockstep(net.handle, tnow, tend, n_integration_steps, _ptr(constants),
                                       _ptr(equations), nsys, net.N, absolute, relative, const_stride,
                                       C.byref(res)))
    return res.value
