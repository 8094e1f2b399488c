"""The C-ABI without a GPU: the libraries load, export every entry point
include/wind.h declares (and the reference's mangled names in the shims),
and argument errors are reported before any device call."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from wind_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "wind_amd", "lib")


def declared_functions():
    text = open(os.path.join(ROOT, "include", "wind.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wind_[a-z_0-9]+)\s*\(", text)))


def dynamic_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_header_functions_exported():
    names = declared_functions()
    assert "wind_integrate" in names and len(names) >= 9
    syms = dynamic_symbols(os.path.join(LIB, "libwind_amd.so"))
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    lib = api.load()
    for n in names:
        getattr(lib, n)


@pytest.mark.parametrize("net", ["nr_test", "katz96", "katz96_numconst2", "hhe6"])
def test_shims_export_reference_symbols(net):
    """The names python/wind_harness.py:24-53 looks up with getattr()."""
    d = os.path.join(LIB, net)
    assert "_Z19cudaIntegrateSystemffiPfS_iiff" in dynamic_symbols(os.path.join(d, "sie.so"))
    assert "_Z19cudaIntegrateSystemffiPfS_iiff" in dynamic_symbols(os.path.join(d, "rk2.so"))
    host = dynamic_symbols(os.path.join(d, "sie_host.so"))
    assert "_Z16cudaIntegrateSIEffiPfS_iiff" in host and "_Z26initializeCublasExternallyv" in host
    for lib in ("sie.so", "rk2.so"):
        getattr(C.CDLL(os.path.join(d, lib)), "_Z19cudaIntegrateSystemffiPfS_iiff")


def test_version_and_no_error():
    lib = api.load()
    assert lib.wind_version().decode().startswith("wind_amd")


def test_argument_errors_before_device():
    lib = api.load()
    with pytest.raises(api.WindError, match="unknown builtin"):
        api.Network.builtin(7)
    with pytest.raises(api.WindError, match="24 network parameters"):
        api.Network.builtin(api.NET_HHE6)
    net = api.Network.builtin(api.NET_NR_TEST)
    y = np.zeros(4, np.float32)
    c = np.zeros(8, np.float32)
    o = api.WindOptions()
    lib.wind_options_init(C.byref(o))
    assert lib.wind_integrate(net.handle, 0.0, 1.0, 1, c.ctypes.data, y.ctypes.data, 1, 3, 1e-3, 1e-3,
                              C.byref(o)) == -1
    assert b"does not match the network" in lib.wind_last_error()
    assert lib.wind_integrate(None, 0.0, 1.0, 1, None, None, 1, 2, 1e-3, 1e-3, C.byref(o)) == -1
    assert b"NULL" in lib.wind_last_error()
    rc = lib.wind_integrate(net.handle, 0.0, 1.0, -1, c.ctypes.data, y.ctypes.data, 2, 2, 1e-3, 1e-3, C.byref(o))
    assert rc == -1 and b"n_integration_steps" in lib.wind_last_error()
    o.method = 5
    assert lib.wind_integrate(net.handle, 0.0, 1.0, 1, c.ctypes.data, y.ctypes.data, 2, 2, 1e-3, 1e-3,
                              C.byref(o)) == -1
    o.method = 0
    o.controller = 9
    assert lib.wind_integrate(net.handle, 0.0, 1.0, 1, c.ctypes.data, y.ctypes.data, 2, 2, 1e-3, 1e-3,
                              C.byref(o)) == -1
    o.controller = 0
    # nothing to integrate: the reference makes no attempt (common_gold.c:77), no device call either
    y[:] = [1, 2, 3, 4]
    assert lib.wind_integrate(net.handle, 1.0, 1.0, 1, c.ctypes.data, y.ctypes.data, 2, 2, 1e-3, 1e-3,
                              C.byref(o)) == 0
    assert lib.wind_integrate(net.handle, 0.0, 1.0, 1, c.ctypes.data, y.ctypes.data, 0, 2, 1e-3, 1e-3,
                              C.byref(o)) == 0
    assert list(y) == [1, 2, 3, 4]


def test_mass_action_table_validation():
    with pytest.raises(api.WindError, match="species index"):
        api.Network.mass_action(4, np.array([0, 1, 2, 9], np.int32), np.ones(1, np.float32))
    with pytest.raises(api.WindError, match="N out of range"):
        api.Network.mass_action(0, np.zeros(4, np.int32), np.ones(1, np.float32))


def test_options_struct_matches_header():
    """api.WindOptions (ctypes) has the members of wind_options in
    include/wind.h, in order and with matching kinds (pointer, int, int64),
    so wind_options_init and wind_integrate read what Python writes."""
    text = open(os.path.join(ROOT, "include", "wind.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = re.search(r"typedef struct wind_options\s*\{(.*?)\}\s*wind_options;", text, flags=re.S).group(1)
    members = []
    for decl in body.split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        ctype, name = decl.rsplit(" ", 1)
        ptr = "*" in ctype or name.startswith("*")
        name = name.lstrip("*")
        kind = "ptr" if ptr else ("int64" if "int64_t" in ctype else "int")
        members.append((name, kind))
    kinds = {C.c_void_p: "ptr", C.c_int64: "int64", C.c_int: "int"}
    fields = [(n, kinds[t]) for n, t in api.WindOptions._fields_]
    assert fields == members


def test_noalloc_argument_errors_before_device():
    """wind_integrate_noalloc rejects what would need an allocation before
    any device call: host pointers, kernel timing, more systems than its
    static scratch holds."""
    lib = api.load()
    net = api.Network.builtin(api.NET_NR_TEST)
    o = api.WindOptions()
    lib.wind_options_init(C.byref(o))
    y = np.zeros(4, np.float32)
    c = np.zeros(8, np.float32)
    rc = lib.wind_integrate_noalloc(net.handle, 0.0, 1.0, 1, c.ctypes.data, y.ctypes.data, 2, 2, 1e-3, 1e-3,
                                    C.byref(o))
    assert rc == -1 and b"device pointers" in lib.wind_last_error()
    o.device_pointers = 1
    kms = C.c_float(0.0)
    o.kernel_ms = C.cast(C.byref(kms), C.c_void_p)
    rc = lib.wind_integrate_noalloc(net.handle, 0.0, 1.0, 1, c.ctypes.data, y.ctypes.data, 2, 2, 1e-3, 1e-3,
                                    C.byref(o))
    assert rc == -1 and b"kernel_ms" in lib.wind_last_error()
    o.kernel_ms = None
    text = open(os.path.join(ROOT, "include", "wind.h")).read()
    limit = eval(re.search(r"#define WIND_NOALLOC_MAX_SYSTEMS \((.*?)\)", text).group(1))
    rc = lib.wind_integrate_noalloc(net.handle, 0.0, 1.0, 1, c.ctypes.data, y.ctypes.data, limit + 1, 2, 1e-3,
                                    1e-3, C.byref(o))
    assert rc == -1 and b"exceeds its scratch maximum" in lib.wind_last_error()
    rc = lib.wind_integrate_noalloc(net.handle, 0.0, 1.0, 1, c.ctypes.data, y.ctypes.data, 2, 3, 1e-3, 1e-3,
                                    C.byref(o))
    assert rc == -1 and b"does not match the network" in lib.wind_last_error()
    # nothing to integrate: no device call
    assert lib.wind_integrate_noalloc(net.handle, 1.0, 1.0, 1, c.ctypes.data, y.ctypes.data, 2, 2, 1e-3, 1e-3,
                                      C.byref(o)) == 0
