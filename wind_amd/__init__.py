"""wind_amd: batched stiff-ODE integration (SIE / RK2 step doubling) on MI355X.

Native code lives in wind_amd/lib (built by ``python -m wind_amd.build``):
the core library with the C-ABI of include/wind.h, the per-network drop-in
libraries that WIND's own Python frontend loads (python/wind_harness.py), and
the generated network plugins.  ``wind_amd.api`` binds the C-ABI with ctypes,
``wind_amd.codegen`` compiles network descriptions, and ``wind_amd.networks``
holds the benchmark workloads.
"""
__version__ = "0.2.0"
