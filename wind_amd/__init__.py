"""wind_amd: batched stiff-ODE integration (SIE / RK2 step doubling) on MI355X.

Native code lives in wind_amd/lib (built by ``python -m wind_amd.build``);
``wind_amd.api`` binds the C-ABI of include/wind.h and ``wind_amd.harness``
mirrors the reference's Python calling convention (python/wind_harness.py,
python/ode_systems/ode_base.py).
"""
__version__ = "0.1.0"
