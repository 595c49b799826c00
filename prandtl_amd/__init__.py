"""prandtl_amd -- MI355X (gfx950) DGSEM right-hand side and RK stage behind a C ABI.

The product is ``libprandtl_amd.so`` (HIP kernels + C ABI, see include/prandtl_amd.h).
This package holds its sources (``csrc/``), its build script, the ctypes binding
used by tests and the benchmark (``capi``; the library is loaded on first use), and
the synthetic workloads (``mesh``, ``cases``).
"""
from . import capi, cases, mesh  # noqa: F401

__all__ = ["mesh", "cases", "capi"]
