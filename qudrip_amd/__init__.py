"""MI355X-native sparse-Hamiltonian Krylov path of QuDrip.

The product is ``libqudrip_hip.so`` (HIP kernels for gfx950 behind the C ABI
in include/qudrip_hip.h) and the C++ drop-in headers in include/qudrip/.
This Python package binds the same C ABI for tests and benchmarks:
``operators`` mirrors the reference's operator interface (term programs),
``runtime`` owns the library handles.
"""
from . import abi  # noqa: F401
from . import operators  # noqa: F401

__all__ = ["abi", "operators"]
