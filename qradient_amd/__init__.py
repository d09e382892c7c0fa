"""MI355X-native drop-in for the statevector grad_run path of qradient.

Same import surface as the reference (qradient/__init__.py:1-3,
qradient/physical_components/__init__.py:1-2, qradient/circuit_logic/__init__.py:1-2)::

    from qradient_amd.circuit_logic import McClean, Qaoa
    from qradient_amd.physical_components import State, Gates, Observable

All statevector arithmetic runs in HIP kernels on gfx950 through the C ABI in
include/qradient_amd.h (libqradient_amd.so, bound with ctypes).
"""
from . import physical_components, circuit_logic, optimization_problems, optimization  # noqa: F401

__all__ = ['physical_components', 'circuit_logic', 'optimization_problems', 'optimization']
