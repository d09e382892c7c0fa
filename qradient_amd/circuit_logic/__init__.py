"""Circuit classes (reference qradient/circuit_logic/__init__.py:1-2)."""
from .mc_clean import McClean
from .qaoa import Qaoa

__all__ = ['McClean', 'Qaoa']
