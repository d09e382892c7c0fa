"""State, Gates and Observable (reference qradient/physical_components/__init__.py:1-2,
plus the Gates builder that the reference's circuits import but never define)."""
from .state import State
from .observable import Observable
from .gates import Gates

__all__ = ['State', 'Observable', 'Gates']
