"""ParametrizedCircuit base (reference qradient/circuit_logic/base.py:7-58)."""
import ctypes

import numpy as np

from .. import _native
from ..physical_components import Observable, State

_BACKEND_KEYS = ('dtype', 'entangler', 'device', 'ckpt_every', 'tune', 'graph')


def progbar_range(hide_progbar):
    """base.py:51-58.  The GPU path runs every layer in one launch sequence, so
    there is nothing to report per layer; kept for callers that use it."""
    if hide_progbar:
        return np.arange
    try:
        from tqdm import trange
        return trange
    except ImportError:  # pragma: no cover
        return np.arange


class ParametrizedCircuit:
    """Parent class for VQE circuits. Not meant for instantiation."""

    def init(self, qubit_number, observable, use_observable_components=False, backend=None):
        backend = dict(backend or {})
        self.qnum = int(qubit_number)
        self.dtype = backend.get('dtype', 'complex128')
        if self.dtype not in _native.DTYPES:
            raise ValueError('dtype must be complex128 or complex64')
        self.entangler = backend.get('entangler', 'cnot')
        if self.entangler not in _native.ENTANGLERS:
            raise ValueError("entangler must be 'cnot' or 'cz'")
        self.device = int(backend.get('device', 0))
        self.ckpt_every = int(backend.get('ckpt_every', 0))
        self.tune = bool(backend.get('tune', False))
        self.graph = bool(backend.get('graph', False))
        self.observable = Observable(self.qnum, observable, store_components=use_observable_components)
        self._state = None
        self._gates = None
        self._plan = None
        self._ctx = None
        self.has_loaded_projectors = False

    @staticmethod
    def split_backend(kwargs):
        return {k: kwargs.pop(k) for k in list(kwargs) if k in _BACKEND_KEYS}

    # The reference allocates State(n) eagerly (base.py:11); here the device
    # state is created on first access so grad_run (which works in plan
    # workspaces) does not hold an unused 2^n vector.
    @property
    def state(self):
        if self._state is None:
            self._state = State(self.qnum, dtype=self.dtype, device=self.device)
            self._state.gates = self._gates
            self._init_state(self._state)
        return self._state

    @state.setter
    def state(self, value):
        self._state = value

    def _init_state(self, st):
        pass

    def context(self):
        if self._ctx is None:
            self._ctx = _native.Context.default(self.device)
        return self._ctx

    def expec_val(self):
        """<vec|O|vec> of the current state (base.py:17-20), FP64 on the device."""
        return self.state.expec_val(self.observable)

    def sample_expec_val(self, shot_num):
        raise NotImplementedError('finite-shot sampling is outside the GPU grad_run path')

    def _destroy_plan(self):
        if self._plan is not None and _native._lib is not None:
            _native._lib.qg_plan_destroy(self._plan)
        self._plan = None

    def __del__(self):
        try:
            self._destroy_plan()
        except Exception:  # pragma: no cover - interpreter shutdown
            pass

    @staticmethod
    def _ini(ini_state, n):
        if ini_state is None:
            return None, None
        v = np.ascontiguousarray(ini_state, dtype=np.complex128)
        if v.shape != (2 ** n,):
            raise ValueError('ini_state must have shape ({},)'.format(2 ** n))
        return v, v.view(np.float64).ctypes.data_as(_native._dp)

    def last_device_ms(self):
        """Device time of the last grad_run / run_expec_val launch sequence."""
        ms = ctypes.c_double()
        _native.call('qg_plan_last_ms', self._plan, ctypes.byref(ms))
        return ms.value
