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


def draw_uniforms(count, shot_num, random_state=None):
    """[count, shot_num] uniforms: the numbers ``count`` successive
    rv_discrete.rvs calls draw (one ``uniform(size=shot_num)`` each; the
    legacy RandomState fills one larger request in the same order)."""
    rs = np.random.mtrand._rand if random_state is None else random_state
    return rs.uniform(size=(int(count), int(shot_num)))


def projector_estimates(probs, weights, uniforms):
    """sample_projectors for many states at once: probs [M, K], uniforms
    [M, K, shots] in draw order; returns the M estimates, each summed over
    its K terms in order (the reference's accumulation, base.py:27-33)."""
    probs = np.clip(np.asarray(probs, dtype=np.float64), 0.0, 1.0)
    q1 = probs + (1. - probs)   # np.cumsum([p, 1 - p])[1]
    u = np.asarray(uniforms)
    first = u <= probs[..., None]
    second = u <= q1[..., None]
    w = np.asarray(weights, dtype=np.float64)[None, :, None]
    vals = np.where(first | ~second, w, -w)
    means = vals.mean(axis=-1)
    total = np.zeros(means.shape[0])
    for k in range(means.shape[1]):
        total = total + means[:, k]
    return total


def sample_projectors(probs, weights, shot_num, random_state=None, uniforms=None):
    """Finite-shot estimate of sum_k w_k <P_k> (base.py:22-34).

    For each term in order: shot_num draws of the two-point distribution
    P(+w) = prob, P(-w) = 1 - prob, averaged, summed over the terms.  The draws
    are those of scipy's ``rv_discrete(values=([0, 1], [prob, 1 - prob])).rvs``
    that the reference makes: one uniform per shot from the legacy global
    numpy RandomState (or ``random_state``), outcome 0 where it is at most the
    first cumulative probability.  So a run seeded like the reference's gives
    the reference's estimate.  A probability that rounding puts just outside
    [0, 1] is clipped (the reference's rv_discrete would raise).
    """
    rs = np.random.mtrand._rand if random_state is None else random_state
    shot_num = int(shot_num)
    if shot_num < 1:
        raise ValueError('shot_num must be at least 1')
    total = 0.
    for k, (prob, w) in enumerate(zip(probs, weights)):
        prob = min(max(float(prob), 0.0), 1.0)
        qvals = np.cumsum([prob, 1. - prob])
        u = rs.uniform(size=shot_num) if uniforms is None else uniforms[k]
        idx = np.where(u <= qvals[0], 0, np.where(u <= qvals[1], 1, 0))
        total += np.array([w, -w])[idx].mean()
    return total


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
        self._plan_state = False  # the last run_expec_val's final state is the current state
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
    # After run_expec_val the final state stays in the plan's workspace; it is
    # copied into ``state`` only when ``state`` is read (the reference's
    # run_expec_val leaves it in self.state, mc_clean.py:27-45).
    @property
    def state(self):
        if self._state is None:
            self._state = State(self.qnum, dtype=self.dtype, device=self.device)
            self._state.gates = self._gates
            self._init_state(self._state)
        if self._plan_state:
            self._plan_state = False
            out = np.empty(2 ** self.qnum, dtype=np.complex128)
            _native.call('qg_plan_get_state', self._plan, out.view(np.float64).ctypes.data_as(_native._dp))
            self._state.vec = out
        return self._state

    @state.setter
    def state(self, value):
        self._plan_state = False
        self._state = value

    def _init_state(self, st):
        pass

    def context(self):
        if self._ctx is None:
            self._ctx = _native.Context.default(self.device)
        return self._ctx

    def expec_val(self):
        """<vec|O|vec> of the current state (base.py:17-20), FP64 on the device."""
        if self._plan_state:
            # the last run's final state is still in the plan's workspace
            arr, k = self.observable.c_terms()
            E = ctypes.c_double()
            _native.call('qg_plan_expec_val', self._plan, arr, k, ctypes.byref(E))
            return np.float64(E.value)
        return self.state.expec_val(self.observable)

    def _term_probs(self, terms):
        """Projector probabilities of the terms in the current state."""
        arr, k = _native.terms_array(terms)
        out = np.empty(k)
        if k:
            if self._plan_state:
                _native.call('qg_plan_term_prob', self._plan, arr, k, _native.dptr(out))
            else:
                _native.call('qg_state_term_prob', self.state._h, arr, k, _native.dptr(out))
        return out

    def sample_expec_val(self, shot_num, random_state=None):
        """Finite-shot expectation value of the observable in the current state
        (base.py:22-34): projector probabilities on the device, shot_num shots
        per term drawn as by the reference."""
        terms = self.observable.terms
        return sample_projectors(self._term_probs(terms), [t[3] for t in terms], shot_num, random_state)

    def sample_component_expec_val(self, shot_num, component, random_state=None):
        """Finite-shot expectation value of one observable term (base.py:36-46)."""
        t = self.observable.terms[component]
        return sample_projectors(self._term_probs([t]), [t[3]], shot_num, random_state)

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
