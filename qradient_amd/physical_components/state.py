"""Device-resident statevector with the reference's gate methods.

Mirrors qradient/physical_components/state.py:8-332 plus the State methods the
reference's circuits call but its State lacks (multiply_matrix,
exp_ham_classical, x_summed, ham_classical, reset(ini)): every method runs a
HIP kernel on the GPU through the C ABI.  ``vec`` reads or writes the whole
vector as a host numpy complex128 array (one device copy each way).
"""
import ctypes

import numpy as np

from .. import _native
from .observable import Observable

_AX = {'x': 0, 'y': 1, 'z': 2}


class State:
    """``State(qubit_number, ini='0')`` (state.py:34-37); extra keywords pick the
    storage precision (``dtype='complex128'|'complex64'``) and the device."""

    def __init__(self, qubit_number, ini='0', dtype='complex128', device=0):
        if ini not in ('0', '+'):
            raise ValueError('Invalid initialization format {}.'.format(ini))
        self.qnum = int(qubit_number)
        self._ini = ini
        self.dtype = dtype
        self._ctx = _native.Context.default(device)
        h = ctypes.c_void_p()
        _native.call('qg_state_create', self._ctx.handle, self.qnum, _native.DTYPES[dtype],
                     ini.encode(), ctypes.byref(h))
        self._h = h
        self.gates = None

    def __del__(self):
        h = getattr(self, '_h', None)
        if h and _native._lib is not None:
            _native._lib.qg_state_destroy(h)
            self._h = None

    # ---- vector access ------------------------------------------------------
    @property
    def vec(self):
        out = np.empty(2 ** self.qnum, dtype=np.complex128)
        _native.call('qg_state_get', self._h, out.view(np.float64).ctypes.data_as(_native._dp))
        return out

    @vec.setter
    def vec(self, value):
        v = np.ascontiguousarray(value, dtype=np.complex128)
        if v.shape != (2 ** self.qnum,):
            raise ValueError('state vector must have shape ({},)'.format(2 ** self.qnum))
        _native.call('qg_state_set', self._h, v.view(np.float64).ctypes.data_as(_native._dp))

    def copy_from(self, other):
        """vec <- other.vec on the device (same qubit number and dtype)."""
        _native.call('qg_state_copy', self._h, other._h)
        return self

    def sample_basis(self, uniforms, order=None):
        """Basis positions drawn at the given uniforms: the inverse CDF of
        |vec|^2 over the positions of ``order`` (natural order if None), as
        scipy's rv_discrete(values=(arange(2^n), dist)).rvs maps its uniforms
        (qaoa.py:196-198).  Returns int32 positions."""
        u = np.ascontiguousarray(uniforms, dtype=np.float64).ravel()
        idx = np.empty(u.size, dtype=np.int32)
        o = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
        if o is not None and o.shape != (2 ** self.qnum,):
            raise ValueError('order must have shape ({},)'.format(2 ** self.qnum))
        if u.size:
            _native.call('qg_state_sample_basis', self._h, None if o is None else _native.iptr(o), int(u.size),
                         _native.dptr(u), _native.iptr(idx))
        return idx

    def reset(self, ini=None):
        """state.py:61-71; ``reset('+')`` also changes the stored initial state (qaoa.py:17)."""
        if ini is not None:
            if ini not in ('0', '+'):
                raise ValueError('Invalid initialization format {}.'.format(ini))
            self._ini = ini
        _native.call('qg_state_reset', self._h, self._ini.encode())

    # ---- rotations (state.py:90-97, 142-149, 168-175) ------------------------
    def _rot(self, axis, angle, i, deriv):
        _native.call('qg_state_rot', self._h, _AX[axis], float(angle), int(i), int(deriv))

    def xrot(self, angle, i):
        self._rot('x', angle, i, 0)

    def yrot(self, angle, i):
        self._rot('y', angle, i, 0)

    def zrot(self, angle, i):
        self._rot('z', angle, i, 0)

    def dxrot(self, angle, i):
        self._rot('x', angle, i, 1)

    def dyrot(self, angle, i):
        self._rot('y', angle, i, 1)

    def dzrot(self, angle, i):
        self._rot('z', angle, i, 1)

    def xrot_all(self, angle=None, *args):
        """With an angle: RX(angle) on every qubit (Qaoa.__xrot_all, qaoa.py:192-194).
        Without: multiply by .5 * sum_q (-i X_q) (state.py:121-123)."""
        if angle is None:
            self.x_summed()
        else:
            _native.call('qg_state_rot_all', self._h, 0, float(angle))

    def x_summed(self):
        """vec <- .5 * sum_q (-i X_q) vec (state.py:107-123; qaoa.py:62)."""
        arr, k = _native.terms_array([(_native.TERM_X, q, -1, 0.5) for q in range(self.qnum)])
        _native.call('qg_state_multiply_observable', self._h, arr, k)
        _native.call('qg_state_scale', self._h, 0.0, -1.0)

    # ---- entanglers (state.py:198-199, 243-251) ------------------------------
    def cnot(self, i, j):
        _native.call('qg_state_cnot', self._h, int(i), int(j))

    def cnot_ladder(self, stacking):
        _native.call('qg_state_cnot_ladder', self._h, int(stacking))

    def cz_ladder(self):
        _native.call('qg_state_cz_ladder', self._h)

    # ---- observables and the classical Hamiltonian ---------------------------
    def multiply_matrix(self, matrix):
        """vec <- O vec (mc_clean.py:65).  ``matrix`` is an Observable, its dict,
        or the ``matrix`` attribute of an Observable."""
        obs = matrix if isinstance(matrix, Observable) else (
            Observable(matrix) if isinstance(matrix, dict) else Observable.from_matrix(matrix))
        if obs is None:
            raise TypeError('multiply_matrix takes an Observable (or its .matrix / dict) on the GPU path')
        arr, k = obs.c_terms()
        _native.call('qg_state_multiply_observable', self._h, arr, k)

    def expec_val(self, observable):
        arr, k = observable.c_terms()
        out = ctypes.c_double()
        _native.call('qg_state_expec_val', self._h, arr, k, ctypes.byref(out))
        return out.value

    def term_probs(self, observable):
        """|P_k^+ vec|^2 for each term of the observable, in its term order (the
        projector probabilities of base.py:22-46; FP64 reductions on the device)."""
        arr, k = observable.c_terms()
        out = np.empty(k)
        if k:
            _native.call('qg_state_term_prob', self._h, arr, k, _native.dptr(out))
        return out

    def _classical(self):
        if self.gates is None or getattr(self.gates, '_classical_terms', None) is None:
            raise AttributeError('classical Hamiltonian not loaded: use Gates(n).add_classical_ham(observable)')
        return self.gates._classical_terms

    def exp_ham_classical(self, angle):
        """vec <- exp(-i angle C) vec (qaoa.py:33,51; HEAD rot_classical_ham, state.py:299-301)."""
        arr, k = _native.terms_array(self._classical())
        _native.call('qg_state_exp_diag', self._h, arr, k, float(angle))

    rot_classical_ham = exp_ham_classical

    def exp_ham_classical_component(self, angle, i):
        """vec <- exp(-i angle C_i) vec for the i-th component (qaoa.py:141)."""
        arr, k = _native.terms_array([self._classical()[i]])
        _native.call('qg_state_exp_diag', self._h, arr, k, float(angle))

    rot_classical_ham_component = exp_ham_classical_component

    def ham_classical(self):
        """vec <- -i C vec (qaoa.py:67; HEAD classical_ham(), state.py:319-321)."""
        arr, k = _native.terms_array(self._classical())
        _native.call('qg_state_multiply_observable', self._h, arr, k)
        _native.call('qg_state_scale', self._h, 0.0, -1.0)

    def norm_error(self):
        """1 - ||vec|| (state.py:331-332), FP64 reduction on the device."""
        out = ctypes.c_double()
        _native.call('qg_state_norm_error', self._h, ctypes.byref(out))
        return out.value
