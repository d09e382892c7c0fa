"""QAOA circuit on the GPU (reference qradient/circuit_logic/qaoa.py:8-70).

|+>^n, then per layer exp(-i gamma_i C) and RX(beta_i) on every qubit, with C
the z / zz part of the observable (state.py:261-297).  ``grad_run(betas,
gammas)`` returns (E, grad[p, 2]), column 0 = dE/d beta_i, column 1 =
dE/d gamma_i.  Unlike the reference, an ``ini_state`` array is copied, never
mutated.
"""
import ctypes
import warnings

import numpy as np

from .. import _native
from ..physical_components import Gates
from .base import ParametrizedCircuit


class Qaoa(ParametrizedCircuit):
    def __init__(self, qubit_number, observable, layer_number, **kwargs):
        backend = ParametrizedCircuit.split_backend(kwargs)
        if kwargs:
            raise TypeError('unexpected keyword arguments: {}'.format(sorted(kwargs)))
        ParametrizedCircuit.init(self, qubit_number, observable, False, backend)
        self._gates = Gates(self.qnum).add_xrots().add_x_summed() \
            .add_classical_ham(self.observable, include_individual_components=True)
        self.lnum = int(layer_number)
        self.state_history = None
        self.has_loaded_eigensystem = False

    def _init_state(self, st):
        st.reset('+')  # uniform superposition (qaoa.py:17)

    def __check_parameters(self, betas, gammas):
        """qaoa.py:186-191."""
        betas = np.asarray(betas, dtype=np.float64)
        gammas = np.asarray(gammas, dtype=np.float64)
        if (betas.size != self.lnum) or (gammas.size != self.lnum):
            raise ValueError(('Wrong amount of parameters. Expected {} and {},'.format(self.lnum, self.lnum),
                              ' found {} and {}.'.format(betas.size, gammas.size)))
        return np.ascontiguousarray(betas.ravel()), np.ascontiguousarray(gammas.ravel())

    def _get_plan(self):
        if self._plan is None:
            h = ctypes.c_void_p()
            _native.call('qg_qaoa_plan', self.context().handle, self.qnum, self.lnum,
                         _native.DTYPES[self.dtype], 1, ctypes.byref(h))
            self._plan = h
        return self._plan

    def _terms(self):
        terms = self.observable.classical_terms()
        arr, k = _native.terms_array(terms)
        off = np.array([0, k], dtype=np.int32)
        return arr, off

    def grad_run(self, betas, gammas, hide_progbar=True, ini_state=None):
        """(E, grad[p, 2]): qaoa.py:40-70."""
        betas, gammas = self.__check_parameters(betas, gammas)
        arr, off = self._terms()
        E = np.empty(1)
        grad = np.empty((self.lnum, 2), dtype='double')
        if ini_state is None:
            _native.call('qg_qaoa_grad_run', self._get_plan(), 1, arr, _native.iptr(off), _native.dptr(betas),
                         _native.dptr(gammas), _native.dptr(E), _native.dptr(grad))
        else:
            # the library only reads ini_state (the reference scales it in place)
            _keep, ini = self._ini(ini_state, self.qnum)
            _native.call('qg_qaoa_grad_run_ini', self._get_plan(), 1, arr, _native.iptr(off), _native.dptr(betas),
                         _native.dptr(gammas), ini, _native.dptr(E), _native.dptr(grad))
        self._plan_state = False
        return np.float64(E[0]), grad

    def sample_grad(self, betas, gammas, shot_num=1, hide_progbar=True, exact_expec_val=True, ini_state=None):
        """Not implemented in the reference either (qaoa.py:72-81): warns and
        returns None."""
        warnings.warn('Not implemented yet.')

    def run_expec_val(self, betas, gammas, hide_progbar=True, exact_expec_val=True, shot_num=1, ini_state=None):
        """Forward circuit, then the exact expectation or, with
        exact_expec_val=False, its shot_num-shot estimate (qaoa.py:23-38)."""
        betas, gammas = self.__check_parameters(betas, gammas)
        arr, off = self._terms()
        E = np.empty(1)
        if ini_state is None:
            _native.call('qg_qaoa_expec_val', self._get_plan(), 1, arr, _native.iptr(off), _native.dptr(betas),
                         _native.dptr(gammas), _native.dptr(E))
        else:
            _keep, ini = self._ini(ini_state, self.qnum)
            _native.call('qg_qaoa_expec_val_ini', self._get_plan(), 1, arr, _native.iptr(off), _native.dptr(betas),
                         _native.dptr(gammas), ini, _native.dptr(E))
        self._plan_state = True
        if not exact_expec_val:
            return self.sample_expec_val(shot_num)
        return np.float64(E[0])
