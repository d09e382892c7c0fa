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
from .base import ParametrizedCircuit, sample_projectors


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

    def sample_grad_dense(self, betas, gammas, shot_num=1, hide_progbar=True, exact_expec_val=True,
                          ini_state=None):
        """(E, grad[p, 2]) from finite-shot measurements of shifted circuits
        in the computational basis (qaoa.py:83-158).

        For layer i: each classical-Hamiltonian component C_j phase-shifted by
        exp(-+i pi/4 C_j) after the cost phase gives grad[i, 1] = sum_j
        (o_plus_j - o_minus_j); each qubit's mixer rotation shifted by +-pi/2
        gives grad[i, 0] = sum_q (o_plus_q - o_minus_q) / 2.  An estimate o is
        the mean of C(x) over shot_num basis draws of the shifted circuit's
        final state.  The reference carries each suffix circuit as a dense
        left-hand-side matrix; here the shifted states run the suffix on the
        device from device copies of the forward history, and are sampled
        there by the inverse CDF at uniforms drawn on the host in the
        reference's order.  E is <C> (or, with exact_expec_val=False, the
        shot estimate of the whole observable).  The reference mutates
        ``ini_state``; it is only read here."""
        from ..physical_components import State
        betas, gammas = self.__check_parameters(betas, gammas)
        n, L = self.qnum, self.lnum
        cterms = self.observable.classical_terms()
        arr, k = _native.terms_array(cterms)
        C = self._classical_values(cterms)

        def new_state():
            st = State(n, '+', dtype=self.dtype, device=self.device)
            st.gates = self._gates
            return st

        st = new_state()
        if ini_state is not None:
            st.vec = ini_state
        hist = []
        for i in range(L):
            st.exp_ham_classical(gammas[i])
            hist.append(new_state().copy_from(st))
            st.xrot_all(betas[i])
            hist.append(new_state().copy_from(st))
        terms = self.observable.terms
        if exact_expec_val:
            Eo = ctypes.c_double()
            _native.call('qg_state_expec_val', st._h, arr, k, ctypes.byref(Eo))
            E = np.float64(Eo.value)
        else:
            E = sample_projectors(st.term_probs(self.observable), [t[3] for t in terms], shot_num)
        work = new_state()
        shots = int(shot_num)

        def estimate(first_layer, mixer_first):
            """Runs the suffix from ``work`` and returns the shot mean of C."""
            if mixer_first:
                work.xrot_all(betas[first_layer - 1])
            for j in range(first_layer, L):
                work.exp_ham_classical(gammas[j])
                work.xrot_all(betas[j])
            u = np.random.uniform(size=shots)
            return C[work.sample_basis(u)].mean()

        grad = np.empty((L, 2), dtype='double')
        for i in range(L):
            dg = 0.
            for j in range(len(cterms)):
                out = []
                for angle in (np.pi / 4, -np.pi / 4):
                    work.copy_from(hist[2 * i])
                    work.exp_ham_classical_component(angle, j)
                    out.append(estimate(i + 1, True))
                dg += out[0] - out[1]
            db = 0.
            for q in range(n):
                out = []
                for angle in (np.pi / 2, -np.pi / 2):
                    work.copy_from(hist[2 * i + 1])
                    work.xrot(angle, q)
                    out.append(estimate(i + 1, False))
                db += out[0] - out[1]
            grad[i, 0] = .5 * db
            grad[i, 1] = dg
        return E, grad

    def _classical_values(self, cterms):
        """C(x) in natural order, summed in term order (state.py:273-294)."""
        n = self.qnum
        x = np.arange(2 ** n, dtype=np.int64)
        C = np.zeros(2 ** n)
        for kind, i, j, w in cterms:
            s = 1.0 - 2.0 * ((x >> (n - 1 - i)) & 1)
            if kind == _native.TERM_ZZ:
                s = s * (1.0 - 2.0 * ((x >> (n - 1 - j)) & 1))
            C = C + w * s
        return C

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
