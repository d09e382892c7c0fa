"""McClean circuit on the GPU (reference qradient/circuit_logic/mc_clean.py:8-78).

RY(pi/4) on every qubit, then per layer the CNOT ladder (odd bonds first,
state.py:229-241) and one rotation per qubit about axes[i, q] in {0: X, 1: Y,
2: Z} by angles[i, q].  ``grad_run`` returns (E, grad[L, n]) with
grad[i, q] = dE/d angles[i, q], computed by the adjoint sweep in HIP kernels.

Backend keywords (not in the reference): ``dtype='complex128'`` (the plain
method, FP64 throughout) or ``'complex64'`` (complex64 state, FP64 arithmetic
and reductions); ``entangler='cnot'`` (reference) or ``'cz'``; ``device``;
``ckpt_every``; ``tune=True`` (apply the autotuner's cached tile shape and,
for circuits whose per-pass history does not fit, checkpoint period for this
circuit shape, timing the candidates on the device first if none is cached); ``graph=True`` (replay the pass launches as a HIP graph; for callers
that make no concurrent device-synchronising HIP calls from other threads).
"""
import ctypes
import warnings

import numpy as np

from .. import _native
from ..physical_components import Gates
from .base import ParametrizedCircuit, draw_uniforms, projector_estimates, sample_projectors


class McClean(ParametrizedCircuit):
    def __init__(self, qubit_number, observable, layer_number, use_observable_components=False, **kwargs):
        backend = ParametrizedCircuit.split_backend(kwargs)
        ParametrizedCircuit.init(self, qubit_number, observable, use_observable_components, backend)
        self.lnum = int(layer_number)
        # both defaults are drawn from the global RNG, axes first, even when
        # the caller passes them (kwargs.get evaluates eagerly, mc_clean.py:13-14)
        default_axes = (3 * np.random.rand(self.lnum, self.qnum)).astype('int')
        default_angles = 2 * np.pi * np.random.rand(self.lnum, self.qnum)
        self.axes = kwargs.get('axes', default_axes)
        self.angles = kwargs.get('angles', default_angles)
        self._gates = Gates(self.qnum).add_xrots().add_yrots().add_zrots().add_cnot_ladder()
        # the history lives in device plan workspaces, never on the host
        self.state_history = None
        self.has_loaded_eigensystem = False
        self.has_loaded_component_eigensystems = False

    def _get_plan(self):
        if self._plan is None:
            ckpt = self.ckpt_every
            if self.tune and ckpt == 0:
                # the tuned checkpoint period of this shape, when its history does not fit
                from .. import autotune
                k = autotune.lookup(autotune.ckpt_key(self.qnum, self.lnum, self.dtype, self.entangler))
                if k is None:
                    cand = autotune.ckpt_candidates(self.context(), self.qnum, self.lnum, self.dtype)
                    k = cand[0] if cand == [0] else autotune.tune_ckpt(
                        self.qnum, self.lnum, self.dtype, self.entangler, device=self.device)[0]
                ckpt = int(k)
            h = ctypes.c_void_p()
            _native.call('qg_mcclean_plan', self.context().handle, self.qnum, self.lnum,
                         _native.DTYPES[self.dtype], _native.ENTANGLERS[self.entangler], ckpt,
                         ctypes.byref(h))
            self._plan = h
            if self.graph:
                _native.call('qg_plan_set_graph', h, 1)
            if self.tune:
                from .. import autotune
                shape = autotune.cached('mcclean', self.qnum, self.lnum, self.dtype, self.entangler)
                if shape is None:
                    shape, _ = autotune.tune_mcclean(self.qnum, self.lnum, self.dtype, self.entangler,
                                                     device=self.device)
                autotune.apply(h, shape)
        return self._plan

    def _params(self):
        axes = np.ascontiguousarray(self.axes, dtype=np.int64)
        angles = np.ascontiguousarray(self.angles, dtype=np.float64)
        if axes.shape != (self.lnum, self.qnum) or angles.shape != (self.lnum, self.qnum):
            raise ValueError('axes and angles must have shape ({}, {})'.format(self.lnum, self.qnum))
        bad = axes[(axes < 0) | (axes > 2)]
        if bad.size:
            raise ValueError('Invalid axis {}'.format(bad[0]))
        return np.ascontiguousarray(axes, dtype=np.int32), angles

    def grad_run(self, hide_progbar=True, ini_state=None):
        """(E, grad[L, n]): mc_clean.py:47-78."""
        axes, angles = self._params()
        _keep, ini = self._ini(ini_state, self.qnum)
        arr, k = self.observable.c_terms()
        E = ctypes.c_double()
        grad = np.empty((self.lnum, self.qnum), dtype='double')
        _native.call('qg_mcclean_grad_run', self._get_plan(), _native.iptr(axes), _native.dptr(angles), arr, k,
                     ini, ctypes.byref(E), _native.dptr(grad))
        self._plan_state = False
        return np.float64(E.value), grad

    def run_expec_val(self, hide_progbar=True, exact_expec_val=True, shot_num=1, ini_state=None):
        """Forward circuit, then the exact expectation or, with
        exact_expec_val=False, its shot_num-shot estimate (mc_clean.py:27-45)."""
        axes, angles = self._params()
        _keep, ini = self._ini(ini_state, self.qnum)
        # the shot estimate needs only the final state: its forward run takes
        # the z / zz terms, so x / y terms are sampled at every size
        arr, k = self.observable.c_terms() if exact_expec_val else _native.terms_array(
            self.observable.classical_terms())
        E = ctypes.c_double()
        _native.call('qg_mcclean_expec_val', self._get_plan(), _native.iptr(axes), _native.dptr(angles), arr, k,
                     ini, ctypes.byref(E))
        self._plan_state = True
        if not exact_expec_val:
            return self.sample_expec_val(shot_num)
        return np.float64(E.value)

    def grad_run_with_component_sampling(self, hide_progbar=True, ini_state=None):
        """(E, grad[L, n]) with the adjoint sweep seeded by one sampled
        observable component (mc_clean.py:80-115).

        E is the whole observable's expectation.  The gradient is that of
        <psi|w_k P_k|psi> for the term k drawn by
        ``np.random.choice(num_components, p=weight_distribution)`` from the
        global RNG, one draw per call as in the reference (which draws after
        the forward sweep; the sweep takes no random numbers, so the stream is
        the same).  Needs ``use_observable_components=True``.
        """
        obs = self.observable
        k = np.random.choice(np.arange(obs.num_components), p=obs.weight_distribution)
        axes, angles = self._params()
        _keep, ini = self._ini(ini_state, self.qnum)
        arr, nk = _native.terms_array([obs.terms[k]])
        E = ctypes.c_double()
        grad = np.empty((self.lnum, self.qnum), dtype='double')
        plan = self._get_plan()
        _native.call('qg_mcclean_grad_run', plan, _native.iptr(axes), _native.dptr(angles), arr, nk, ini,
                     ctypes.byref(E), _native.dptr(grad))
        self._plan_state = False
        full, nf = obs.c_terms()
        status = _native.load().qg_plan_expec_val(plan, full, nf, ctypes.byref(E))
        if status == _native.QG_ERR_INVALID:
            # a checkpointed adjoint sweep uncomputed the final state: run the
            # forward sweep again for the whole observable
            _native.call('qg_mcclean_expec_val', plan, _native.iptr(axes), _native.dptr(angles), full, nf, ini,
                         ctypes.byref(E))
        else:
            _native.check(status)
        return np.float64(E.value), grad

    def shifted_term_probs(self, shift_param, shift_delta, terms, ini_state=None):
        """Projector probabilities of terms (``observable.terms`` entries) in
        the final states of the McClean circuits with angles.flat[p] shifted
        by delta, one row per (p, delta) pair; afterwards ``state`` is the
        unshifted circuit's final state."""
        axes, angles = self._params()
        _keep, ini = self._ini(ini_state, self.qnum)
        sp_ = np.ascontiguousarray(shift_param, dtype=np.int32)
        sd = np.ascontiguousarray(shift_delta, dtype=np.float64)
        if sp_.shape != sd.shape or sp_.ndim != 1:
            raise ValueError('shift_param and shift_delta must be 1-d arrays of one length')
        arr, k = _native.terms_array(terms)
        probs = np.empty((sp_.size, k))
        _native.call('qg_mcclean_shift_term_prob', self._get_plan(), _native.iptr(axes), _native.dptr(angles), ini,
                     int(sp_.size), _native.iptr(sp_), _native.dptr(sd), arr, k, _native.dptr(probs))
        self._plan_state = True
        return probs

    def _shift_run(self, shot_num, exact_expec_val, ini_state):
        """Shared part of the parameter-shift methods: the unshifted circuit's
        E (exact, or its shot estimate), then the probabilities of every term
        in the +pi/2 and -pi/2 shifted circuit of every parameter, rows in
        the reference's loop order (layer, qubit, then +, -)."""
        if int(shot_num) < 1:
            raise ValueError('shot_num must be at least 1')
        npar = self.lnum * self.qnum
        params = np.repeat(np.arange(npar, dtype=np.int32), 2)
        deltas = np.tile(np.array([np.pi / 2, -np.pi / 2]), npar)
        terms = self.observable.terms
        probs = self.shifted_term_probs(params, deltas, terms, ini_state)
        if exact_expec_val:
            E = self.expec_val()
        else:
            E = self.sample_expec_val(shot_num)
        return E, probs, np.array([t[3] for t in terms])

    def sample_grad(self, hide_progbar=True, shot_num=1, exact_expec_val=True, ini_state=None):
        """(E, grad[L, n]) by the parameter-shift rule on finite-shot
        expectation values (mc_clean.py:117-155): grad[i, q] =
        (o_plus - o_minus) / 2 with o_+- the shot_num-shot estimates of the
        circuit with angles[i, q] shifted by +-pi/2.  The shifted circuits run
        on the device and give the projector probabilities of every term; the
        shots are drawn on the host from the global RNG in the reference's
        order (E's shots first when exact_expec_val is False, then per
        parameter o_plus's, then o_minus's), so a seeded call gives the
        reference's estimate."""
        E, probs, w = self._shift_run(shot_num, exact_expec_val, ini_state)
        npar = self.lnum * self.qnum
        # every shifted circuit's draws in the reference's order: per
        # parameter o_plus's terms, then o_minus's, shot_num each
        u = np.random.uniform(size=(2 * npar, len(w), int(shot_num))) if npar and len(w) else None
        o = projector_estimates(probs, w, u) if u is not None else np.zeros(2 * npar)
        grad = (.5 * (o[0::2] - o[1::2])).reshape(self.lnum, self.qnum)
        return E, grad

    def sample_grad_with_component_sampling(self, hide_progbar=True, shot_num=1, exact_expec_val=True,
                                            ini_state=None):
        """As ``sample_grad``, but each parameter's two estimates measure one
        observable component, drawn per parameter with
        ``np.random.choice(num_components, p=weight_distribution)`` before its
        shots (mc_clean.py:157-198).  Needs ``use_observable_components=True``."""
        obs = self.observable
        ncomp, dist = obs.num_components, obs.weight_distribution
        E, probs, w = self._shift_run(shot_num, exact_expec_val, ini_state)
        grad = np.empty((self.lnum, self.qnum), dtype='double')
        for k in range(self.lnum * self.qnum):
            c = np.random.choice(np.arange(ncomp), p=dist)
            o_plus = sample_projectors(probs[2 * k, c:c + 1], w[c:c + 1], shot_num)
            o_minus = sample_projectors(probs[2 * k + 1, c:c + 1], w[c:c + 1], shot_num)
            grad.flat[k] = .5 * (o_plus - o_minus)
        return E, grad

    def _diagonal_values(self):
        """D(x) of a diagonal observable, summed in term order as the
        reference's observable matrix sums its terms (observable.py:44-73);
        np.linalg.eigh of that matrix returns np.sort(D) exactly."""
        n = self.qnum
        x = np.arange(2 ** n, dtype=np.int64)
        D = np.zeros(2 ** n)
        for kind, i, j, w in self.observable.terms:
            s = 1.0 - 2.0 * ((x >> (n - 1 - i)) & 1)
            if kind == _native.TERM_ZZ:
                s = s * (1.0 - 2.0 * ((x >> (n - 1 - j)) & 1))
            D = D + w * s
        return D

    def sample_grad_dense(self, shot_num=1, hide_progbar=True, exact_expec_val=True, ini_state=None):
        """(E, grad[L, n]) by the parameter-shift rule with each shifted
        circuit measured shot_num times in the observable's eigenbasis
        (mc_clean.py:207-268).

        The reference forms the eigenbasis with a dense eigh and carries the
        suffix circuit as dense 2^n x 2^n left-hand-side matrices.  For a
        diagonal (z / zz) observable the eigenbasis is the computational basis
        sorted by D(x) ascending, so here every shifted circuit runs on the
        device and is sampled there by the inverse CDF over that order at
        uniforms drawn on the host in the reference's order (E's shots
        first when exact_expec_val is False, then per parameter the +pi/2
        circuit's, then the -pi/2 circuit's).  Observables with x / y terms
        raise NotImplementedError."""
        if not self.observable.is_diagonal:
            raise NotImplementedError('sample_grad_dense: only diagonal (z / zz) observables on the GPU path')
        if int(shot_num) < 1:
            raise ValueError('shot_num must be at least 1')
        D = self._diagonal_values()
        order = np.argsort(D, kind='stable').astype(np.int32)
        eig = D[order]
        terms = self.observable.terms
        w = np.array([t[3] for t in terms])
        u_E = None if exact_expec_val else draw_uniforms(len(terms), shot_num)
        npar = self.lnum * self.qnum
        u = draw_uniforms(2 * npar, shot_num).ravel()
        axes, angles = self._params()
        _keep, ini = self._ini(ini_state, self.qnum)
        params = np.repeat(np.arange(npar, dtype=np.int32), 2)
        deltas = np.tile(np.array([np.pi / 2, -np.pi / 2]), npar)
        idx = np.empty(u.size, dtype=np.int32)
        _native.call('qg_mcclean_shift_sample', self._get_plan(), _native.iptr(axes), _native.dptr(angles), ini,
                     2 * npar, _native.iptr(params), _native.dptr(deltas), _native.iptr(order), int(shot_num),
                     _native.dptr(u), _native.iptr(idx))
        self._plan_state = True
        E = self.expec_val() if exact_expec_val else sample_projectors(self._term_probs(terms), w, shot_num,
                                                                        uniforms=u_E)
        means = eig[idx].reshape(2 * npar, int(shot_num)).mean(axis=1)
        grad = ((means[0::2] - means[1::2]) / 2.).reshape(self.lnum, self.qnum)
        return E, grad

    def sample_grad_dense_with_component_sampling(self, shot_num=1, hide_progbar=True, exact_expec_val=True,
                                                  ini_state=None):
        """As sample_grad_dense, measuring one sampled component w_k P_k in
        its eigenbasis (mc_clean.py:277-350); the component is drawn once
        per call, before E.  The eigenvalues of w_k P_k are -|w_k| and
        +|w_k|, each on half of the space, so a shot is -|w_k| when its
        uniform is at most the mass of the -|w_k| eigenspace (from the device's
        projector probability of P_k) and +|w_k| otherwise.  Needs
        ``use_observable_components=True``."""
        if int(shot_num) < 1:
            raise ValueError('shot_num must be at least 1')
        obs = self.observable
        c = np.random.choice(np.arange(obs.num_components), p=obs.weight_distribution)
        terms = obs.terms
        w = np.array([t[3] for t in terms])
        u_E = None if exact_expec_val else draw_uniforms(len(terms), shot_num)
        npar = self.lnum * self.qnum
        u = draw_uniforms(2 * npar, shot_num)
        params = np.repeat(np.arange(npar, dtype=np.int32), 2)
        deltas = np.tile(np.array([np.pi / 2, -np.pi / 2]), npar)
        p_plus = self.shifted_term_probs(params, deltas, [terms[c]], ini_state)[:, 0]
        E = self.expec_val() if exact_expec_val else sample_projectors(self._term_probs(terms), w, shot_num,
                                                                        uniforms=u_E)
        wc = abs(w[c])
        # mass of the -|w| eigenspace: P_k = -1 for w > 0, +1 for w < 0
        m_minus = 1.0 - p_plus if w[c] > 0 else p_plus
        means = np.array([np.where(uk <= m, -wc, wc).mean() for uk, m in zip(u, m_minus)])
        grad = ((means[0::2] - means[1::2]) / 2.).reshape(self.lnum, self.qnum)
        return E, grad

    def sample_grad_observable(self, *args):
        """Renamed in the reference (mc_clean.py:200-205): warns, returns None."""
        warnings.warn('Method sample_grad_observable is now called sample_grad_dense.', DeprecationWarning,
                      stacklevel=2)

    def sample_grad_observable_with_component_sampling(self, *args):
        """Renamed in the reference (mc_clean.py:270-275): warns, returns None."""
        warnings.warn('Method sample_grad_observable_with_component_sampling is now called '
                      'sample_grad_dense_with_component_sampling.', DeprecationWarning, stacklevel=2)

    def final_state(self):
        """The state after the last forward run (complex128 host copy)."""
        out = np.empty(2 ** self.qnum, dtype=np.complex128)
        _native.call('qg_plan_get_state', self._get_plan(), out.view(np.float64).ctypes.data_as(_native._dp))
        return out
