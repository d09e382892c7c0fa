"""McClean circuit on the GPU (reference qradient/circuit_logic/mc_clean.py:8-78).

RY(pi/4) on every qubit, then per layer the CNOT ladder (odd bonds first,
state.py:229-241) and one rotation per qubit about axes[i, q] in {0: X, 1: Y,
2: Z} by angles[i, q].  ``grad_run`` returns (E, grad[L, n]) with
grad[i, q] = dE/d angles[i, q], computed by the adjoint sweep in HIP kernels.

Backend keywords (not in the reference): ``dtype='complex128'`` (the plain
method, FP64 throughout) or ``'complex64'`` (complex64 state, FP64 arithmetic
and reductions); ``entangler='cnot'`` (reference) or ``'cz'``; ``device``;
``ckpt_every``; ``tune=True`` (apply the autotuner's cached tile shape for
this circuit shape, timing the candidates on the device first if none is
cached); ``graph=True`` (replay the pass launches as a HIP graph; for callers
that make no concurrent device-synchronising HIP calls from other threads).
"""
import ctypes

import numpy as np

from .. import _native
from ..physical_components import Gates
from .base import ParametrizedCircuit


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
            h = ctypes.c_void_p()
            _native.call('qg_mcclean_plan', self.context().handle, self.qnum, self.lnum,
                         _native.DTYPES[self.dtype], _native.ENTANGLERS[self.entangler], self.ckpt_every,
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

    def final_state(self):
        """The state after the last forward run (complex128 host copy)."""
        out = np.empty(2 ** self.qnum, dtype=np.complex128)
        _native.call('qg_plan_get_state', self._get_plan(), out.view(np.float64).ctypes.data_as(_native._dp))
        return out
