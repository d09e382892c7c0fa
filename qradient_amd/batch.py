"""Batched grad_run: many circuits of one shape in one launch sequence.

The reference evaluates one circuit per Python call (qradient/circuit_logic/
qaoa.py:40-70, mc_clean.py:47-78); here B circuits of the same shape share
every kernel launch (blockIdx.y selects the circuit), so small states still
fill the GPU.  ``QaoaBatch``: B QAOA circuits, each with its own graph and
angles.  ``McCleanBatch``: B parameter sets (axes, angles) of one McClean
shape and one observable, each result bitwise what a single McClean plan
gives it.
"""
import ctypes

import numpy as np

from . import _native
from .physical_components import Observable


class PackedObservables:
    """Classical Hamiltonians of a batch in the C-ABI term layout (see QaoaBatch.pack)."""

    def __init__(self, arr, off):
        self.arr, self.off = arr, off

    def __len__(self):
        return len(self.off) - 1


class QaoaBatch:
    """Plan for up to ``batch`` QAOA circuits of ``n`` qubits and ``p`` layers."""

    def __init__(self, n, p, batch, dtype='complex64', device=0, tune=False):
        """``tune=True`` applies the autotuner's batch shape for (n, p, batch,
        dtype) (tile shape, chunk, workgroups per CU), timing the candidates
        on the device first if none is cached; ``shape`` holds it."""
        self.n, self.p, self.batch, self.dtype = int(n), int(p), int(batch), dtype
        self._ctx = _native.Context.default(device)
        self.shape = None
        if tune:
            from . import autotune
            shape = autotune.lookup(autotune.qaoa_key(self.n, self.p, self.batch, dtype))
            if shape is None:
                shape, _ = autotune.tune_qaoa(self.n, self.p, self.batch, dtype, device=device)
            self.shape = shape
        h = ctypes.c_void_p()
        _native.call('qg_qaoa_plan', self._ctx.handle, self.n, self.p, _native.DTYPES[dtype], self.batch,
                     ctypes.byref(h))
        self._plan = h
        if self.shape is not None:
            from . import autotune
            autotune.apply_qaoa(h, self.shape)

    def __del__(self):
        h = getattr(self, '_plan', None)
        if h and _native._lib is not None:
            _native._lib.qg_plan_destroy(h)
            self._plan = None

    def pack(self, observables):
        """Pre-packs the classical Hamiltonians of a batch; the result can be passed to
        grad_run / expec_val in place of the observable list (an optimiser loop over fixed
        graphs then converts them once)."""
        arr, off = self._terms(observables)
        return PackedObservables(arr, off)

    def _terms(self, observables):
        if isinstance(observables, PackedObservables):
            return observables.arr, observables.off
        terms, off = [], [0]
        for o in observables:
            ob = o if isinstance(o, Observable) else Observable(o)
            if ob.qnum != self.n:
                raise ValueError('observable has {} qubits, plan has {}'.format(ob.qnum, self.n))
            terms.extend(ob.classical_terms())
            off.append(len(terms))
        arr, _ = _native.terms_array(terms)
        return arr, np.array(off, dtype=np.int32)

    def _params(self, betas, gammas, B):
        b = np.ascontiguousarray(betas, dtype=np.float64).reshape(B, self.p)
        g = np.ascontiguousarray(gammas, dtype=np.float64).reshape(B, self.p)
        return b, g

    def _ini(self, ini_states, B):
        """B initial states as one contiguous complex128 block (read only)."""
        v = np.ascontiguousarray(ini_states, dtype=np.complex128)
        if v.shape != (B, 2 ** self.n):
            raise ValueError('ini_states must have shape ({}, {})'.format(B, 2 ** self.n))
        return v, v.view(np.float64).ctypes.data_as(_native._dp)

    def grad_run(self, observables, betas, gammas, ini_states=None):
        """(E[B], grad[B, p, 2]) with grad[b, i] = (dE/dbeta_i, dE/dgamma_i) of circuit b.
        ini_states: optional (B, 2**n) initial states (default |+>^n per circuit)."""
        B = len(observables)
        if B > self.batch:
            raise ValueError('{} circuits, the plan holds {}'.format(B, self.batch))
        arr, off = self._terms(observables)
        b, g = self._params(betas, gammas, B)
        E = np.empty(B)
        grad = np.empty((B, self.p, 2))
        if ini_states is None:
            _native.call('qg_qaoa_grad_run', self._plan, B, arr, _native.iptr(off), _native.dptr(b),
                         _native.dptr(g), _native.dptr(E), _native.dptr(grad))
        else:
            _keep, ini = self._ini(ini_states, B)
            _native.call('qg_qaoa_grad_run_ini', self._plan, B, arr, _native.iptr(off), _native.dptr(b),
                         _native.dptr(g), ini, _native.dptr(E), _native.dptr(grad))
        return E, grad

    def expec_val(self, observables, betas, gammas, ini_states=None):
        B = len(observables)
        if B > self.batch:
            raise ValueError('{} circuits, the plan holds {}'.format(B, self.batch))
        arr, off = self._terms(observables)
        b, g = self._params(betas, gammas, B)
        E = np.empty(B)
        if ini_states is None:
            _native.call('qg_qaoa_expec_val', self._plan, B, arr, _native.iptr(off), _native.dptr(b),
                         _native.dptr(g), _native.dptr(E))
        else:
            _keep, ini = self._ini(ini_states, B)
            _native.call('qg_qaoa_expec_val_ini', self._plan, B, arr, _native.iptr(off), _native.dptr(b),
                         _native.dptr(g), ini, _native.dptr(E))
        return E

    def last_device_ms(self):
        ms = ctypes.c_double()
        _native.call('qg_plan_last_ms', self._plan, ctypes.byref(ms))
        return ms.value


class McCleanBatch:
    """Plan for up to ``batch`` McClean circuits of ``n`` qubits and ``L`` layers."""

    def __init__(self, n, L, batch, dtype='complex64', entangler='cnot', device=0):
        self.n, self.L, self.batch, self.dtype = int(n), int(L), int(batch), dtype
        self._ctx = _native.Context.default(device)
        h = ctypes.c_void_p()
        _native.call('qg_mcclean_batch_plan', self._ctx.handle, self.n, self.L, _native.DTYPES[dtype],
                     _native.ENTANGLERS[entangler], self.batch, ctypes.byref(h))
        self._plan = h

    def __del__(self):
        h = getattr(self, '_plan', None)
        if h and _native._lib is not None:
            _native._lib.qg_plan_destroy(h)
            self._plan = None

    def grad_run(self, axes, angles, observable):
        """axes, angles: [B, L, n]; one observable for all.  Returns (E[B], grad[B, L, n])."""
        axes = np.ascontiguousarray(axes, dtype=np.int32)
        angles = np.ascontiguousarray(angles, dtype=np.float64)
        if axes.ndim != 3 or axes.shape[1:] != (self.L, self.n) or angles.shape != axes.shape:
            raise ValueError('axes and angles must have shape (B, {}, {})'.format(self.L, self.n))
        B = axes.shape[0]
        if B > self.batch:
            raise ValueError('{} circuits, the plan holds {}'.format(B, self.batch))
        ob = observable if isinstance(observable, Observable) else Observable(observable)
        arr, k = ob.c_terms()
        E = np.empty(B)
        grad = np.empty((B, self.L, self.n))
        _native.call('qg_mcclean_batch_grad_run', self._plan, B, _native.iptr(axes), _native.dptr(angles), arr, k,
                     _native.dptr(E), _native.dptr(grad))
        return E, grad

    def last_device_ms(self):
        ms = ctypes.c_double()
        _native.call('qg_plan_last_ms', self._plan, ctypes.byref(ms))
        return ms.value
