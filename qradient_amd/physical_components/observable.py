"""Observable: dict of Pauli weights -> terms for the GPU kernels.

Mirrors qradient/physical_components/observable.py:7-79.  The dict may hold
'x', 'y', 'z' (1-d arrays of weights or None) and 'zz' (n x n object array,
strictly upper triangular, None elsewhere).  ``terms`` is the flattened form
the C ABI takes; ``matrix`` is the reference's sparse matrix, built only on
request (it is a representation for callers, never used by grad_run).
"""
import warnings

import numpy as np

from .. import _native

_KNOWN = ('x', 'y', 'z', 'zz')
_KIND = {'x': _native.TERM_X, 'y': _native.TERM_Y, 'z': _native.TERM_Z}


def _weight_check(weight, component):
    """observable.py:119-124."""
    if abs(weight) < 10. ** -15:
        warnings.warn(('Weight in observable {} is zero or almost zero.'.format(component),
                       " If you dont't want to include it, set it to None."))


class Observable:
    """``Observable(dict)`` (observable.py:8) or ``Observable(n, dict, store_components=)``
    (the call form of qradient/circuit_logic/base.py:12)."""

    _matrices = {}

    def __init__(self, *args, store_components=False):
        if len(args) == 1:
            observable = args[0]
        elif len(args) == 2:
            observable = args[1]
        else:
            raise TypeError('Observable(dict) or Observable(qubit_number, dict)')
        if not isinstance(observable, dict) or not observable:
            raise ValueError('observable must be a non-empty dict')
        arrays = {k: np.asarray(v, dtype=object) for k, v in observable.items()}
        self.qnum = next(iter(arrays.values())).shape[0]
        for v in arrays.values():
            for s in v.shape:
                if s != self.qnum:
                    raise ValueError(('Inconsistent shapes in observable dictionary.',
                                      'Cannot infer qubit_number.'))
        if len(args) == 2 and int(args[0]) != self.qnum:
            raise ValueError('qubit_number {} does not match the observable ({} qubits)'.format(args[0], self.qnum))
        self.dict = observable
        self.store_components = store_components
        for key in arrays:
            if key not in _KNOWN:
                warnings.warn('Unknown element of observable {} will be ignored.'.format(key))
        self.terms = self._flatten(arrays)
        self.is_diagonal = all(t[0] in (_native.TERM_Z, _native.TERM_ZZ) for t in self.terms)
        if store_components:
            self.num_components = len(self.terms)
            w = np.array([t[3] for t in self.terms])
            self.component_weights = w
            # |w|-proportional: a valid distribution also for negative weights
            self.weight_distribution = np.abs(w) / np.abs(w).sum() if len(w) else w

    def _flatten(self, arrays):
        """(kind, i, j, w) in the reference's accumulation order (observable.py:44-73)."""
        n = self.qnum
        terms = []
        for key in ('x', 'y', 'z'):
            if key in arrays:
                for i, w in enumerate(arrays[key]):
                    if w is not None:
                        _weight_check(w, key)
                        terms.append((_KIND[key], i, -1, float(w)))
        if 'zz' in arrays:
            zz = arrays['zz']
            if zz.ndim != 2:
                raise ValueError('zz of observable must be a {} by {} matrix.'.format(n, n))
            for i in range(n):
                for j in range(i + 1):
                    if zz[i, j] is not None:
                        raise ValueError(('zz of observable should be a upper triangular {} '.format(n),
                                          'by {} matrix. Diagonal and lower triangle should'.format(n),
                                          "contain None's, not {}.".format(zz[i, j])))
                for j in range(i + 1, n):
                    if zz[i, j] is not None:
                        _weight_check(zz[i, j], 'zz')
                        terms.append((_native.TERM_ZZ, i, j, float(zz[i, j])))
        return terms

    def c_terms(self):
        return _native.terms_array(self.terms)

    def classical_terms(self):
        """The z / zz terms: the classical Hamiltonian of state.py:261-297."""
        return [t for t in self.terms if t[0] in (_native.TERM_Z, _native.TERM_ZZ)]

    @property
    def matrix(self):
        """Sparse CSR matrix of the observable (observable.py:34-79), built on request."""
        if not hasattr(self, '_matrix'):
            import scipy.sparse as sp
            n = self.qnum
            P = {_native.TERM_X: sp.csr_matrix([[0., 1.], [1., 0.]], dtype='complex'),
                 _native.TERM_Y: sp.csr_matrix([[0., -1.j], [1.j, 0.]], dtype='complex'),
                 _native.TERM_Z: sp.csr_matrix([[1., 0.], [0., -1.]], dtype='complex')}
            m = sp.csr_matrix((2 ** n, 2 ** n), dtype='complex')
            for kind, i, j, w in self.terms:
                if kind == _native.TERM_ZZ:
                    op = sp.kron(sp.kron(sp.identity(2 ** i), P[_native.TERM_Z]), sp.identity(2 ** (j - i - 1)))
                    op = sp.kron(sp.kron(op, P[_native.TERM_Z]), sp.identity(2 ** (n - j - 1)))
                else:
                    op = sp.kron(sp.kron(sp.identity(2 ** i), P[kind]), sp.identity(2 ** (n - i - 1)))
                m = m + w * op
            self._matrix = m.asformat('csr')
            Observable._matrices[id(self._matrix)] = self
        return self._matrix

    @classmethod
    def from_matrix(cls, m):
        """The Observable whose ``matrix`` is ``m`` (for State.multiply_matrix), or None."""
        return cls._matrices.get(id(m))
