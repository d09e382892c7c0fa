"""Gates: the gate-cache builder the reference's circuits import.

The reference's circuits construct ``Gates(n).add_xrots().add_yrots()
.add_zrots().add_cnot_ladder()`` (mc_clean.py:16-20) and
``Gates(n).add_xrots().add_x_summed().add_classical_ham(obs, ...)``
(qaoa.py:11-14), but the class is absent from the reference snapshot; its API
is restated from those call sites with state.py as the semantic source.

On the GPU no operator is precomputed: rotations, ladders and the classical
Hamiltonian are evaluated from the index bits inside the kernels.  The builder
records what was requested; the operator attributes of the reference
(``xrot[q]``, ``cnot_ladder[s]``, ``classical_ham`` ...) are materialised on
first access for callers that inspect them, and ``classical_ham`` is computed
by a GPU kernel.
"""
import numpy as np

from .. import _native


class _Lazy(dict):
    def __init__(self, fn):
        super().__init__()
        self._fn = fn

    def __missing__(self, key):
        v = self._fn(key)
        self[key] = v
        return v


def _one_qubit_csr(n, i, P):
    import scipy.sparse as sp
    return sp.kron(sp.kron(sp.identity(2 ** i, dtype='complex'), P), sp.identity(2 ** (n - i - 1), dtype='complex'),
                   format='csr')


class Gates:
    def __init__(self, qubit_number):
        self.qnum = int(qubit_number)
        self.loaded = set()
        self._classical_terms = None
        self.periodic = False

    # ---- builder (returns self, as chained at mc_clean.py:16-20, qaoa.py:11-14)
    def add_xrots(self):
        self.loaded.add('xrots')
        return self

    def add_yrots(self):
        self.loaded.add('yrots')
        return self

    def add_zrots(self):
        self.loaded.add('zrots')
        return self

    def add_cnot_ladder(self, periodic=False):
        """state.py:209-215: a periodic ladder needs an even qubit number."""
        if periodic and self.qnum % 2 != 0:
            raise ValueError(('CNOT gates in a ladder structure with periodic boundaries',
                              'are ambiguous for uneven number of qubits.'))
        self.periodic = periodic
        self.loaded.add('cnot_ladder')
        return self

    def add_x_summed(self):
        self.loaded.add('x_summed')
        return self

    def add_classical_ham(self, observable, include_individual_components=False):
        """state.py:261-297: the z / zz part of the observable; other keys warn."""
        import warnings
        other = [t for t in observable.terms if t[0] in (_native.TERM_X, _native.TERM_Y)]
        if other:
            warnings.warn("Non-classical observable component found. Only 'z' and 'zz' are accepted in this method.")
        self._classical_terms = observable.classical_terms()
        self._components = include_individual_components
        self.loaded.add('classical_ham')
        return self

    # ---- reference attributes, materialised on request -------------------------
    @property
    def xrot(self):
        import scipy.sparse as sp
        X = sp.coo_matrix([[0., 1.], [1., 0.]], dtype='complex')
        return _Lazy(lambda q: -1.j * _one_qubit_csr(self.qnum, q, X))

    @property
    def yrot(self):
        import scipy.sparse as sp
        Y = sp.coo_matrix([[0., -1.j], [1.j, 0.]], dtype='complex')
        return _Lazy(lambda q: -1.j * _one_qubit_csr(self.qnum, q, Y))

    def _zmask(self, q, val):
        x = np.arange(2 ** self.qnum)
        return (((x >> (self.qnum - 1 - q)) & 1) == val).astype('complex')

    @property
    def zrot_pos(self):
        return _Lazy(lambda q: self._zmask(q, 0))

    @property
    def zrot_neg(self):
        return _Lazy(lambda q: self._zmask(q, 1))

    @property
    def classical_ham(self):
        """C(x) for every basis state, evaluated by the diagonal GPU kernel."""
        if self._classical_terms is None:
            raise AttributeError('add_classical_ham was not called')
        return self._diag_of(self._classical_terms)

    @property
    def classical_ham_components(self):
        if self._classical_terms is None:
            raise AttributeError('add_classical_ham was not called')
        return np.array([self._diag_of([t]) for t in self._classical_terms])

    def _diag_of(self, terms):
        from .state import State
        s = State(self.qnum, ini='+')
        arr, k = _native.terms_array(terms)
        _native.call('qg_state_multiply_observable', s._h, arr, k)
        return (s.vec.real * 2.0 ** (0.5 * self.qnum))
