"""The reference's own formulation of grad_run: scipy CSR operators + history.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  This is the CPU
baseline that ``bench.py`` times (``cpu_baseline.kind = "port"``) and a
second, independently constructed cross-check of ``oracle.statevector``:

* per-qubit X/Y rotation generators ``-1j*kron(kron(I_{2^i}, P), I)`` as CSR
  (qradient/physical_components/state.py:81-88, 133-140),
* 0/1 Z masks ``zrot_pos/zrot_neg`` (state.py:159-166),
* the CNOT ladder as a product of CSR CNOT matrices (state.py:209-241, 336-356),
* the observable as a CSR sum of Kronecker products (observable.py:34-73),
* QAOA's diagonal classical Hamiltonian and summed-X generator
  (state.py:107-123, 261-297),
* ``state_history`` of L+1 (McClean) or 2p+1 (QAOA) stored states
  (mc_clean.py:22, qaoa.py:19).

Operators are built lazily per qubit so a bounded sample of a large circuit
only builds what it touches.  Construction is excluded from timings, as in the
reference's own benchmarks (qradient-vs-pennylane-benchmark.py:62-78).
"""
import numpy as np
import scipy.sparse as sp

_X = sp.coo_matrix([[0., 1.], [1., 0.]], dtype='complex')
_Y = sp.coo_matrix([[0., -1.j], [1.j, 0.]], dtype='complex')
_Z = sp.coo_matrix([[1., 0.], [0., -1.]], dtype='complex')
_P0 = sp.coo_matrix([[1., 0.], [0., 0.]], dtype='complex')
_P1 = sp.coo_matrix([[0., 0.], [0., 1.]], dtype='complex')


def _id(k):
    return sp.identity(2 ** k, dtype='complex', format='coo')


def _one_qubit(n, i, P):
    return sp.kron(sp.kron(_id(i), P), _id(n - i - 1), format='csr')


def _cnot(n, i, j):
    """CNOT(i -> j) as CSR (state.py:336-356)."""
    if i < j < n:
        out1 = sp.kron(sp.kron(_id(i), _P0), _id(n - i - 1))
        out2 = sp.kron(sp.kron(sp.kron(sp.kron(_id(i), _P1), _id(j - i - 1)), _X), _id(n - j - 1))
    elif j < i < n:
        out1 = sp.kron(sp.kron(_id(i), _P0), _id(n - i - 1))
        out2 = sp.kron(sp.kron(sp.kron(sp.kron(_id(j), _X), _id(i - j - 1)), _P1), _id(n - i - 1))
    else:
        raise ValueError('Invalid CNOT indecies {} and {}, for {} qubits.'.format(i, j, n))
    return (out1 + out2).asformat('csr')


class Gates:
    """Lazy CSR gate cache with the reference's operator definitions."""

    def __init__(self, n):
        self.n = n
        self._xrot, self._yrot, self._zpos, self._zneg = {}, {}, {}, {}
        self._ladder = None

    def xrot_op(self, i):
        if i not in self._xrot:
            self._xrot[i] = -1.j * _one_qubit(self.n, i, _X)
        return self._xrot[i]

    def yrot_op(self, i):
        if i not in self._yrot:
            self._yrot[i] = -1.j * _one_qubit(self.n, i, _Y)
        return self._yrot[i]

    def zmasks(self, i):
        if i not in self._zpos:
            n = self.n
            interval = 2 ** (n - 1 - i)
            pos = np.zeros(2 ** n, dtype='complex')
            neg = np.zeros(2 ** n, dtype='complex')
            for j in range(0, 2 ** n, 2 * interval):
                pos[j:j + interval] += 1.
                neg[j + interval:j + 2 * interval] += 1.
            self._zpos[i], self._zneg[i] = pos, neg
        return self._zpos[i], self._zneg[i]

    def ladder(self):
        """[L0, L1] with the product order of state.py:229-241 (non-periodic)."""
        if self._ladder is None:
            n = self.n
            l0 = sp.identity(2 ** n, dtype='complex', format='csr')
            l1 = sp.identity(2 ** n, dtype='complex', format='csr')
            upper = n - 1
            for i in range(0, upper, 2):
                l0 = l0.dot(_cnot(n, i, i + 1))
            for i in range(1, upper, 2):
                c = _cnot(n, i, i + 1)
                l0 = l0.dot(c)
                l1 = l1.dot(c)
            for i in range(0, upper, 2):
                l1 = l1.dot(_cnot(n, i, i + 1))
            self._ladder = (l0.asformat('csr'), l1.asformat('csr'))
        return self._ladder


def observable_matrix(obs):
    """CSR matrix of the dict observable (observable.py:34-73)."""
    n = next(iter(obs.values())).shape[0]
    m = sp.csr_matrix((2 ** n, 2 ** n), dtype='complex')
    for key, P in (('x', _X), ('y', _Y), ('z', _Z)):
        if key in obs:
            for i, w in enumerate(obs[key]):
                if w is not None:
                    m = m + w * sp.kron(sp.kron(sp.identity(2 ** i), P), sp.identity(2 ** (n - i - 1)))
    if 'zz' in obs:
        zz = obs['zz']
        for i in range(n):
            for j in range(i + 1, n):
                if zz[i, j] is not None:
                    op = sp.kron(sp.kron(sp.identity(2 ** i), _Z), sp.identity(2 ** (j - i - 1)))
                    op = sp.kron(sp.kron(op, _Z), sp.identity(2 ** (n - j - 1)))
                    m = m + zz[i, j] * op
    return m.asformat('csr')


class McCleanRef:
    """McClean.grad_run in the reference formulation (mc_clean.py:47-78, 383-414)."""

    def __init__(self, n, obs, axes, angles):
        self.n, self.axes, self.angles = n, np.asarray(axes), np.asarray(angles)
        self.L = self.axes.shape[0]
        self.g = Gates(n)
        self.O = observable_matrix(obs)
        self.hist = None

    def build(self, layers=None):
        """Materialise the operators the first ``layers`` layers touch."""
        L = self.L if layers is None else layers
        self.g.ladder()
        for q in range(self.n):
            self.g.yrot_op(q)
            for i in range(L):
                a = self.axes[i, q]
                (self.g.xrot_op if a == 0 else self.g.yrot_op if a == 1 else self.g.zmasks)(q)
        self.hist = np.empty((L + 1, 2 ** self.n), dtype='complex')
        self.tmp = np.empty(2 ** self.n, dtype='complex')

    def _rot(self, v, i, q, sign=1.):
        a, t = self.axes[i, q], sign * self.angles[i, q]
        if a == 0:
            return np.sin(.5 * t) * self.g.xrot_op(q).dot(v) + np.cos(.5 * t) * v
        if a == 1:
            return np.sin(.5 * t) * self.g.yrot_op(q).dot(v) + np.cos(.5 * t) * v
        if a == 2:
            pos, neg = self.g.zmasks(q)
            return np.exp(-.5j * t) * pos * v + np.exp(.5j * t) * neg * v
        raise ValueError('Invalid axis {}'.format(a))

    def _drot(self, v, i, q, sign=1.):
        a, t = self.axes[i, q], sign * self.angles[i, q]
        if a == 0:
            return .5 * np.cos(.5 * t) * self.g.xrot_op(q).dot(v) - .5 * np.sin(.5 * t) * v
        if a == 1:
            return .5 * np.cos(.5 * t) * self.g.yrot_op(q).dot(v) - .5 * np.sin(.5 * t) * v
        if a == 2:
            pos, neg = self.g.zmasks(q)
            return -.5j * np.exp(-.5j * t) * pos * v + .5j * np.exp(.5j * t) * neg * v
        raise ValueError('Invalid axis {}'.format(a))

    def grad_run(self, layers=None):
        """Full grad_run (layers=None) or a bounded sample over the first ``layers``."""
        L = self.L if layers is None else layers
        if self.hist is None or self.hist.shape[0] != L + 1:
            self.build(L)
        n = self.n
        l0, l1 = self.g.ladder()
        v = np.zeros(2 ** n, dtype='complex')
        v[0] = 1.
        for q in range(n):
            v = np.sin(np.pi / 8) * self.g.yrot_op(q).dot(v) + np.cos(np.pi / 8) * v
        for i in range(L):
            v = l0.dot(v)
            self.hist[i] = v
            for q in range(n):
                v = self._rot(v, i, q)
        self.hist[L] = v
        v = self.O.dot(v)
        E = self.hist[L].conj().dot(v).real
        grad = np.empty((L, n))
        for i in range(L - 1, -1, -1):
            for q in range(n):
                v = self._rot(v, i, q, -1.)
            self.tmp[:] = v
            for q in range(n):
                v = self._rot(v, i, q)
                v = self._drot(v, i, q, -1.)
                grad[i, q] = -2. * self.hist[i].conj().dot(v).real
                v = self.tmp.copy()
            v = l1.dot(v)
        return E, grad


class QaoaRef:
    """Qaoa.grad_run in the reference formulation (qaoa.py:40-70)."""

    def __init__(self, n, obs, p):
        from .statevector import classical_ham
        self.n, self.p = n, p
        self.C = classical_ham(n, obs)
        self.x_ops = [-1.j * _one_qubit(n, q, _X) for q in range(n)]
        xs = sp.csr_matrix((2 ** n, 2 ** n), dtype='complex')
        for q in range(n):
            xs = xs + .5 * self.x_ops[q]
        self.x_summed = xs.asformat('csr')
        self.hist = np.empty((2 * p + 1, 2 ** n), dtype='complex')

    def _xrot_all(self, v, b):
        for op in self.x_ops:
            v = np.sin(.5 * b) * op.dot(v) + np.cos(.5 * b) * v
        return v

    def grad_run(self, betas, gammas):
        p = self.p
        v = np.full(2 ** self.n, 2. ** (-.5 * self.n), dtype='complex')
        for i in range(p):
            self.hist[2 * i] = v
            v = v * np.exp(-1.j * gammas[i] * self.C)
            self.hist[2 * i + 1] = v
            v = self._xrot_all(v, betas[i])
        self.hist[2 * p] = v
        v = v * self.C
        E = self.hist[2 * p].conj().dot(v).real
        grad = np.empty((p, 2))
        for i in range(p - 1, -1, -1):
            v = self._xrot_all(v, -betas[i])
            tmp = v.copy()
            u = self.x_summed.dot(v)
            grad[i, 0] = -2. * self.hist[2 * i + 1].conj().dot(u).real
            v = tmp * np.exp(1.j * gammas[i] * self.C)
            u = v * (-1.j * self.C)
            grad[i, 1] = -2. * self.hist[2 * i].conj().dot(u).real
        return E, grad


# --------------------------------------------------------------------------
# Dense sampled gradients in the reference's formulation: eigenbasis from a
# dense eigh, suffix circuits as left-hand-side matrices, rv_discrete draws
# (mc_clean.py:207-268, 277-350; qaoa.py:83-158, 196-198).  Tiny n only.
# --------------------------------------------------------------------------
def _rv_mean(dist, values, shot_num, exact_means):
    """values[rv_discrete(values=(arange(N), dist)).rvs(size=shot_num)].mean(),
    or with ``exact_means`` the distribution's mean sum(dist * values) (no
    random numbers): the deterministic form used to pin the shifted states."""
    if exact_means:
        return float(np.dot(dist, values))
    from scipy import stats
    rnd_var = stats.rv_discrete(values=(np.arange(dist.size), dist))
    return values[rnd_var.rvs(size=shot_num)].mean()


def _rot_matrix(g, axis, angle, q):
    """The LeftHandSide.rot factor (mc_clean.py:361-376)."""
    ident = sp.identity(2 ** g.n, dtype='complex', format='csr')
    if axis == 0:
        return np.sin(.5 * angle) * g.xrot_op(q) + np.cos(.5 * angle) * ident
    if axis == 1:
        return np.sin(.5 * angle) * g.yrot_op(q) + np.cos(.5 * angle) * ident
    pos, neg = g.zmasks(q)
    return np.exp(-.5j * angle) * sp.diags(pos) + np.exp(.5j * angle) * sp.diags(neg)


def _component_matrices(obs):
    """component_array of observable.py:41-73: weight * operator, in order."""
    n = next(iter(obs.values())).shape[0]
    out = []
    for key, P in (('x', _X), ('y', _Y), ('z', _Z)):
        if key in obs:
            for i, w in enumerate(obs[key]):
                if w is not None:
                    out.append(w * sp.kron(sp.kron(sp.identity(2 ** i), P), sp.identity(2 ** (n - i - 1))))
    if 'zz' in obs:
        zz = obs['zz']
        for i in range(n):
            for j in range(i + 1, n):
                if zz[i, j] is not None:
                    op = sp.kron(sp.kron(sp.identity(2 ** i), _Z), sp.identity(2 ** (j - i - 1)))
                    op = sp.kron(sp.kron(op, _Z), sp.identity(2 ** (n - j - 1)))
                    out.append(zz[i, j] * op)
    return out


def mcclean_sample_grad_dense(n, obs, axes, angles, shot_num, exact_expec_val=True, component_sampling=False,
                              exact_means=False):
    """McClean.sample_grad_dense (mc_clean.py:207-268) and, with
    ``component_sampling``, sample_grad_dense_with_component_sampling
    (mc_clean.py:277-350), in the reference's dense formulation, with numpy's
    global RandomState in the reference's order."""
    from . import statevector as sv
    axes, angles = np.asarray(axes), np.asarray(angles)
    L = axes.shape[0]
    ref = McCleanRef(n, obs, axes, angles)
    g = ref.g
    if component_sampling:
        comps = _component_matrices(obs)
        c = np.random.choice(np.arange(len(comps)), p=sv.weight_distribution(obs))
        M = comps[c]
    else:
        M = ref.O
    eigenvalues, eigenvectors = np.linalg.eigh(M.toarray())
    lhs = sp.csr_matrix(eigenvectors.transpose().conj())
    lhs_history = [lhs.toarray()]
    l0, _ = g.ladder()
    for i in range(L - 1):
        i_inv = L - i - 1
        for q in range(n):
            lhs = lhs.dot(_rot_matrix(g, axes[i_inv, q], angles[i_inv, q], q))
        lhs = lhs.dot(l0)
        lhs_history.append(lhs.toarray())
    # forward sweep, history after each layer's rotations (mc_clean.py:226-233)
    v = np.zeros(2 ** n, dtype='complex')
    v[0] = 1.
    for q in range(n):
        v = _rot_matrix(g, 1, np.pi / 4, q).dot(v)
    hist = []
    for i in range(L):
        v = l0.dot(v)
        for q in range(n):
            v = _rot_matrix(g, axes[i, q], angles[i, q], q).dot(v)
        hist.append(v)
    E = float(v.conj().dot(ref.O.dot(v)).real) if exact_expec_val else sv.sample_expec_val(v, obs, shot_num)
    grad = np.empty((L, n))
    for i in range(L):
        for q in range(n):
            s = _rot_matrix(g, axes[i, q], np.pi / 2, q).dot(hist[i])
            d1 = np.abs(lhs_history[L - i - 1].dot(s)) ** 2
            s1 = _rv_mean(d1, eigenvalues, shot_num, exact_means)
            s = _rot_matrix(g, axes[i, q], -np.pi, q).dot(s)
            d2 = np.abs(lhs_history[L - i - 1].dot(s)) ** 2
            s2 = _rv_mean(d2, eigenvalues, shot_num, exact_means)
            grad[i, q] = (s1 - s2) / 2.
    return E, grad


def qaoa_sample_grad_dense(n, obs, betas, gammas, shot_num, exact_expec_val=True, ini_state=None, exact_means=False):
    """Qaoa.sample_grad_dense (qaoa.py:83-158) in the reference's dense
    formulation: eigenvalues = the classical Hamiltonian in natural order
    (qaoa.py:94), suffix circuits as left-hand-side matrices
    (qaoa.py:127-133), rv_discrete draws (qaoa.py:196-198)."""
    from . import statevector as sv
    p = len(betas)
    C = sv.classical_ham(n, obs)
    _, terms = sv.observable_terms(obs)
    comps = [sv.diagonal(n, [t]) for t in terms if t[0] in ('z', 'zz')]
    ident = sp.identity(2 ** n, dtype='complex', format='csr')
    x_ops = [-1.j * _one_qubit(n, q, _X) for q in range(n)]

    def xrot_all(m_or_v, b, right):
        for op in x_ops:
            f = np.sin(.5 * b) * op + np.cos(.5 * b) * ident
            m_or_v = m_or_v.dot(f) if right else f.dot(m_or_v)
        return m_or_v

    lhs = ident
    lhs_history = [None] * (2 * p)
    for i in range(p):
        i_inv = p - i - 1
        lhs_history[2 * i] = lhs.toarray()
        lhs = xrot_all(lhs, betas[i_inv], True)
        lhs_history[2 * i + 1] = lhs.toarray()
        lhs = lhs.dot(sp.diags(np.exp(-1.j * gammas[i_inv] * C), format='csr'))
    v = np.full(2 ** n, 2. ** (-.5 * n), dtype='complex') if ini_state is None else \
        np.array(ini_state, dtype='complex')
    hist = []
    for i in range(p):
        v = v * np.exp(-1.j * gammas[i] * C)
        hist.append(v)
        v = xrot_all(v, betas[i], False)
        hist.append(v)
    E = float(v.conj().dot(C * v).real) if exact_expec_val else sv.sample_expec_val(v, obs, shot_num)
    grad = np.empty((p, 2))
    for i in range(p):
        total = 0.
        for comp in comps:
            s = hist[2 * i] * np.exp(-1.j * (np.pi / 4) * comp)
            plus = _rv_mean(np.abs(lhs_history[2 * (p - 1 - i) + 1].dot(s)) ** 2, C, shot_num, exact_means)
            s = s * np.exp(-1.j * (-np.pi / 2) * comp)
            minus = _rv_mean(np.abs(lhs_history[2 * (p - 1 - i) + 1].dot(s)) ** 2, C, shot_num, exact_means)
            total += plus - minus
        grad[i, 1] = total
        total = 0.
        for q in range(n):
            f = np.sin(.25 * np.pi) * x_ops[q] + np.cos(.25 * np.pi) * ident
            s = f.dot(hist[2 * i + 1])
            plus = _rv_mean(np.abs(lhs_history[2 * (p - 1 - i)].dot(s)) ** 2, C, shot_num, exact_means)
            f = np.sin(-.5 * np.pi) * x_ops[q] + np.cos(-.5 * np.pi) * ident
            s = f.dot(s)
            minus = _rv_mean(np.abs(lhs_history[2 * (p - 1 - i)].dot(s)) ** 2, C, shot_num, exact_means)
            total += plus - minus
        grad[i, 0] = .5 * total
    return E, grad
