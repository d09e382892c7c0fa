"""Matrix-free numpy restatement of qradient's statevector ``grad_run`` path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the product package
never imports this module.

Conventions restated from the reference (paths relative to the reference
repository root):

* qubit ``q`` is index bit ``n-1-q`` (qubit 0 is the most significant bit):
  single-qubit operators are ``kron(kron(I_{2^q}, P), I_{2^{n-q-1}})``
  (qradient/physical_components/state.py:84-87) and the Z masks use
  ``interval = 2**(n-1-q)`` (state.py:163-166).  In a C-order reshape to
  ``(2,)*n`` qubit ``q`` is axis ``q``.
* ``xrot(t, q)``: ``v <- sin(t/2) (-iX_q) v + cos(t/2) v`` (state.py:90-92);
  ``yrot`` likewise with Y (state.py:142-144); ``zrot``:
  ``e^{-it/2}`` on bit 0, ``e^{+it/2}`` on bit 1 (state.py:168-170).
  ``d?rot`` are the exact angle derivatives (state.py:94-97, 146-149, 172-175).
* CNOT(i, j): control ``i``, target ``j`` (state.py:336-356).  The ladder
  uses bonds ``(i, i+1)``; ``stacking=0`` is the product built at
  state.py:235-239, whose right-most factors are the odd bonds, so odd bonds
  act first; ``stacking=1`` (state.py:237-241) acts even bonds first and is the
  inverse.
* McClean layout (qradient/circuit_logic/mc_clean.py:47-78): ``RY(pi/4)`` on
  every qubit, then per layer the ladder ``stacking=0`` followed by one
  rotation per qubit about axis ``axes[i,q]`` (0 X, 1 Y, 2 Z; mc_clean.py:383-392).
* QAOA layout (qradient/circuit_logic/qaoa.py:40-70): ``|+>^n``; per layer the
  cost phase ``exp(-i gamma C)`` then ``RX(beta)`` on every qubit.
"""
import numpy as np

AXIS_X, AXIS_Y, AXIS_Z = 0, 1, 2


# --------------------------------------------------------------------------
# states (state.py:61-71)
# --------------------------------------------------------------------------
def reset_state(n, ini='0', dtype=np.complex128):
    """|0...0> for ini='0', 2^{-n/2}(1,...,1) for ini='+' (state.py:61-71)."""
    if ini == '0':
        v = np.zeros(2 ** n, dtype=dtype)
        v[0] = 1.0
    elif ini == '+':
        v = np.full(2 ** n, 2.0 ** (-0.5 * n), dtype=dtype)
    else:
        raise ValueError('Invalid initialization format {}.'.format(ini))
    return v


def _view(v, q):
    n = int(v.size).bit_length() - 1
    return v.reshape(2 ** q, 2, 2 ** (n - q - 1))


def _real(dtype):
    return np.float32 if np.dtype(dtype) == np.complex64 else np.float64


# --------------------------------------------------------------------------
# single-qubit rotations and derivatives (state.py:79-175)
# --------------------------------------------------------------------------
def xrot(v, angle, q):
    """RX: sin(a/2)(-iX)v + cos(a/2)v (state.py:90-92)."""
    r = _real(v.dtype)
    c, s = r(np.cos(0.5 * angle)), r(np.sin(0.5 * angle))
    w = _view(v, q)
    a0, a1 = w[:, 0, :].copy(), w[:, 1, :].copy()
    # (-iX v)_0 = -i v_1, (-iX v)_1 = -i v_0
    w[:, 0, :] = c * a0 + s * (-1j * a1).astype(v.dtype)
    w[:, 1, :] = c * a1 + s * (-1j * a0).astype(v.dtype)
    return v


def yrot(v, angle, q):
    """RY: sin(a/2)(-iY)v + cos(a/2)v (state.py:142-144); -iY = [[0,-1],[1,0]]."""
    r = _real(v.dtype)
    c, s = r(np.cos(0.5 * angle)), r(np.sin(0.5 * angle))
    w = _view(v, q)
    a0, a1 = w[:, 0, :].copy(), w[:, 1, :].copy()
    w[:, 0, :] = c * a0 - s * a1
    w[:, 1, :] = c * a1 + s * a0
    return v


def zrot(v, angle, q):
    """RZ: e^{-ia/2} on qubit value 0, e^{+ia/2} on 1 (state.py:168-170)."""
    ph0 = np.exp(-0.5j * angle).astype(v.dtype)
    ph1 = np.exp(0.5j * angle).astype(v.dtype)
    w = _view(v, q)
    w[:, 0, :] *= ph0
    w[:, 1, :] *= ph1
    return v


def dxrot(v, angle, q):
    """d/da of xrot: .5cos(a/2)(-iX)v - .5sin(a/2)v (state.py:94-97)."""
    w = _view(v, q)
    a0, a1 = w[:, 0, :].copy(), w[:, 1, :].copy()
    c, s = 0.5 * np.cos(0.5 * angle), 0.5 * np.sin(0.5 * angle)
    w[:, 0, :] = c * (-1j * a1) - s * a0
    w[:, 1, :] = c * (-1j * a0) - s * a1
    return v


def dyrot(v, angle, q):
    """d/da of yrot (state.py:146-149)."""
    w = _view(v, q)
    a0, a1 = w[:, 0, :].copy(), w[:, 1, :].copy()
    c, s = 0.5 * np.cos(0.5 * angle), 0.5 * np.sin(0.5 * angle)
    w[:, 0, :] = c * (-a1) - s * a0
    w[:, 1, :] = c * a0 - s * a1
    return v


def dzrot(v, angle, q):
    """d/da of zrot (state.py:172-175)."""
    w = _view(v, q)
    w[:, 0, :] *= -0.5j * np.exp(-0.5j * angle)
    w[:, 1, :] *= 0.5j * np.exp(0.5j * angle)
    return v


_ROT = {AXIS_X: xrot, AXIS_Y: yrot, AXIS_Z: zrot}
_DROT = {AXIS_X: dxrot, AXIS_Y: dyrot, AXIS_Z: dzrot}


def rot(v, axis, angle, q):
    """Axis dispatch of McClean.__rot (mc_clean.py:383-392)."""
    if axis not in _ROT:
        raise ValueError('Invalid axis {}'.format(axis))
    return _ROT[axis](v, angle, q)


def drot(v, axis, angle, q):
    """Axis dispatch of McClean.__drot (mc_clean.py:405-414)."""
    if axis not in _DROT:
        raise ValueError('Invalid axis {}'.format(axis))
    return _DROT[axis](v, angle, q)


# --------------------------------------------------------------------------
# entanglers (state.py:183-251, 336-356)
# --------------------------------------------------------------------------
def cnot(v, i, j):
    """CNOT with control i, target j (state.py:336-356)."""
    n = int(v.size).bit_length() - 1
    if not ((i < j < n) or (j < i < n)) or i < 0 or j < 0:
        raise ValueError('Invalid CNOT indecies {} and {}, for {} qubits.'.format(i, j, n))
    t = v.reshape((2,) * n)
    idx_c1 = [slice(None)] * n
    idx_c1[i] = 1
    sub = t[tuple(idx_c1)]                      # control = 1 block (n-1 axes)
    tj = j if j < i else j - 1                  # target axis inside the block
    sub[...] = np.flip(sub, axis=tj).copy()
    return v


def ladder_bonds(n, periodic=False):
    """Bonds (i, (i+1)%n) in the two groups of state.py:231-241."""
    upper = n if periodic else n - 1
    even = [(i, (i + 1) % n) for i in range(0, upper, 2)]
    odd = [(i, (i + 1) % n) for i in range(1, upper, 2)]
    return even, odd


def cnot_ladder(v, stacking, periodic=False):
    """Apply the CNOT ladder (state.py:209-251).

    stacking 0: matrix C_even * C_odd (state.py:235-239) -> odd bonds act
    first.  stacking 1: C_odd * C_even (state.py:237-241) -> even bonds act
    first; it is the inverse of stacking 0.
    """
    n = int(v.size).bit_length() - 1
    if periodic and n % 2 != 0:
        raise ValueError('CNOT gates in a ladder structure with periodic boundaries '
                         'are ambiguous for uneven number of qubits.')
    even, odd = ladder_bonds(n, periodic)
    first, second = (odd, even) if stacking == 0 else (even, odd)
    for (a, b) in first:
        cnot(v, a, b)
    for (a, b) in second:
        cnot(v, a, b)
    return v


def cz_ladder(v):
    """Nearest-neighbour CZ ladder: phase (-1)^{sum_i b_i b_{i+1}} (self-inverse)."""
    n = int(v.size).bit_length() - 1
    x = np.arange(v.size, dtype=np.int64)
    par = np.zeros(v.size, dtype=np.int64)
    for i in range(n - 1):
        bi = (x >> (n - 1 - i)) & 1
        bj = (x >> (n - 2 - i)) & 1
        par ^= bi & bj
    v *= np.where(par == 1, -1.0, 1.0).astype(v.dtype)
    return v


def entangle(v, entangler, inverse=False):
    if entangler == 'cnot':
        return cnot_ladder(v, 1 if inverse else 0)
    if entangler == 'cz':
        return cz_ladder(v)
    raise ValueError('unknown entangler {}'.format(entangler))


# --------------------------------------------------------------------------
# observables (observable.py:7-79) and classical Hamiltonian (state.py:261-297)
# --------------------------------------------------------------------------
def _obs_qnum(obs):
    return next(iter(obs.values())).shape[0]


def z_signs(n):
    """z[q, x] = +1 if qubit q (bit n-1-q) of x is 0, else -1 (state.py:163-166)."""
    x = np.arange(2 ** n, dtype=np.int64)
    return np.stack([1.0 - 2.0 * ((x >> (n - 1 - q)) & 1) for q in range(n)])


def observable_terms(obs):
    """Flatten an observable dict into [(kind, i, j, w)] (observable.py:44-73)."""
    n = _obs_qnum(obs)
    terms = []
    for key in ('x', 'y', 'z'):
        if key in obs and obs[key] is not None:
            for i, w in enumerate(obs[key]):
                if w is not None:
                    terms.append((key, i, -1, float(w)))
    if 'zz' in obs and obs['zz'] is not None:
        zz = obs['zz']
        for i in range(n):
            for j in range(i + 1):
                if zz[i, j] is not None:
                    raise ValueError('zz of observable should be a upper triangular matrix.')
            for j in range(i + 1, n):
                if zz[i, j] is not None:
                    terms.append(('zz', i, j, float(zz[i, j])))
    return n, terms


def diagonal(n, terms):
    """Diagonal of the z/zz part: sum w z_i (+ sum w z_i z_j) (state.py:273-294)."""
    zs = z_signs(n)
    d = np.zeros(2 ** n)
    for kind, i, j, w in terms:
        if kind == 'z':
            d += w * zs[i]
        elif kind == 'zz':
            d += w * zs[i] * zs[j]
    return d


def apply_observable(v, obs):
    """lambda = O v for the dict observable (observable.py:34-79 matrix, applied)."""
    n, terms = observable_terms(obs)
    out = (diagonal(n, terms) * v).astype(v.dtype)
    for kind, i, _, w in terms:
        if kind in ('x', 'y'):
            u = v.copy()
            w_ = _view(u, i)
            a0, a1 = w_[:, 0, :].copy(), w_[:, 1, :].copy()
            if kind == 'x':
                w_[:, 0, :], w_[:, 1, :] = a1, a0
            else:
                w_[:, 0, :], w_[:, 1, :] = -1j * a1, 1j * a0
            out += w * u
    return out


def expec_val(v, obs):
    """<v|O|v> (qradient/circuit_logic/base.py:17-20)."""
    return float(np.vdot(v.astype(np.complex128), apply_observable(v, obs).astype(np.complex128)).real)


# --------------------------------------------------------------------------
# finite-shot expectation values (base.py:22-46, observable.py:82-104, 126-177)
# --------------------------------------------------------------------------
_PROJ = {'x': np.array([[.5, .5], [.5, .5]], dtype=complex),      # observable.py:130
         'y': np.array([[.5, -.5j], [.5j, .5]], dtype=complex)}   # observable.py:137


def projector_prob(v, kind, i, j=-1):
    """|P^+ v|^2 for one term: the +1-eigenspace projector of Projector
    (observable.py:126-177) applied to v, then the squared norm (base.py:28)."""
    v = np.asarray(v, dtype=np.complex128)
    n = int(v.size).bit_length() - 1
    if kind in ('x', 'y'):
        w = _view(v, i)
        P = _PROJ[kind]
        p0 = P[0, 0] * w[:, 0, :] + P[0, 1] * w[:, 1, :]
        p1 = P[1, 0] * w[:, 0, :] + P[1, 1] * w[:, 1, :]
        return float((np.abs(p0) ** 2).sum() + (np.abs(p1) ** 2).sum())
    zs = z_signs(n)
    if kind == 'z':
        keep = zs[i] > 0
    elif kind == 'zz':
        keep = zs[i] * zs[j] > 0   # up-up + down-down (observable.py:150-164)
    else:
        raise ValueError('Unknown key for projector {}.'.format(kind))
    return float((np.abs(v[keep]) ** 2).sum())


def term_probs(v, obs):
    """Projector probabilities of every term, in load_projectors order
    (observable.py:82-104: x, y, z by qubit, then zz row-major)."""
    _, terms = observable_terms(obs)
    return np.array([projector_prob(v, k, i, j) for k, i, j, _ in terms]), np.array([t[3] for t in terms])


def sample_expec_val(v, obs, shot_num, random_state=None):
    """ParametrizedCircuit.sample_expec_val (base.py:22-34), with scipy's
    rv_discrete as the reference calls it (random_state=None: numpy's global
    RandomState, as there)."""
    from scipy import stats
    probs, weights = term_probs(v, obs)
    total = 0.
    for prob, weight in zip(probs, weights):
        rnd_var = stats.rv_discrete(values=([0, 1], [prob, 1 - prob]))
        samples = np.array([weight, -weight])[rnd_var.rvs(size=shot_num, random_state=random_state)]
        total += samples.mean()
    return total


# --------------------------------------------------------------------------
# McClean (mc_clean.py:9-78)
# --------------------------------------------------------------------------
def mcclean_random_params(L, n, rng=None):
    """axes = int(3U), angles = 2 pi U, axes drawn first (mc_clean.py:13-14)."""
    rand = np.random.rand if rng is None else rng.rand
    axes = (3 * rand(L, n)).astype('int')
    angles = 2 * np.pi * rand(L, n)
    return axes, angles


def mcclean_prepare(n, ini_state=None, dtype=np.complex128):
    """Reset (or take ini_state) then RY(pi/4) on every qubit (mc_clean.py:48-56)."""
    v = reset_state(n, '0', dtype) if ini_state is None else np.array(ini_state, dtype=dtype)
    for q in range(n):
        yrot(v, np.pi / 4.0, q)
    return v


def mcclean_forward(n, axes, angles, ini_state=None, entangler='cnot', dtype=np.complex128):
    """Forward circuit of run_expec_val (mc_clean.py:27-41); returns the final state."""
    v = mcclean_prepare(n, ini_state, dtype)
    for i in range(axes.shape[0]):
        entangle(v, entangler)
        for q in range(n):
            rot(v, axes[i, q], angles[i, q], q)
    return v


def mcclean_run_expec_val(n, axes, angles, obs, ini_state=None, entangler='cnot'):
    """run_expec_val with exact expectation (mc_clean.py:27-43, base.py:17-20)."""
    return expec_val(mcclean_forward(n, axes, angles, ini_state, entangler), obs)


def mcclean_grad_run(n, axes, angles, obs, ini_state=None, entangler='cnot'):
    """Restatement of McClean.grad_run (mc_clean.py:47-78), complex128.

    Stores the history of pre-rotation states, computes lambda = O psi_L and
    E = Re <psi_L|lambda>, then walks the layers backwards: un-rotate lambda,
    and for every qubit form ``-2 Re <hist_i | drot(-a) rot(a) lambda>``
    (mc_clean.py:69-76) before undoing the entangler (mc_clean.py:77).
    """
    L = axes.shape[0]
    v = mcclean_prepare(n, ini_state)
    hist = np.empty((L + 1, 2 ** n), dtype=np.complex128)
    for i in range(L):
        entangle(v, entangler)
        hist[i] = v
        for q in range(n):
            rot(v, axes[i, q], angles[i, q], q)
    hist[L] = v
    lam = apply_observable(v, obs)
    E = float(np.vdot(hist[L], lam).real)
    grad = np.empty((L, n))
    for i in range(L - 1, -1, -1):
        for q in range(n):
            rot(lam, axes[i, q], -angles[i, q], q)
        tmp = lam.copy()
        for q in range(n):
            rot(lam, axes[i, q], angles[i, q], q)
            drot(lam, axes[i, q], -angles[i, q], q)
            grad[i, q] = -2.0 * np.vdot(hist[i], lam).real
            lam[:] = tmp
        entangle(lam, entangler, inverse=True)
    return E, grad


def pauli_overlap_im(lam, psi, axis, q):
    """Im <lam | P_q | psi> accumulated in float64 (P = X, Y, Z by axis)."""
    lam = lam.astype(np.complex128)
    psi = psi.astype(np.complex128)
    l = _view(lam, q)
    p = _view(psi, q)
    if axis == AXIS_Z:
        z = np.vdot(l[:, 0, :], p[:, 0, :]) - np.vdot(l[:, 1, :], p[:, 1, :])
    elif axis == AXIS_X:
        z = np.vdot(l[:, 0, :], p[:, 1, :]) + np.vdot(l[:, 1, :], p[:, 0, :])
    elif axis == AXIS_Y:
        z = np.vdot(l[:, 0, :], -1j * p[:, 1, :]) + np.vdot(l[:, 1, :], 1j * p[:, 0, :])
    else:
        raise ValueError('Invalid axis {}'.format(axis))
    return float(z.imag)


def mcclean_grad_run_adjoint(n, axes, angles, obs, ini_state=None, entangler='cnot',
                             dtype=np.complex128, checkpoint_every=0):
    """The same gradient through the identity dE/da_{iq} = Im<lam_i|P_q|psi_i>.

    psi is uncomputed alongside lambda instead of being stored; with
    ``checkpoint_every=k > 0`` psi is restarted from a stored copy every k
    layers.  ``dtype=np.complex64`` rounds every gate to complex64 while the
    inner products are accumulated in float64 -- the numerical model of the
    GPU's complex64 mode.  Equal to ``mcclean_grad_run`` in exact arithmetic
    (derivation: dR(-a)R(a) = -(i/2)P, mc_clean.py:72-75).
    """
    L = axes.shape[0]
    psi = mcclean_prepare(n, ini_state, dtype)
    ckpt = {}
    for i in range(L):
        entangle(psi, entangler)
        for q in range(n):
            rot(psi, axes[i, q], angles[i, q], q)
        if checkpoint_every and (i + 1) % checkpoint_every == 0:
            ckpt[i + 1] = psi.copy()
    lam = apply_observable(psi, obs).astype(dtype)
    E = float(np.vdot(psi.astype(np.complex128), lam.astype(np.complex128)).real)
    grad = np.empty((L, n))
    for i in range(L - 1, -1, -1):
        if (i + 1) in ckpt:
            psi = ckpt[i + 1].copy()
        for q in range(n):
            grad[i, q] = pauli_overlap_im(lam, psi, axes[i, q], q)
        for q in range(n):
            rot(psi, axes[i, q], -angles[i, q], q)
            rot(lam, axes[i, q], -angles[i, q], q)
        entangle(psi, entangler, inverse=True)
        entangle(lam, entangler, inverse=True)
    return E, grad


# --------------------------------------------------------------------------
# Sampled McClean gradients (mc_clean.py:80-198)
# --------------------------------------------------------------------------
def component_observable(obs, k):
    """The dict observable of term k alone, with its weight: the reference's
    ``component_array[k]`` = weight * operator (observable.py:44-73)."""
    n, terms = observable_terms(obs)
    kind, i, j, w = terms[k]
    if kind == 'zz':
        zz = np.full((n, n), None)
        zz[i, j] = w
        return {'zz': zz}
    arr = np.array([None] * n, dtype=object)
    arr[i] = w
    return {kind: arr}


def weight_distribution(obs):
    """Component sampling distribution.  The reference's w / sum(w)
    (observable.py:75-79) for non-negative weights; |w| / sum|w| in general
    (the reference's is not a distribution when a weight is negative)."""
    _, terms = observable_terms(obs)
    w = np.abs(np.array([t[3] for t in terms]))
    return w / w.sum()


def mcclean_grad_run_component(n, axes, angles, obs, component, ini_state=None, entangler='cnot'):
    """McClean.grad_run_with_component_sampling (mc_clean.py:80-115) for a
    given component: the forward sweep with history, E of the whole
    observable, then lambda = component_array[k] psi_L and the reverse sweep
    of grad_run (mc_clean.py:67-78)."""
    E = mcclean_run_expec_val(n, axes, angles, obs, ini_state, entangler)
    _, grad = mcclean_grad_run(n, axes, angles, component_observable(obs, component), ini_state, entangler)
    return E, grad


def mcclean_shift_states(n, axes, angles, ini_state=None, entangler='cnot'):
    """The shifted final states of McClean.sample_grad (mc_clean.py:117-155),
    as the reference forms them: state_history[i] is the state after layer
    i's rotations (mc_clean.py:126-131); for each (i, dq) the state restarts
    from state_history[i], gets the manual rotation of +pi/2 about axes[i, dq]
    on qubit dq (mc_clean.py:394-403) and runs layers i+1..L-1, then the same
    with -pi/2.  Yields (i, dq, psi_plus, psi_minus)."""
    L = axes.shape[0]
    v = mcclean_prepare(n, ini_state)
    hist = []
    for i in range(L):
        entangle(v, entangler)
        for q in range(n):
            rot(v, axes[i, q], angles[i, q], q)
        hist.append(v.copy())
    for i in range(L):
        for dq in range(n):
            out = []
            for shift in (np.pi / 2, -np.pi / 2):
                u = hist[i].copy()
                rot(u, axes[i, dq], shift, dq)
                for j in range(i + 1, L):
                    entangle(u, entangler)
                    for q in range(n):
                        rot(u, axes[j, q], angles[j, q], q)
                out.append(u)
            yield i, dq, out[0], out[1]


def sample_component_expec_val(v, obs, shot_num, component, random_state=None):
    """ParametrizedCircuit.sample_component_expec_val (base.py:36-46)."""
    from scipy import stats
    probs, weights = term_probs(v, obs)
    prob, weight = probs[component], weights[component]
    rnd_var = stats.rv_discrete(values=([0, 1], [prob, 1 - prob]))
    return np.array([weight, -weight])[rnd_var.rvs(size=shot_num, random_state=random_state)].mean()


def mcclean_sample_grad(n, axes, angles, obs, shot_num, exact_expec_val=True, ini_state=None, entangler='cnot',
                        component_sampling=False):
    """McClean.sample_grad (mc_clean.py:117-155) and, with
    ``component_sampling``, sample_grad_with_component_sampling
    (mc_clean.py:157-198): random numbers from numpy's global RandomState in
    the reference's order (E's shots, then per parameter the component draw,
    o_plus's shots, o_minus's shots)."""
    L = axes.shape[0]
    final = mcclean_forward(n, axes, angles, ini_state, entangler)
    E = expec_val(final, obs) if exact_expec_val else sample_expec_val(final, obs, shot_num)
    grad = np.empty((L, n))
    dist = weight_distribution(obs) if component_sampling else None
    for i, dq, vp, vm in mcclean_shift_states(n, axes, angles, ini_state, entangler):
        if component_sampling:
            c = np.random.choice(np.arange(dist.size), p=dist)
            o_plus = sample_component_expec_val(vp, obs, shot_num, c)
            o_minus = sample_component_expec_val(vm, obs, shot_num, c)
        else:
            o_plus = sample_expec_val(vp, obs, shot_num)
            o_minus = sample_expec_val(vm, obs, shot_num)
        grad[i, dq] = .5 * (o_plus - o_minus)
    return E, grad


# --------------------------------------------------------------------------
# QAOA (qaoa.py:9-70; classical Hamiltonian state.py:261-321)
# --------------------------------------------------------------------------
def classical_ham(n, obs):
    """C(x) = sum w_i z_i + sum w_ij z_i z_j (state.py:273-294); other keys ignored."""
    _, terms = observable_terms(obs)
    return diagonal(n, [t for t in terms if t[0] in ('z', 'zz')])


def xrot_all(v, angle):
    """Qaoa.__xrot_all: RX(angle) on every qubit (qaoa.py:192-194)."""
    n = int(v.size).bit_length() - 1
    for q in range(n):
        xrot(v, angle, q)
    return v


def x_summed(v):
    """v <- .5 * sum_q (-i X_q) v (state.py:107-123)."""
    n = int(v.size).bit_length() - 1
    out = np.zeros_like(v)
    for q in range(n):
        u = v.copy()
        w = _view(u, q)
        a0, a1 = w[:, 0, :].copy(), w[:, 1, :].copy()
        w[:, 0, :], w[:, 1, :] = -1j * a1, -1j * a0
        out += 0.5 * u
    return out


def qaoa_forward(n, obs, betas, gammas, ini_state=None):
    """The state after Qaoa's forward loop (qaoa.py:29-35)."""
    C = classical_ham(n, obs)
    v = reset_state(n, '+') if ini_state is None else np.array(ini_state, dtype=np.complex128)
    for b, g in zip(betas, gammas):
        v *= np.exp(-1j * g * C)
        xrot_all(v, b)
    return v


def qaoa_run_expec_val(n, obs, betas, gammas, ini_state=None):
    """Qaoa.run_expec_val with exact expectation (qaoa.py:23-38)."""
    C = classical_ham(n, obs)
    v = qaoa_forward(n, obs, betas, gammas, ini_state)
    return float(np.vdot(v, C * v).real)


def qaoa_grad_run(n, obs, betas, gammas, ini_state=None):
    """Restatement of Qaoa.grad_run (qaoa.py:40-70).

    Returns (E, grad[p, 2]) with column 0 = d/d beta_i (from the summed-X
    generator, qaoa.py:60-64) and column 1 = d/d gamma_i (from -iC,
    qaoa.py:65-69).
    """
    betas = np.asarray(betas, dtype=float)
    gammas = np.asarray(gammas, dtype=float)
    p = betas.size
    if gammas.size != p:
        raise ValueError('Wrong amount of parameters.')
    C = classical_ham(n, obs)
    v = reset_state(n, '+') if ini_state is None else np.array(ini_state, dtype=np.complex128)
    hist = np.empty((2 * p + 1, 2 ** n), dtype=np.complex128)
    for i in range(p):
        hist[2 * i] = v
        v *= np.exp(-1j * gammas[i] * C)
        hist[2 * i + 1] = v
        xrot_all(v, betas[i])
    hist[2 * p] = v
    v = v * C
    E = float(np.vdot(hist[2 * p], v).real)
    grad = np.empty((p, 2))
    for i in range(p - 1, -1, -1):
        xrot_all(v, -betas[i])
        tmp = v.copy()
        u = x_summed(v)
        grad[i, 0] = -2.0 * np.vdot(hist[2 * i + 1], u).real
        v = tmp
        v *= np.exp(1j * gammas[i] * C)
        tmp = v.copy()
        u = v * (-1j * C)
        grad[i, 1] = -2.0 * np.vdot(hist[2 * i], u).real
        v = tmp
    return E, grad


def qaoa_grad_run_adjoint(n, obs, betas, gammas, ini_state=None, dtype=np.complex128):
    """QAOA gradient through Im<lam|sum X|psi> and 2 Im<lam|C|psi>, psi uncomputed.

    ``dtype=np.complex64`` models the GPU's complex64 state with float64
    reductions.
    """
    betas = np.asarray(betas, dtype=float)
    gammas = np.asarray(gammas, dtype=float)
    p = betas.size
    C = classical_ham(n, obs)
    psi = reset_state(n, '+', dtype) if ini_state is None else np.array(ini_state, dtype=dtype)
    for i in range(p):
        psi *= np.exp(-1j * gammas[i] * C).astype(dtype)
        xrot_all(psi, betas[i])
    lam = (C * psi).astype(dtype)
    E = float(np.vdot(psi.astype(np.complex128), lam.astype(np.complex128)).real)
    grad = np.empty((p, 2))
    for i in range(p - 1, -1, -1):
        grad[i, 0] = sum(pauli_overlap_im(lam, psi, AXIS_X, q) for q in range(n))
        xrot_all(psi, -betas[i])
        xrot_all(lam, -betas[i])
        grad[i, 1] = 2.0 * float(np.vdot(lam.astype(np.complex128), C * psi.astype(np.complex128)).imag)
        psi *= np.exp(1j * gammas[i] * C).astype(dtype)
        lam *= np.exp(1j * gammas[i] * C).astype(dtype)
    return E, grad


# --------------------------------------------------------------------------
# inputs
# --------------------------------------------------------------------------
def zz_chain_observable(n, weights):
    """{'zz': Z} with Z[i, i+1] = weights[i], None elsewhere (observable.py:56-73 format)."""
    zz = np.full((n, n), None)
    for i in range(n - 1):
        zz[i, i + 1] = float(weights[i])
    return {'zz': zz}


def maxcut_observable(n, edges):
    """MaxCut.to_observable (qradient/optimization_problems.py:33-37)."""
    zz = np.full((n, n), None)
    for a, b in edges:
        zz[a, b] = 1.0
    return {'zz': zz}


def random_regular3_edges(n, rng):
    """Seeded pairing-model 3-regular simple graph on n (even) vertices, edges a<b."""
    while True:
        stubs = np.repeat(np.arange(n), 3)
        rng.shuffle(stubs)
        pairs = stubs.reshape(-1, 2)
        ok = True
        seen = set()
        for a, b in pairs:
            if a == b:
                ok = False
                break
            e = (min(a, b), max(a, b))
            if e in seen:
                ok = False
                break
            seen.add(e)
        if ok:
            return sorted(seen)
