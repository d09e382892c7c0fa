"""Pins of the CPU oracle (oracle/) by checks independent of its own formulas.

The reference ships no golden vectors for this path (its tests are `pass`
stubs, reference tests/mc_clean.py:1-11) and its snapshot cannot run, so every
oracle function is pinned by mathematics a plausible slip would fail:

* gates     -- dense Kronecker products of scipy.linalg.expm(-i t P / 2)
               (matrix exponential, not the cos/sin form) and special angles;
* derivatives -- central finite differences of the gates;
* CNOT / ladder -- integer bit manipulation on basis states, and the dense
               product in the reference's factor order (state.py:229-241);
* observable -- dense sums of Kronecker products;
* McClean grad_run -- parameter shift [E(t+pi/2) - E(t-pi/2)]/2 (exact for
               these gates; the same rule as mc_clean.py:139-155) on a forward
               pass pinned by a dense brute-force unitary; analytic zeros; the
               independently built CSR formulation of oracle/reference_algorithm;
* QAOA      -- p=1 MaxCut closed form on triangle-free 3-regular graphs
               (Wang, Hadfield, Jiang, Rieffel, PRA 97, 022304 (2018), mapped to
               C = sum Z Z and RX(beta) = exp(-i beta X / 2)) and its analytic
               derivatives; dense brute force and finite differences for p > 1;
* RNG       -- numpy's frozen legacy RandomState stream (NEP 19).
"""
import numpy as np
import pytest
import scipy.linalg as la

from oracle import reference_algorithm as ra
from oracle import statevector as sv

I2 = np.eye(2)
PX = np.array([[0, 1], [1, 0]], dtype=complex)
PY = np.array([[0, -1j], [1j, 0]])
PZ = np.diag([1.0, -1.0]).astype(complex)
PAULI = {0: PX, 1: PY, 2: PZ}


def op1(n, q, U):
    return np.kron(np.kron(np.eye(2 ** q), U), np.eye(2 ** (n - q - 1)))


def rand_state(n, seed):
    rs = np.random.RandomState(seed)
    v = rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)
    return v / np.linalg.norm(v)


def dense_rot(axis, t):
    return la.expm(-0.5j * t * PAULI[axis])


# ---------------------------------------------------------------- gates -----
@pytest.mark.parametrize('axis,fn', [(0, sv.xrot), (1, sv.yrot), (2, sv.zrot)])
def test_rotations_match_dense_expm(axis, fn):
    n = 5
    v = rand_state(n, axis)
    for q in range(n):
        for t in (0.0, 0.3, -1.7, np.pi, 2 * np.pi, 5.1):
            got = fn(v.copy(), t, q)
            want = op1(n, q, dense_rot(axis, t)) @ v
            assert np.abs(got - want).max() < 1e-14


def test_rotation_special_angles():
    n = 3
    v = rand_state(n, 7)
    for fn in (sv.xrot, sv.yrot, sv.zrot):
        assert np.allclose(fn(v.copy(), 0.0, 1), v)
        assert np.allclose(fn(v.copy(), 2 * np.pi, 1), -v)
    assert np.allclose(sv.xrot(v.copy(), np.pi, 0), -1j * (op1(n, 0, PX) @ v))
    z = sv.reset_state(1)
    assert np.allclose(sv.yrot(z, np.pi / 2, 0), np.array([1, 1]) / np.sqrt(2))


@pytest.mark.parametrize('axis', [0, 1, 2])
def test_derivative_gates_are_finite_differences(axis):
    n, eps = 4, 1e-6
    v = rand_state(n, 11)
    for q in range(n):
        t = 0.9 + q
        fd = (sv.rot(v.copy(), axis, t + eps, q) - sv.rot(v.copy(), axis, t - eps, q)) / (2 * eps)
        assert np.abs(sv.drot(v.copy(), axis, t, q) - fd).max() < 1e-9


def test_rotation_composition():
    v = rand_state(4, 3)
    for axis in range(3):
        a = sv.rot(sv.rot(v.copy(), axis, 0.4, 2), axis, 1.3, 2)
        assert np.allclose(a, sv.rot(v.copy(), axis, 1.7, 2))


# ------------------------------------------------------------- entanglers ---
def _cnot_bits(n, x, c, t):
    if (x >> (n - 1 - c)) & 1:
        x ^= 1 << (n - 1 - t)
    return x


@pytest.mark.parametrize('n', [2, 3, 5])
def test_cnot_basis_brute_force(n):
    for c in range(n):
        for t in range(n):
            if c == t:
                continue
            for x in range(2 ** n):
                e = np.zeros(2 ** n, dtype=complex)
                e[x] = 1
                out = sv.cnot(e, c, t)
                assert out[_cnot_bits(n, x, c, t)] == 1 and np.abs(out).sum() == 1


def _dense_cnot(n, c, t):
    P0, P1 = np.diag([1.0, 0]), np.diag([0, 1.0])
    a = [I2] * n
    a[c] = P0
    b = [I2] * n
    b[c], b[t] = P1, PX

    def kron_all(ms):
        m = np.eye(1)
        for x in ms:
            m = np.kron(m, x)
        return m
    return kron_all(a) + kron_all(b)


@pytest.mark.parametrize('n', [2, 3, 4, 5, 6])
def test_ladder_reference_product_order(n):
    """L0 = C01 C23 ... C12 C34 ... as a matrix product (state.py:235-239); L1 the reverse order."""
    even = [_dense_cnot(n, i, i + 1) for i in range(0, n - 1, 2)]
    odd = [_dense_cnot(n, i, i + 1) for i in range(1, n - 1, 2)]
    L0 = np.eye(2 ** n)
    for m in even + odd:
        L0 = L0 @ m
    L1 = np.eye(2 ** n)
    for m in odd + even:
        L1 = L1 @ m
    v = rand_state(n, n)
    assert np.allclose(sv.cnot_ladder(v.copy(), 0), L0 @ v)
    assert np.allclose(sv.cnot_ladder(v.copy(), 1), L1 @ v)
    assert np.allclose(sv.cnot_ladder(sv.cnot_ladder(v.copy(), 0), 1), v)
    # and the CSR restatement agrees
    l0, l1 = ra.Gates(n).ladder()
    assert np.allclose(l0.toarray(), L0) and np.allclose(l1.toarray(), L1)


def test_periodic_ladder_odd_raises():
    with pytest.raises(ValueError):
        sv.cnot_ladder(rand_state(3, 0), 0, periodic=True)


@pytest.mark.parametrize('n', [2, 3, 5])
def test_cz_ladder_dense(n):
    d = np.ones(2 ** n)
    for i in range(n - 1):  # CZ on (i, i+1) = diag(1, 1, 1, -1) on that pair
        d *= np.diag(np.kron(np.kron(np.eye(2 ** i), np.diag([1, 1, 1, -1])), np.eye(2 ** (n - i - 2))))
    v = rand_state(n, 5)
    assert np.allclose(sv.cz_ladder(v.copy()), d * v)


# ------------------------------------------------------------- observable ---
def test_observable_dense_sum():
    n = 5
    rs = np.random.RandomState(2)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    obs['x'] = np.array([0.3, None, -0.2, None, 0.1], dtype=object)
    obs['y'] = np.array([None, 0.4, None, -0.6, None], dtype=object)
    obs['z'] = np.array([0.2, None, None, None, -0.5], dtype=object)
    M = np.zeros((2 ** n, 2 ** n), dtype=complex)
    for key, P in (('x', PX), ('y', PY), ('z', PZ)):
        for i, w in enumerate(obs[key]):
            if w is not None:
                M += w * op1(n, i, P)
    for i in range(n - 1):
        M += obs['zz'][i, i + 1] * op1(n, i, PZ) @ op1(n, i + 1, PZ)
    v = rand_state(n, 1)
    assert np.allclose(sv.apply_observable(v.copy(), obs), M @ v)
    assert np.allclose(ra.observable_matrix(obs).toarray(), M)
    assert abs(sv.expec_val(v, obs) - np.vdot(v, M @ v).real) < 1e-13


def test_lower_triangular_zz_raises():
    zz = np.full((3, 3), None)
    zz[2, 0] = 1.0
    with pytest.raises(ValueError):
        sv.observable_terms({'zz': zz})


# ---------------------------------------------------------------- McClean ---
def _mcclean_dense_E(n, axes, angles, obs, entangler='cnot'):
    """E by full dense unitaries, built without any oracle gate function."""
    psi = np.zeros(2 ** n, dtype=complex)
    psi[0] = 1
    for q in range(n):
        psi = op1(n, q, dense_rot(1, np.pi / 4)) @ psi
    even = [_dense_cnot(n, i, i + 1) for i in range(0, n - 1, 2)]
    odd = [_dense_cnot(n, i, i + 1) for i in range(1, n - 1, 2)]
    L0 = np.eye(2 ** n)
    for m in even + odd:
        L0 = L0 @ m
    for i in range(axes.shape[0]):
        psi = L0 @ psi
        for q in range(n):
            psi = op1(n, q, dense_rot(axes[i, q], angles[i, q])) @ psi
    M = ra.observable_matrix(obs).toarray()
    return np.vdot(psi, M @ psi).real


@pytest.mark.parametrize('n,L,seed', [(2, 2, 0), (3, 3, 1), (4, 2, 2), (5, 3, 3)])
def test_mcclean_forward_dense(n, L, seed):
    rs = np.random.RandomState(seed)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    assert abs(sv.mcclean_run_expec_val(n, axes, angles, obs) - _mcclean_dense_E(n, axes, angles, obs)) < 1e-13


@pytest.mark.parametrize('n,L,seed', [(2, 3, 0), (4, 3, 1), (6, 4, 2), (7, 2, 3)])
@pytest.mark.parametrize('entangler', ['cnot', 'cz'])
def test_mcclean_gradient_parameter_shift(n, L, seed, entangler):
    rs = np.random.RandomState(seed)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    obs['x'] = np.array([0.3] + [None] * (n - 1), dtype=object)
    E, g = sv.mcclean_grad_run(n, axes, angles, obs, entangler=entangler)
    E2, g2 = sv.mcclean_grad_run_adjoint(n, axes, angles, obs, entangler=entangler)
    ps = np.zeros_like(g)
    for i in range(L):
        for q in range(n):
            a1, a2 = angles.copy(), angles.copy()
            a1[i, q] += np.pi / 2
            a2[i, q] -= np.pi / 2
            ps[i, q] = 0.5 * (sv.mcclean_run_expec_val(n, axes, a1, obs, entangler=entangler)
                              - sv.mcclean_run_expec_val(n, axes, a2, obs, entangler=entangler))
    assert abs(E - sv.mcclean_run_expec_val(n, axes, angles, obs, entangler=entangler)) < 1e-13
    assert np.abs(g - ps).max() < 1e-12
    assert np.abs(g2 - ps).max() < 1e-12 and abs(E2 - E) < 1e-13


def test_mcclean_reference_formulation_agrees():
    n, L = 6, 4
    rs = np.random.RandomState(9)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    E, g = sv.mcclean_grad_run(n, axes, angles, obs)
    Er, gr = ra.McCleanRef(n, obs, axes, angles).grad_run()
    assert abs(E - Er) < 1e-13 and np.abs(g - gr).max() < 1e-13


def test_mcclean_analytic_zeros():
    """Last-layer Z-axis parameters commute with a diagonal observable: gradient 0."""
    n, L = 5, 3
    rs = np.random.RandomState(4)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    axes[-1, :] = [2, 0, 2, 1, 2]
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    _, g = sv.mcclean_grad_run(n, axes, angles, obs)
    assert np.abs(g[-1, [0, 2, 4]]).max() < 1e-14
    assert np.abs(g[-1, [1, 3]]).max() > 1e-6


def test_mcclean_ini_state_still_gets_ry_layer():
    n = 3
    v = rand_state(n, 0)
    prep = sv.mcclean_prepare(n, v)
    want = v
    for q in range(n):
        want = op1(n, q, dense_rot(1, np.pi / 4)) @ want
    assert np.allclose(prep, want)


def test_rng_order_axes_then_angles():
    rs1, rs2 = np.random.RandomState(17), np.random.RandomState(17)
    axes, angles = sv.mcclean_random_params(3, 4, rs1)
    u1, u2 = rs2.rand(3, 4), rs2.rand(3, 4)
    assert np.array_equal(axes, (3 * u1).astype(int)) and np.array_equal(angles, 2 * np.pi * u2)


# ------------------------------------------------------------------- QAOA ---
K33 = [(a, b) for a in range(3) for b in range(3, 6)]
CUBE = sorted({(min(a, a ^ (1 << k)), max(a, a ^ (1 << k))) for a in range(8) for k in range(3)})


@pytest.mark.parametrize('n,edges', [(6, K33), (8, CUBE)])
def test_qaoa_p1_closed_form_and_gradient(n, edges):
    """E = |E| sin(2b) sin(2g) cos^2(2g) for triangle-free 3-regular graphs."""
    obs = sv.maxcut_observable(n, edges)
    m = len(edges)
    for b, g in [(0.3, 0.7), (1.1, -0.4), (2.0, 1.3)]:
        E, grad = sv.qaoa_grad_run(n, obs, [b], [g])
        E2, grad2 = sv.qaoa_grad_run_adjoint(n, obs, [b], [g])
        Ec = m * np.sin(2 * b) * np.sin(2 * g) * np.cos(2 * g) ** 2
        dEb = m * 2 * np.cos(2 * b) * np.sin(2 * g) * np.cos(2 * g) ** 2
        dEg = m * np.sin(2 * b) * (2 * np.cos(2 * g) ** 3 - 4 * np.sin(2 * g) ** 2 * np.cos(2 * g))
        assert abs(E - Ec) < 1e-12 and abs(E2 - Ec) < 1e-12
        assert abs(grad[0, 0] - dEb) < 1e-12 and abs(grad[0, 1] - dEg) < 1e-12
        assert np.abs(grad2 - grad).max() < 1e-12


def test_qaoa_dense_and_finite_differences():
    n, p = 5, 3
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (0, 4), (1, 3)]
    obs = sv.maxcut_observable(n, edges)
    rs = np.random.RandomState(0)
    b, g = np.pi * rs.rand(p), np.pi * rs.rand(p)
    C = np.zeros(2 ** n)
    for (u, v) in edges:
        C += np.diag(op1(n, u, PZ) @ op1(n, v, PZ)).real
    B = sum(op1(n, q, PX) for q in range(n))
    psi = np.full(2 ** n, 2 ** (-n / 2), dtype=complex)
    for i in range(p):
        psi = np.exp(-1j * g[i] * C) * psi
        psi = la.expm(-0.5j * b[i] * B) @ psi
    E_dense = np.vdot(psi, C * psi).real
    E, grad = sv.qaoa_grad_run(n, obs, b, g)
    assert abs(E - E_dense) < 1e-12
    Er, gr = ra.QaoaRef(n, obs, p).grad_run(b, g)
    assert abs(Er - E) < 1e-12 and np.abs(gr - grad).max() < 1e-12
    eps = 1e-6
    for i in range(p):
        bp, bm = b.copy(), b.copy()
        bp[i] += eps
        bm[i] -= eps
        gp, gm = g.copy(), g.copy()
        gp[i] += eps
        gm[i] -= eps
        fdb = (sv.qaoa_run_expec_val(n, obs, bp, g) - sv.qaoa_run_expec_val(n, obs, bm, g)) / (2 * eps)
        fdg = (sv.qaoa_run_expec_val(n, obs, b, gp) - sv.qaoa_run_expec_val(n, obs, b, gm)) / (2 * eps)
        assert abs(fdb - grad[i, 0]) < 1e-7 and abs(fdg - grad[i, 1]) < 1e-7


def test_qaoa_ini_state_dense_and_finite_differences():
    """QAOA from a random initial state: dense unitaries for E, central
    finite differences for both gradient columns; the input is not modified."""
    n, p = 5, 2
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (0, 4)]
    obs = sv.maxcut_observable(n, edges)
    rs = np.random.RandomState(3)
    b, g = np.pi * rs.rand(p), np.pi * rs.rand(p)
    ini = rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)
    ini /= np.linalg.norm(ini)
    keep = ini.copy()
    C = np.zeros(2 ** n)
    for (u, v) in edges:
        C += np.diag(op1(n, u, PZ) @ op1(n, v, PZ)).real
    B = sum(op1(n, q, PX) for q in range(n))
    psi = ini.copy()
    for i in range(p):
        psi = la.expm(-0.5j * b[i] * B) @ (np.exp(-1j * g[i] * C) * psi)
    E, grad = sv.qaoa_grad_run(n, obs, b, g, ini_state=ini)
    assert abs(E - np.vdot(psi, C * psi).real) < 1e-12
    assert abs(sv.qaoa_run_expec_val(n, obs, b, g, ini_state=ini) - E) < 1e-12
    Ea, ga = sv.qaoa_grad_run_adjoint(n, obs, b, g, ini_state=ini)
    assert abs(Ea - E) < 1e-12 and np.abs(ga - grad).max() < 1e-12
    eps = 1e-6
    for i in range(p):
        d = np.zeros(p)
        d[i] = eps
        fdb = (sv.qaoa_run_expec_val(n, obs, b + d, g, ini) - sv.qaoa_run_expec_val(n, obs, b - d, g, ini)) / (2 * eps)
        fdg = (sv.qaoa_run_expec_val(n, obs, b, g + d, ini) - sv.qaoa_run_expec_val(n, obs, b, g - d, ini)) / (2 * eps)
        assert abs(fdb - grad[i, 0]) < 1e-7 and abs(fdg - grad[i, 1]) < 1e-7
    assert np.array_equal(ini, keep)


@pytest.mark.parametrize('seed', [0, 1])
def test_projector_probs_are_dense_expectations(seed):
    """|P^+ v|^2 = (|v|^2 + <v|P|v>) / 2 with <v|P|v> from dense Kronecker
    products, for every term kind, also for a state of norm != 1."""
    n = 4
    rs = np.random.RandomState(seed)
    v = (rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)) * 0.7
    nrm = np.vdot(v, v).real
    P = {'x': PX, 'y': PY, 'z': PZ}
    for q in range(n):
        for kind in ('x', 'y', 'z'):
            e = np.vdot(v, op1(n, q, P[kind]) @ v).real
            assert abs(sv.projector_prob(v, kind, q) - 0.5 * (nrm + e)) < 1e-12
        for j in range(q + 1, n):
            e = np.vdot(v, op1(n, q, PZ) @ op1(n, j, PZ) @ v).real
            assert abs(sv.projector_prob(v, 'zz', q, j) - 0.5 * (nrm + e)) < 1e-12


def test_projector_probs_special_states():
    s = 2 ** -0.5
    assert abs(sv.projector_prob(np.array([s, 1j * s]), 'y', 0) - 1.0) < 1e-15    # |+i>
    assert abs(sv.projector_prob(np.array([s, -s]), 'x', 0)) < 1e-15              # |->
    assert sv.projector_prob(np.array([0, 0, 1, 0], dtype=complex), 'z', 0) == 0.0  # |10>: qubit 0 is 1
    assert sv.projector_prob(np.array([0, 0, 1, 0], dtype=complex), 'z', 1) == 1.0
    assert sv.projector_prob(np.array([0, 0, 0, 1], dtype=complex), 'zz', 0, 1) == 1.0


def test_sampled_expectation_converges():
    """The shot estimate is unbiased: many shots land within 5 sigma of <O>."""
    n = 3
    rs = np.random.RandomState(4)
    v = rs.standard_normal(8) + 1j * rs.standard_normal(8)
    v /= np.linalg.norm(v)
    obs = {'x': np.array([0.5, None, None], dtype=object), 'z': np.array([None, -1.0, 0.25], dtype=object)}
    shots = 200000
    est = sv.sample_expec_val(v, obs, shots, np.random.RandomState(9))
    probs, w = sv.term_probs(v, obs)
    sigma = np.sqrt(np.sum(4 * w ** 2 * probs * (1 - probs) / shots))
    assert abs(est - sv.expec_val(v, obs)) < 5 * sigma


def test_complex64_model_error_is_bounded():
    """The complex64 model (FP32 gates) lands within 1e-5 of complex128 at small n."""
    n, L = 8, 6
    rs = np.random.RandomState(1)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    _, g = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    _, g32 = sv.mcclean_grad_run_adjoint(n, axes, angles, obs, dtype=np.complex64)
    err = np.linalg.norm(g32 - g) / np.linalg.norm(g)
    assert 1e-9 < err < 1e-5


def test_regular3_generator():
    rs = np.random.RandomState(0)
    for _ in range(5):
        e = sv.random_regular3_edges(20, rs)
        assert len(e) == 30 and all(a < b for a, b in e)
        deg = np.bincount(np.array(e).ravel(), minlength=20)
        assert (deg == 3).all()
