"""GPU parity of the finite-shot expectation values (base.py:22-46,
mc_clean.py:27-45, qaoa.py:23-38): the projector probabilities against the
oracle's, and whole shot estimates against the oracle's scipy rv_discrete
draws from the same seed."""
import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _mixed_obs(n, rs):
    """x / y / z terms on about half the qubits (first and last included) and
    a zz term per qubit: more than one launch's worth of terms for n >= 9."""
    def pick(scale):
        return np.array([scale * rs.standard_normal() if (q % 2 == 0 or q == n - 1) else None
                         for q in range(n)], dtype=object)
    obs = {'x': pick(0.5), 'y': pick(0.7), 'z': pick(1.0)}
    if n >= 2:
        zz = np.full((n, n), None)
        for i in range(n - 1):
            zz[i, (i + 1 + 3 * i) % (n - i - 1) + i + 1] = float(rs.standard_normal())
        obs['zz'] = zz
    return obs


@pytest.mark.parametrize('n', [1, 3, 9, 17])
@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-12), ('complex64', 1e-6)])
def test_state_term_probs(gpu, n, dtype, tol):
    from qradient_amd.physical_components import Observable, State
    rs = np.random.RandomState(n)
    v = rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)
    v *= 0.8 / np.linalg.norm(v)   # norm 0.8: the probabilities keep |v|^2
    obs = _mixed_obs(n, rs)
    st = State(n, dtype=dtype)
    st.vec = v
    got = st.term_probs(Observable(obs))
    want, _ = sv.term_probs(v.astype(np.complex64) if dtype == 'complex64' else v, obs)
    assert got.shape == want.shape and np.max(np.abs(got - want)) <= tol


def _seeded(fn, seed):
    np.random.seed(seed)
    return fn()


@pytest.mark.parametrize('n,L', [(6, 3), (16, 2)])
@pytest.mark.parametrize('dtype', ['complex128', 'complex64'])
def test_mcclean_shot_estimate(gpu, n, L, dtype):
    from qradient_amd.circuit_logic import McClean
    rs = np.random.RandomState(30 + n)
    obs = _mixed_obs(n, rs)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    circ = McClean(n, obs, L, axes=axes, angles=angles, dtype=dtype)
    shots = 1000
    got = _seeded(lambda: circ.run_expec_val(exact_expec_val=False, shot_num=shots), 7)
    v = sv.mcclean_forward(n, axes, angles)
    ref = _seeded(lambda: sv.sample_expec_val(v, obs, shots), 7)
    wmax = max(abs(t[3]) for t in sv.observable_terms(obs)[1])
    # complex128: the same draws on probabilities equal to ~1e-15; complex64:
    # a draw within ~1e-7 of a probability may land on the other side
    tol = 1e-12 if dtype == 'complex128' else 4 * wmax / shots + 1e-6
    assert abs(got - ref) <= tol, (got, ref)
    # the state stays current: the sampled probabilities are the final state's
    probs, _ = sv.term_probs(v, obs)
    assert np.max(np.abs(circ._term_probs(circ.observable.terms) - probs)) <= (1e-12 if dtype == 'complex128' else 1e-6)


@pytest.mark.parametrize('n,p', [(8, 3), (16, 2)])
def test_qaoa_shot_estimate_and_state(gpu, n, p):
    from qradient_amd.circuit_logic import Qaoa
    rs = np.random.RandomState(50 + n)
    edges = sv.random_regular3_edges(n, rs)
    obs = sv.maxcut_observable(n, edges)
    b, g = np.pi * rs.rand(p), np.pi * rs.rand(p)
    circ = Qaoa(n, obs, p)
    shots = 2000
    got = _seeded(lambda: circ.run_expec_val(b, g, exact_expec_val=False, shot_num=shots), 11)
    v = sv.qaoa_forward(n, obs, b, g)
    ref = _seeded(lambda: sv.sample_expec_val(v, obs, shots), 11)
    assert abs(got - ref) <= 1e-12, (got, ref)
    # after run_expec_val, circuit.state is the final state (qaoa.py:23-38)
    assert rel_err(circ.state.vec.view(np.float64), v.view(np.float64)) <= 1e-12
    assert abs(circ.expec_val() - sv.expec_val(v, obs)) <= 1e-12 * max(1.0, abs(sv.expec_val(v, obs)))


def test_mcclean_state_and_component_sampling(gpu):
    from qradient_amd.circuit_logic import McClean
    n, L = 10, 3
    rs = np.random.RandomState(4)
    obs = _mixed_obs(n, rs)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    circ = McClean(n, obs, L, axes=axes, angles=angles)
    E = circ.run_expec_val()
    v = sv.mcclean_forward(n, axes, angles)
    assert rel_err(circ.state.vec.view(np.float64), v.view(np.float64)) <= 1e-12
    assert abs(circ.expec_val() - E) <= 1e-12 * max(1.0, abs(E))
    _, terms = sv.observable_terms(obs)
    for comp in (0, len(terms) - 1):
        kind, i, j, w = terms[comp]
        got = _seeded(lambda: circ.sample_component_expec_val(500, comp), 3 + comp)
        one = {kind: np.full((n, n) if kind == 'zz' else n, None, dtype=object)}
        if kind == 'zz':
            one['zz'][i, j] = w
        else:
            one[kind][i] = w
        ref = _seeded(lambda: sv.sample_expec_val(v, one, 500), 3 + comp)
        assert abs(got - ref) <= 1e-12
    # after grad_run the plan's state is no longer the current one
    circ.grad_run()
    st = circ.state.vec
    assert rel_err(st.view(np.float64), v.view(np.float64)) <= 1e-12
