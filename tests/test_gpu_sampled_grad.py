"""GPU parity of the sampled McClean gradients (mc_clean.py:80-198).

* Projector probabilities of the parameter-shifted circuits
  (qg_mcclean_shift_term_prob) against the oracle's shifted states, formed as
  the reference forms them: complex128 within 1e-12, complex64 within 1e-6,
  in history and checkpoint plans, with ini_state, both entanglers.
* At 20 qubits, the parameter-shift identity against the GPU's own adjoint
  gradient (size-independent property).
* Whole seeded shot-based gradients equal to the oracle's scipy draws.
* grad_run_with_component_sampling against the oracle with the same drawn
  component, with the whole observable's E, including a checkpointed plan.
"""
import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = {'complex128': 1e-12, 'complex64': 1e-6}


def _case(n, L, seed, mixed=False):
    rs = np.random.RandomState(seed)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    if mixed:
        zz = np.full((n, n), None)
        zz[0, n - 1] = float(rs.standard_normal())
        zz[1, 2] = 0.8
        obs = {'x': np.array([0.5] + [None] * (n - 1), dtype=object),
               'y': np.array([None] * (n - 1) + [-0.3], dtype=object),
               'z': np.array([None, 1.2] + [None] * (n - 2), dtype=object), 'zz': zz}
    else:
        obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    return axes, angles, obs


def _oracle_shift_probs(n, axes, angles, obs, ini=None, entangler='cnot'):
    rows = []
    for _, _, vp, vm in sv.mcclean_shift_states(n, axes, angles, ini_state=ini, entangler=entangler):
        rows.append(sv.term_probs(vp, obs)[0])
        rows.append(sv.term_probs(vm, obs)[0])
    return np.array(rows)


def _all_shifts(L, n):
    return np.repeat(np.arange(L * n), 2), np.tile([np.pi / 2, -np.pi / 2], L * n)


@pytest.mark.parametrize('n,L,seed,ckpt,entangler', [(5, 3, 0, 0, 'cnot'), (5, 3, 1, 0, 'cz'), (12, 2, 2, 0, 'cnot'),
                                                     (15, 3, 3, 0, 'cnot'), (15, 4, 4, 2, 'cnot'),
                                                     (15, 3, 5, 1, 'cz')])
@pytest.mark.parametrize('dtype', ['complex128', 'complex64'])
def test_shift_probs_match_oracle(gpu, n, L, seed, ckpt, entangler, dtype):
    from qradient_amd.circuit_logic import McClean
    axes, angles, obs = _case(n, L, seed, mixed=True)
    print('case', n, L, seed, ckpt, entangler, dtype, flush=True)
    circ = McClean(n, obs, L, axes=axes, angles=angles, dtype=dtype, entangler=entangler, ckpt_every=ckpt)
    params, deltas = _all_shifts(L, n)
    got = circ.shifted_term_probs(params, deltas, circ.observable.terms)
    want = _oracle_shift_probs(n, axes, angles, obs, entangler=entangler)
    assert got.shape == want.shape
    assert np.max(np.abs(got - want)) <= TOL[dtype]
    # the plan keeps the unshifted circuit's final state
    v = sv.mcclean_forward(n, axes, angles, entangler=entangler)
    assert abs(circ.expec_val() - sv.expec_val(v, obs)) <= TOL[dtype]


@pytest.mark.parametrize('dtype', ['complex128', 'complex64'])
def test_shift_probs_with_ini_state(gpu, dtype):
    from qradient_amd.circuit_logic import McClean
    n, L = 14, 2
    axes, angles, obs = _case(n, L, 6)
    rs = np.random.RandomState(60)
    ini = rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)
    ini /= np.linalg.norm(ini)
    circ = McClean(n, obs, L, axes=axes, angles=angles, dtype=dtype)
    params, deltas = _all_shifts(L, n)
    got = circ.shifted_term_probs(params, deltas, circ.observable.terms, ini_state=ini)
    want = _oracle_shift_probs(n, axes, angles, obs, ini=ini)
    assert np.max(np.abs(got - want)) <= TOL[dtype]


@pytest.mark.parametrize('dtype', ['complex128', 'complex64'])
def test_shift_identity_at_20_qubits(gpu, dtype):
    """(<O>_+ - <O>_-) / 2 from the shifted probabilities equals the adjoint
    gradient of grad_run, for every parameter of a 20-qubit circuit."""
    from qradient_amd.circuit_logic import McClean
    n, L = 20, 3
    axes, angles, obs = _case(n, L, 7)
    circ = McClean(n, obs, L, axes=axes, angles=angles, dtype=dtype)
    _, g = circ.grad_run()
    params, deltas = _all_shifts(L, n)
    terms = circ.observable.terms
    probs = circ.shifted_term_probs(params, deltas, terms)
    w = np.array([t[3] for t in terms])
    e = (2 * probs - 1) @ w          # <O> of each shifted state (norm 1)
    got = .5 * (e[0::2] - e[1::2]).reshape(L, n)
    assert rel_err(got, g) <= (1e-10 if dtype == 'complex128' else 2e-6)


@pytest.mark.parametrize('n,L,components,exact', [(6, 2, False, True), (6, 2, False, False), (6, 2, True, True),
                                                  (14, 2, True, False)])
def test_sample_grad_seeded_equals_oracle(gpu, n, L, components, exact):
    from qradient_amd.circuit_logic import McClean
    axes, angles, obs = _case(n, L, 8 + n, mixed=True)
    circ = McClean(n, obs, L, axes=axes, angles=angles, use_observable_components=components)
    shots = 40
    np.random.seed(5)
    if components:
        E, g = circ.sample_grad_with_component_sampling(shot_num=shots, exact_expec_val=exact)
    else:
        E, g = circ.sample_grad(shot_num=shots, exact_expec_val=exact)
    after = np.random.random_sample()
    np.random.seed(5)
    Eo, go = sv.mcclean_sample_grad(n, axes, angles, obs, shots, exact_expec_val=exact,
                                    component_sampling=components)
    assert np.random.random_sample() == after
    assert abs(E - Eo) <= 1e-12
    assert np.array_equal(g, go)


@pytest.mark.parametrize('n,L,dtype,ckpt,mixed', [(8, 3, 'complex128', 0, True), (16, 3, 'complex128', 0, False),
                                                  (16, 3, 'complex64', 0, False), (16, 4, 'complex128', 2, False),
                                                  (16, 4, 'complex64', 1, False)])
def test_component_sampling_matches_oracle(gpu, n, L, dtype, ckpt, mixed):
    from qradient_amd.circuit_logic import McClean
    axes, angles, obs = _case(n, L, 20 + n + ckpt, mixed=mixed)
    print('case', n, L, dtype, ckpt, mixed, flush=True)
    circ = McClean(n, obs, L, axes=axes, angles=angles, dtype=dtype, ckpt_every=ckpt,
                   use_observable_components=True)
    for seed in (1, 2):
        np.random.seed(seed)
        E, g = circ.grad_run_with_component_sampling()
        after = np.random.random_sample()
        np.random.seed(seed)
        k = np.random.choice(np.arange(circ.observable.num_components), p=circ.observable.weight_distribution)
        assert np.random.random_sample() == after
        Eo, go = sv.mcclean_grad_run_component(n, axes, angles, obs, k)
        tol = TOL[dtype]
        assert abs(E - Eo) <= tol * max(1.0, abs(Eo))
        assert rel_err(g, go) <= tol


def test_optimiser_steps_with_shots(gpu):
    from qradient_amd.circuit_logic import McClean
    from qradient_amd.optimization import McCleanOpt
    n, L = 5, 2
    axes, angles, obs = _case(n, L, 30)
    circ = McClean(n, obs, L, axes=axes, angles=angles.copy(), use_observable_components=True)
    opt = McCleanOpt(circ, {'name': 'GradientDescent', 'step_size': 0.1}, max_iter=5)
    np.random.seed(3)
    opt.step(shot_num=20, dense_mode=False)
    opt.step(shot_num=20, dense_mode=False, component_sampling=True)
    opt.step(component_sampling=True)
    # replay with the oracle: same draws, same updates
    np.random.seed(3)
    a = angles.copy()
    for kind in ('shots', 'shots_comp', 'comp'):
        if kind == 'comp':
            dist = sv.weight_distribution(obs)
            k = np.random.choice(np.arange(dist.size), p=dist)
            _, g = sv.mcclean_grad_run_component(n, axes, a, obs, k)
        else:
            _, g = sv.mcclean_sample_grad(n, axes, a, obs, 20, component_sampling=(kind == 'shots_comp'))
        a = a - 0.1 * g
    assert np.max(np.abs(opt.param_history[3] - a)) <= 1e-10
