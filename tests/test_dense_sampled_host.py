"""CPU pins of the dense sampled gradients (mc_clean.py:207-350,
qaoa.py:83-158) and of the reformulation the GPU path uses.

* With exact means in place of shot means, the reference's dense formulation
  (eigh eigenbasis, left-hand-side suffix matrices) gives the adjoint
  gradient: McClean for any observable, QAOA MaxCut (unit weights, where the
  reference's pi/4 component shift is the exact shift rule), and one sampled
  component's gradient in component mode.  A wrong shift, suffix order or
  history slot fails these.
* The GPU reformulation, restated in numpy (shifted-angle circuits, basis
  sorted by D(x), inverse CDF at the same uniforms), draws what the dense
  formulation draws for the same seed.
* The shot means converge to the exact gradient.
"""
import numpy as np
import pytest

from oracle import reference_algorithm as ra
from oracle import statevector as sv


def _mcclean_case(n, L, seed, mixed=False, positive=False):
    rs = np.random.RandomState(seed)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    w = rs.standard_normal(n - 1)
    if positive:
        w = np.abs(w) + 0.1
    obs = sv.zz_chain_observable(n, w)
    if mixed:
        obs['x'] = np.array([0.4] + [None] * (n - 1), dtype=object)
        obs['z'] = np.array([None] * (n - 1) + [-0.6], dtype=object)
    return axes, angles, obs


def _maxcut(n, edges):
    zz = np.full((n, n), None)
    for a, b in edges:
        zz[a, b] = 1.0
    return {'zz': zz}


@pytest.mark.parametrize('mixed', [False, True])
def test_mcclean_dense_exact_means_are_the_gradient(mixed):
    n, L = 4, 3
    axes, angles, obs = _mcclean_case(n, L, 1, mixed)
    E, g = ra.mcclean_sample_grad_dense(n, obs, axes, angles, 1, exact_means=True)
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    assert abs(E - Eo) < 1e-12
    assert np.max(np.abs(g - go)) < 1e-12


def test_mcclean_dense_component_exact_means():
    n, L = 4, 2
    axes, angles, obs = _mcclean_case(n, L, 2, positive=True)
    np.random.seed(4)
    _, g = ra.mcclean_sample_grad_dense(n, obs, axes, angles, 1, component_sampling=True, exact_means=True)
    np.random.seed(4)
    c = np.random.choice(np.arange(n - 1), p=sv.weight_distribution(obs))
    _, gc = sv.mcclean_grad_run_component(n, axes, angles, obs, c)
    assert np.max(np.abs(g - gc)) < 1e-12


@pytest.mark.parametrize('p', [1, 2])
def test_qaoa_dense_exact_means_are_the_gradient(p):
    n = 5
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (0, 4), (1, 3)]
    obs = _maxcut(n, edges)
    rs = np.random.RandomState(p)
    betas, gammas = np.pi * rs.rand(p), np.pi * rs.rand(p)
    E, g = ra.qaoa_sample_grad_dense(n, obs, betas, gammas, 1, exact_means=True)
    Eo, go = sv.qaoa_grad_run(n, obs, betas, gammas)
    assert abs(E - Eo) < 1e-12
    assert np.max(np.abs(g - go)) < 1e-12


def _reformulated_mcclean(n, axes, angles, obs, shot_num, exact_expec_val, component):
    """The GPU path's algorithm in numpy: shifted-angle circuits, the basis
    sorted by D(x) (or the component's two eigenvalue blocks), inverse CDF."""
    L = axes.shape[0]
    _, terms = sv.observable_terms(obs)
    if component:
        c = np.random.choice(np.arange(len(terms)), p=sv.weight_distribution(obs))
    final = sv.mcclean_forward(n, axes, angles)
    if exact_expec_val:
        E = sv.expec_val(final, obs)
    else:
        E = sv.sample_expec_val(final, obs, shot_num)
    D = sv.diagonal(n, terms)
    order = np.argsort(D, kind='stable')
    grad = np.empty((L, n))
    for k in range(L * n):
        means = []
        for d in (np.pi / 2, -np.pi / 2):
            a = angles.copy()
            a.flat[k] += d
            v = sv.mcclean_forward(n, axes, a)
            u = np.random.uniform(size=shot_num)
            if component:
                kind, i, j, w = terms[c]
                p_plus = sv.projector_prob(v, kind, i, j)
                m_minus = 1 - p_plus if w > 0 else p_plus
                means.append(np.where(u <= m_minus, -abs(w), abs(w)).mean())
            else:
                cdf = np.cumsum(np.abs(v[order]) ** 2)
                idx = np.array([np.argmax(cdf >= x) for x in u])
                means.append(D[order][idx].mean())
        grad.flat[k] = (means[0] - means[1]) / 2
    return E, grad


@pytest.mark.parametrize('exact,component', [(True, False), (False, False), (True, True), (False, True)])
def test_mcclean_reformulation_draws_like_the_dense_reference(exact, component):
    n, L = 4, 2
    axes, angles, obs = _mcclean_case(n, L, 5, positive=component)
    shots = 25
    np.random.seed(11)
    E, g = _reformulated_mcclean(n, axes, angles, obs, shots, exact, component)
    after = np.random.random_sample()
    np.random.seed(11)
    Eo, go = ra.mcclean_sample_grad_dense(n, obs, axes, angles, shots, exact_expec_val=exact,
                                          component_sampling=component)
    assert np.random.random_sample() == after
    assert abs(E - Eo) < 1e-12
    assert np.max(np.abs(g - go)) < 1e-12


def test_qaoa_dense_shot_means_converge():
    n = 4
    obs = _maxcut(n, [(0, 1), (1, 2), (2, 3), (0, 3)])
    betas, gammas = np.array([0.7]), np.array([0.4])
    _, go = sv.qaoa_grad_run(n, obs, betas, gammas)
    np.random.seed(2)
    shots = 20000
    _, g = ra.qaoa_sample_grad_dense(n, obs, betas, gammas, shots)
    # |C| <= 4: each mean has sd <= 4 / sqrt(shots); 4 components + 4 qubits
    assert np.max(np.abs(g - go)) < 6 * 8 * 4 / np.sqrt(shots)
