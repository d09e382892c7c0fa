"""CPU pins of the sampled McClean gradients (mc_clean.py:80-198) and the
host side of their GPU path.

Oracle pins (each fails for a plausible slip: a wrong shift sign, restarting
from the pre-rotation history, a dropped weight, the wrong component):
* the shifted states of sample_grad, formed as the reference forms them
  (restart from state_history[i] plus a manual rotation), equal full forward
  runs with angles[i, q] shifted by +-pi/2;
* their exact expectations satisfy the parameter-shift identity against the
  adjoint gradient (itself pinned by brute force in test_oracle_pins.py);
* component gradients are linear: the sum over components equals the whole
  observable's gradient;
* the shot estimates converge to the exact gradient (statistical bound);
* the component distribution is the reference's w / sum(w) for non-negative
  weights.
Host pins: McClean.sample_grad / sample_grad_with_component_sampling, given
the oracle's probabilities in place of the device's, return exactly the
oracle's scipy-drawn estimates for the same seed (same random numbers in the
same order).
"""
import numpy as np
import pytest

from oracle import statevector as sv


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


@pytest.mark.parametrize('entangler', ['cnot', 'cz'])
def test_shift_states_are_shifted_angle_runs(entangler):
    n, L = 4, 3
    axes, angles, _ = _case(n, L, 1)
    for i, q, vp, vm in sv.mcclean_shift_states(n, axes, angles, entangler=entangler):
        for v, d in ((vp, np.pi / 2), (vm, -np.pi / 2)):
            a = angles.copy()
            a[i, q] += d
            w = sv.mcclean_forward(n, axes, a, entangler=entangler)
            assert np.max(np.abs(v - w)) < 1e-13


@pytest.mark.parametrize('mixed', [False, True])
def test_shift_states_parameter_shift_identity(mixed):
    n, L = 4, 3
    axes, angles, obs = _case(n, L, 2, mixed)
    _, g = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    got = np.empty_like(g)
    for i, q, vp, vm in sv.mcclean_shift_states(n, axes, angles):
        got[i, q] = .5 * (sv.expec_val(vp, obs) - sv.expec_val(vm, obs))
    assert np.max(np.abs(got - g)) < 1e-12


def test_shift_states_with_ini_state():
    n, L = 3, 2
    axes, angles, obs = _case(n, L, 3)
    rs = np.random.RandomState(9)
    ini = rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)
    ini /= np.linalg.norm(ini)
    _, g = sv.mcclean_grad_run(n, axes, angles, obs, ini_state=ini)
    for i, q, vp, vm in sv.mcclean_shift_states(n, axes, angles, ini_state=ini):
        assert abs(.5 * (sv.expec_val(vp, obs) - sv.expec_val(vm, obs)) - g[i, q]) < 1e-12


@pytest.mark.parametrize('mixed', [False, True])
def test_component_gradients_are_linear(mixed):
    n, L = 4, 3
    axes, angles, obs = _case(n, L, 4, mixed)
    E, g = sv.mcclean_grad_run(n, axes, angles, obs)
    _, terms = sv.observable_terms(obs)
    total = np.zeros_like(g)
    for k in range(len(terms)):
        Ek, gk = sv.mcclean_grad_run_component(n, axes, angles, obs, k)
        assert abs(Ek - E) < 1e-12          # E is the whole observable's
        total += gk
    assert np.max(np.abs(total - g)) < 1e-12


def test_component_weights_keep_the_reference_distribution():
    n = 4
    zz = np.full((n, n), None)
    zz[0, 1], zz[1, 2], zz[2, 3] = 0.5, 1.5, 2.0
    obs = {'zz': zz}
    w = np.array([0.5, 1.5, 2.0])
    assert np.array_equal(sv.weight_distribution(obs), w / w.sum())


def test_sample_grad_converges():
    n, L = 3, 2
    axes, angles, obs = _case(n, L, 5)
    _, g = sv.mcclean_grad_run(n, axes, angles, obs)
    np.random.seed(0)
    shots = 20000
    _, gs = sv.mcclean_sample_grad(n, axes, angles, obs, shots)
    # each estimate is a difference of two means of +-w draws, over n-1 terms
    wsum = sum(abs(t[3]) for t in sv.observable_terms(obs)[1])
    assert np.max(np.abs(gs - g)) < 6 * wsum / np.sqrt(shots)


# ---------------------------------------------------------------------------
# host side: the GPU path's draws, with oracle probabilities standing in for
# the device's
# ---------------------------------------------------------------------------
def _host_circuit(n, L, axes, angles, obs, components):
    from qradient_amd.circuit_logic import McClean

    circ = McClean(n, obs, L, axes=axes, angles=angles, use_observable_components=components)

    def shifted_term_probs(params, deltas, terms, ini_state=None):
        rows = []
        for p, d in zip(params, deltas):
            a = angles.copy()
            a.flat[p] += d
            rows.append(sv.term_probs(sv.mcclean_forward(n, axes, a), obs)[0])
        circ._final = sv.mcclean_forward(n, axes, angles)
        return np.array(rows)

    circ.shifted_term_probs = shifted_term_probs
    circ.expec_val = lambda: sv.expec_val(circ._final, obs)
    circ._term_probs = lambda terms: sv.term_probs(circ._final, obs)[0] if len(terms) > 1 else \
        sv.term_probs(circ._final, obs)[0][[obs_index(obs, terms[0])]]
    return circ


def obs_index(obs, term):
    from qradient_amd.physical_components import Observable
    return Observable(obs).terms.index(term)


@pytest.mark.parametrize('exact', [True, False])
@pytest.mark.parametrize('components', [False, True])
def test_host_draws_equal_the_oracle(exact, components):
    n, L = 3, 2
    axes, angles, obs = _case(n, L, 6, mixed=True)
    shots = 7
    for seed in (0, 1):
        circ = _host_circuit(n, L, axes, angles, obs, components)
        np.random.seed(seed)
        if components:
            E, g = circ.sample_grad_with_component_sampling(shot_num=shots, exact_expec_val=exact)
        else:
            E, g = circ.sample_grad(shot_num=shots, exact_expec_val=exact)
        after = np.random.random_sample()
        np.random.seed(seed)
        Eo, go = sv.mcclean_sample_grad(n, axes, angles, obs, shots, exact_expec_val=exact,
                                        component_sampling=components)
        assert np.random.random_sample() == after      # same count of random numbers
        assert abs(E - Eo) < 1e-12
        assert np.array_equal(g, go)
