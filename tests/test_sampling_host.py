"""The host sampler of the finite-shot expectation (base.py:22-34) draws what
the reference draws: for the same probabilities and the same seed it gives the
oracle's scipy rv_discrete estimate exactly (CPU only)."""
import numpy as np
import pytest

from oracle import statevector as sv
from qradient_amd.circuit_logic.base import sample_projectors


def _case(seed):
    n = 5
    rs = np.random.RandomState(seed)
    v = rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)
    v /= np.linalg.norm(v)
    zz = np.full((n, n), None)
    zz[0, 3] = -0.7
    zz[1, 2] = 1.0
    obs = {'x': np.array([0.5, None, None, None, 2.0], dtype=object),
           'y': np.array([None, 0.3, None, None, None], dtype=object),
           'z': np.array([None, None, -1.5, None, None], dtype=object), 'zz': zz}
    return v, obs


@pytest.mark.parametrize('seed,shots', [(0, 1), (1, 17), (2, 1000)])
def test_same_draws_as_scipy_rv_discrete(seed, shots):
    v, obs = _case(seed)
    probs, w = sv.term_probs(v, obs)
    ours = sample_projectors(probs, w, shots, np.random.RandomState(100 + seed))
    ref = sv.sample_expec_val(v, obs, shots, np.random.RandomState(100 + seed))
    assert ours == ref


def test_global_random_state_like_the_reference():
    v, obs = _case(3)
    probs, w = sv.term_probs(v, obs)
    np.random.seed(42)
    ours = sample_projectors(probs, w, 500)
    after_ours = np.random.random_sample()
    np.random.seed(42)
    ref = sv.sample_expec_val(v, obs, 500)
    after_ref = np.random.random_sample()
    assert ours == ref and after_ours == after_ref


def test_edge_probabilities():
    rs = np.random.RandomState(0)
    assert sample_projectors([1.0, 0.0], [2.0, 3.0], 50, rs) == 2.0 - 3.0
    # rounding just outside [0, 1] is clipped
    assert sample_projectors([1.0 + 1e-16, -1e-17], [1.0, 1.0], 10, rs) == 0.0
    with pytest.raises(ValueError):
        sample_projectors([0.5], [1.0], 0)


@pytest.mark.parametrize('seed,shots', [(0, 1), (1, 13), (2, 200)])
def test_vectorised_estimates_equal_the_per_term_sampler(seed, shots):
    """projector_estimates over many states at once, with one uniform draw of
    all their shots, equals sample_projectors state after state (the same
    random numbers in the same order, the same sums)."""
    from qradient_amd.circuit_logic.base import projector_estimates
    rs = np.random.RandomState(seed)
    M, K = 6, 5
    probs = rs.uniform(size=(M, K))
    probs[0, 0], probs[1, 1] = 0.0, 1.0            # the edges of [0, 1]
    w = rs.standard_normal(K)
    one = [sample_projectors(probs[m], w, shots, np.random.RandomState(50 + seed)) for m in range(1)]
    rs_a = np.random.RandomState(50 + seed)
    per_state = np.array([sample_projectors(probs[m], w, shots, rs_a) for m in range(M)])
    rs_b = np.random.RandomState(50 + seed)
    u = rs_b.uniform(size=(M, K, shots))
    assert np.array_equal(projector_estimates(probs, w, u), per_state)
    assert per_state[0] == one[0]
