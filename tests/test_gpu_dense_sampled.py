"""GPU parity of the dense sampled gradients (mc_clean.py:207-350,
qaoa.py:83-158) and of the basis sampler behind them.

Seeded runs draw the same shots as the reference's dense formulation (the
oracle's eigh + left-hand-side matrices + scipy rv_discrete), so E and every
gradient entry agree to 1e-12 (complex128 states); larger circuits are
checked against the numpy restatement of the GPU's reformulation.
"""
import numpy as np
import pytest

from oracle import reference_algorithm as ra
from oracle import statevector as sv
from tests.test_dense_sampled_host import _maxcut, _mcclean_case, _reformulated_mcclean

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('n', [3, 10, 14])
@pytest.mark.parametrize('with_order', [False, True])
def test_state_sample_basis_is_the_inverse_cdf(gpu, n, with_order):
    from qradient_amd.physical_components import State
    rs = np.random.RandomState(n)
    v = rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)
    v /= np.linalg.norm(v)
    st = State(n)
    st.vec = v
    order = rs.permutation(2 ** n).astype(np.int32) if with_order else None
    u = rs.uniform(size=3000)
    u[:3] = [0.0, 1.0, 0.5]
    got = st.sample_basis(u, order)
    p = np.abs(v if order is None else v[order]) ** 2
    cdf = np.cumsum(p)
    want = np.array([np.argmax(cdf >= x) for x in u])
    # positions agree except where u falls within rounding of a boundary
    close = np.abs(cdf[want] - u) < 1e-12
    assert np.all((got == want) | close)


@pytest.mark.parametrize('exact', [True, False])
def test_mcclean_dense_seeded_equals_reference_formulation(gpu, exact):
    from qradient_amd.circuit_logic import McClean
    n, L = 5, 2
    axes, angles, obs = _mcclean_case(n, L, 21)
    circ = McClean(n, obs, L, axes=axes, angles=angles)
    shots = 30
    np.random.seed(8)
    E, g = circ.sample_grad_dense(shot_num=shots, exact_expec_val=exact)
    after = np.random.random_sample()
    np.random.seed(8)
    Eo, go = ra.mcclean_sample_grad_dense(n, obs, axes, angles, shots, exact_expec_val=exact)
    assert np.random.random_sample() == after
    assert abs(E - Eo) <= 1e-12
    assert np.max(np.abs(g - go)) <= 1e-12


@pytest.mark.parametrize('exact', [True, False])
def test_mcclean_dense_component_seeded_equals_reference_formulation(gpu, exact):
    from qradient_amd.circuit_logic import McClean
    n, L = 5, 2
    axes, angles, obs = _mcclean_case(n, L, 22, positive=True)
    circ = McClean(n, obs, L, axes=axes, angles=angles, use_observable_components=True)
    shots = 30
    np.random.seed(9)
    E, g = circ.sample_grad_dense_with_component_sampling(shot_num=shots, exact_expec_val=exact)
    after = np.random.random_sample()
    np.random.seed(9)
    Eo, go = ra.mcclean_sample_grad_dense(n, obs, axes, angles, shots, exact_expec_val=exact,
                                          component_sampling=True)
    assert np.random.random_sample() == after
    assert abs(E - Eo) <= 1e-12
    assert np.max(np.abs(g - go)) <= 1e-12


@pytest.mark.parametrize('n,L,ckpt', [(12, 2, 0), (15, 3, 1)])
def test_mcclean_dense_larger_equals_reformulation(gpu, n, L, ckpt):
    from qradient_amd.circuit_logic import McClean
    axes, angles, obs = _mcclean_case(n, L, 23)
    circ = McClean(n, obs, L, axes=axes, angles=angles, ckpt_every=ckpt)
    shots = 20
    np.random.seed(10)
    E, g = circ.sample_grad_dense(shot_num=shots)
    np.random.seed(10)
    Eo, go = _reformulated_mcclean(n, axes, angles, obs, shots, True, False)
    assert abs(E - Eo) <= 1e-12
    assert np.max(np.abs(g - go)) <= 1e-12


def test_mcclean_dense_rejects_xy_observables(gpu):
    from qradient_amd.circuit_logic import McClean
    n, L = 4, 1
    axes, angles, obs = _mcclean_case(n, L, 24, mixed=True)
    with pytest.raises(NotImplementedError):
        McClean(n, obs, L, axes=axes, angles=angles).sample_grad_dense(shot_num=3)


@pytest.mark.parametrize('n,p,exact,ini', [(5, 2, True, False), (5, 2, False, False), (8, 1, True, True)])
def test_qaoa_dense_seeded_equals_reference_formulation(gpu, n, p, exact, ini):
    from qradient_amd.circuit_logic import Qaoa
    rs = np.random.RandomState(40 + n)
    edges = sorted({tuple(sorted(rs.choice(n, 2, replace=False))) for _ in range(2 * n)})
    obs = _maxcut(n, edges)
    betas, gammas = np.pi * rs.rand(p), np.pi * rs.rand(p)
    ini_state = None
    if ini:
        ini_state = rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)
        ini_state /= np.linalg.norm(ini_state)
    circ = Qaoa(n, obs, p)
    shots = 25
    np.random.seed(12)
    E, g = circ.sample_grad_dense(betas, gammas, shot_num=shots, exact_expec_val=exact, ini_state=ini_state)
    after = np.random.random_sample()
    np.random.seed(12)
    Eo, go = ra.qaoa_sample_grad_dense(n, obs, betas, gammas, shots, exact_expec_val=exact, ini_state=ini_state)
    assert np.random.random_sample() == after
    assert abs(E - Eo) <= 1e-12
    assert np.max(np.abs(g - go)) <= 1e-12
