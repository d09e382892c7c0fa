"""GPU State methods against the oracle's gate restatements."""
import numpy as np
import pytest

from oracle import statevector as sv

pytestmark = pytest.mark.gpu


def _rand_state(n, seed):
    rs = np.random.RandomState(seed)
    v = rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-14), ('complex64', 2e-6)])
def test_rotations_and_derivatives(gpu, dtype, tol):
    from qradient_amd.physical_components import State
    n = 6
    v0 = _rand_state(n, 0)
    for name in ('xrot', 'yrot', 'zrot', 'dxrot', 'dyrot', 'dzrot'):
        for q in range(n):
            s = State(n, dtype=dtype)
            s.vec = v0
            getattr(s, name)(0.7 + q, q)
            ref = getattr(sv, name)(v0.copy(), 0.7 + q, q)
            assert np.abs(s.vec - ref).max() <= tol, (name, q)


def test_reset_and_norm(gpu):
    from qradient_amd.physical_components import State
    s = State(5)
    assert np.allclose(s.vec, sv.reset_state(5, '0'))
    s.reset('+')
    assert np.allclose(s.vec, sv.reset_state(5, '+'))
    assert abs(s.norm_error()) < 1e-14
    with pytest.raises(ValueError):
        State(3, ini='1')


@pytest.mark.parametrize('n', [2, 3, 6, 7])
def test_ladders_and_cnot(gpu, n):
    from qradient_amd.physical_components import State
    v0 = _rand_state(n, n)
    for stacking in (0, 1):
        s = State(n)
        s.vec = v0
        s.cnot_ladder(stacking)
        assert np.allclose(s.vec, sv.cnot_ladder(v0.copy(), stacking), atol=1e-15)
    s = State(n)
    s.vec = v0
    s.cz_ladder()
    assert np.allclose(s.vec, sv.cz_ladder(v0.copy()), atol=1e-15)
    for (c, t) in [(0, n - 1), (n - 1, 0), (0, 1)]:
        s = State(n)
        s.vec = v0
        s.cnot(c, t)
        assert np.allclose(s.vec, sv.cnot(v0.copy(), c, t), atol=1e-15)


def test_observable_multiply_and_expec(gpu):
    from qradient_amd.physical_components import Observable, State
    n = 6
    rs = np.random.RandomState(3)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    obs['x'] = np.array([0.3, None, -0.2, None, None, 0.5], dtype=object)
    obs['y'] = np.array([None, 0.4, None, None, -0.1, None], dtype=object)
    obs['z'] = np.array([None, None, 0.25, None, None, None], dtype=object)
    v0 = _rand_state(n, 9)
    s = State(n)
    s.vec = v0
    o = Observable(obs)
    assert abs(s.expec_val(o) - sv.expec_val(v0, obs)) <= 1e-13
    s.multiply_matrix(o)
    assert np.abs(s.vec - sv.apply_observable(v0.copy(), obs)).max() <= 1e-13
    s.vec = v0
    s.multiply_matrix(o.matrix)
    assert np.abs(s.vec - o.matrix.dot(v0)).max() <= 1e-13
