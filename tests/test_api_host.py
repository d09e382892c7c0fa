"""Host-side behaviour of the drop-in API that needs no GPU: argument checks,
error types and RNG conventions of the reference."""
import warnings

import numpy as np
import pytest

from oracle import statevector as sv
from qradient_amd.circuit_logic import McClean, Qaoa
from qradient_amd.optimization_problems import MaxCut
from qradient_amd.physical_components import Gates, Observable


def _chain(n):
    return sv.zz_chain_observable(n, np.ones(n - 1))


def test_mcclean_draws_axes_then_angles_even_when_passed():
    """mc_clean.py:13-14 evaluates both defaults eagerly from the global RNG."""
    np.random.seed(3)
    c = McClean(4, _chain(4), 5)
    np.random.seed(3)
    u1, u2 = np.random.rand(5, 4), np.random.rand(5, 4)
    assert np.array_equal(c.axes, (3 * u1).astype(int))
    assert np.array_equal(c.angles, 2 * np.pi * u2)
    np.random.seed(3)
    given = np.zeros((5, 4), dtype=int)
    c2 = McClean(4, _chain(4), 5, axes=given)
    assert c2.axes is given
    assert np.array_equal(c2.angles, 2 * np.pi * u2)   # the axes draw still happened first
    assert np.array_equal(np.random.rand(1), np.random.RandomState(3).rand(41)[-1:])


def test_backend_keywords_are_checked():
    with pytest.raises(ValueError):
        McClean(3, _chain(3), 2, dtype='float16')
    with pytest.raises(ValueError):
        McClean(3, _chain(3), 2, entangler='iswap')


def test_observable_validation():
    zz = np.full((3, 3), None)
    zz[1, 0] = 1.0
    with pytest.raises(ValueError):
        Observable({'zz': zz})
    with pytest.raises(ValueError):
        Observable({'z': np.array([1.0, 2.0]), 'x': np.array([1.0, 2.0, 3.0])})
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter('always')
        Observable({'z': np.array([1.0, 0.0, None], dtype=object), 'foo': np.array([1, 2, 3])})
        msgs = [str(x.message) for x in w]
    assert any('Unknown element' in m for m in msgs)
    assert any('zero or almost zero' in m for m in msgs)
    o = Observable(3, {'zz': sv.zz_chain_observable(3, [0.5, -1.0])['zz'], 'x': np.array([None, 2.0, None])})
    assert o.terms == [(0, 1, -1, 2.0), (3, 0, 1, 0.5), (3, 1, 2, -1.0)]
    assert not o.is_diagonal
    with pytest.raises(ValueError):
        Observable(4, _chain(3))


def test_observable_matrix_matches_oracle():
    obs = sv.zz_chain_observable(4, [0.5, -1.0, 2.0])
    obs['y'] = np.array([None, 0.3, None, None], dtype=object)
    M = Observable(obs).matrix.toarray()
    v = np.random.RandomState(0).standard_normal(16) + 0j
    assert np.allclose(M @ v, sv.apply_observable(v.copy(), obs))


def test_qaoa_parameter_count_raises_before_gpu():
    q = Qaoa(4, MaxCut(4, edge_set=np.array([[0, 1], [1, 2]])).to_observable(), 3)
    with pytest.raises(ValueError):
        q.grad_run(np.zeros(2), np.zeros(3))


def test_gates_builder():
    g = Gates(4).add_xrots().add_yrots().add_zrots().add_cnot_ladder()
    assert {'xrots', 'yrots', 'zrots', 'cnot_ladder'} <= g.loaded
    with pytest.raises(ValueError):
        Gates(3).add_cnot_ladder(periodic=True)
    zp = g.zrot_pos[1]
    assert np.array_equal(zp.real, (((np.arange(16) >> 2) & 1) == 0).astype(float))


def test_maxcut_and_regular3():
    m = MaxCut(5, edge_set=np.array([[0, 1], [2, 4]]))
    zz = m.to_observable()['zz']
    assert zz[0, 1] == 1.0 and zz[2, 4] == 1.0 and zz[1, 0] is None
    g = MaxCut.regular3(20, np.random.RandomState(1))
    deg = np.bincount(g.edge_set.ravel(), minlength=20)
    assert (deg == 3).all() and len(g.edge_set) == 30
    with pytest.raises(ValueError):
        MaxCut(4)


def test_maxcut_random_edges_follow_the_reference_draw_sequence():
    """optimization_problems.py:15-31: the candidate pairs come from the nested
    loops i < j; edge_num picks indices with one np.random.choice call,
    edge_probability draws one uniform per pair in that order.  Replayed here
    by hand from the same seed."""
    n = 7
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]
    np.random.seed(5)
    m = MaxCut(n, edge_num=6)
    np.random.seed(5)
    idx = np.random.choice(range(len(pairs)), size=6, replace=False)
    assert [tuple(e) for e in m.edge_set] == [pairs[k] for k in idx]
    np.random.seed(6)
    m = MaxCut(n, edge_probability=0.4)
    np.random.seed(6)
    want = [p for p in pairs if np.random.rand() < 0.4]
    assert [tuple(e) for e in m.edge_set] == want
    # edge_set wins over the random sources, edge_num over edge_probability
    assert MaxCut(3, edge_set=[[0, 2]], edge_num=1).edge_set == [[0, 2]]
    np.random.seed(7)
    a = MaxCut(n, edge_num=2, edge_probability=0.9).edge_set
    np.random.seed(7)
    assert np.array_equal(a, MaxCut(n, edge_num=2).edge_set)
    zz = MaxCut(3, edge_set=np.zeros((0, 2), dtype=int)).to_observable()['zz']
    assert all(v is None for v in zz.ravel())
