"""Batched McClean grad runs (qg_mcclean_batch_plan / _grad_run): every
circuit of a batch gets bitwise the E and gradient a single McClean plan of
the shape gives it (same tiles, grids and reduction order), on the
whole-circuit kernel (n <= 13) and on the pass kernels, both entanglers and
both dtypes; and the complex128 results are within 1e-12 of the oracle."""
import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('n,L,B', [(10, 6, 5), (16, 5, 7), (20, 3, 3)])
@pytest.mark.parametrize('dtype', ['complex128', 'complex64'])
@pytest.mark.parametrize('entangler', ['cnot', 'cz'])
def test_batch_equals_single_runs(gpu, n, L, B, dtype, entangler):
    from qradient_amd.batch import McCleanBatch
    from qradient_amd.circuit_logic import McClean
    rs = np.random.RandomState(1000 + n + L + B)
    axes = (3 * rs.rand(B, L, n)).astype(np.int32)
    angles = 2 * np.pi * rs.rand(B, L, n)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    print('batch n={} L={} B={} {} {}'.format(n, L, B, dtype, entangler), flush=True)
    mb = McCleanBatch(n, L, B, dtype=dtype, entangler=entangler)
    E, g = mb.grad_run(axes, angles, obs)
    del mb
    for b in range(B):
        c = McClean(n, obs, L, axes=axes[b], angles=angles[b], dtype=dtype, entangler=entangler)
        Eb, gb = c.grad_run()
        c._destroy_plan()
        assert E[b] == Eb and np.array_equal(g[b], gb), b
        if dtype == 'complex128' and b < 2:
            Eo, go = sv.mcclean_grad_run_adjoint(n, axes[b], angles[b], obs, entangler=entangler)
            assert abs(E[b] - Eo) <= 1e-12 * max(1.0, abs(Eo)) and rel_err(g[b], go) <= 1e-12


def test_batch_smaller_than_plan_and_rejects_xy(gpu):
    from qradient_amd.batch import McCleanBatch
    rs = np.random.RandomState(5)
    n, L = 15, 4
    axes = (3 * rs.rand(6, L, n)).astype(np.int32)
    angles = 2 * np.pi * rs.rand(6, L, n)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    mb = McCleanBatch(n, L, 6, dtype='complex128')
    E6, g6 = mb.grad_run(axes, angles, obs)
    E2, g2 = mb.grad_run(axes[:2], angles[:2], obs)
    assert np.array_equal(E6[:2], E2) and np.array_equal(g6[:2], g2)
    obs_x = dict(obs)
    obs_x['x'] = np.array([0.5] + [None] * (n - 1), dtype=object)
    with pytest.raises(NotImplementedError):
        mb.grad_run(axes, angles, obs_x)
