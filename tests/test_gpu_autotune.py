"""The autotuner on the GPU: the QAOA batch shapes it tries (tile size,
spectator bits, chunk, workgroups per CU) all give the oracle's results, the
tuned shape is cached and applied by QaoaBatch(tune=True), and the
checkpoint-period tuner picks one of its candidates, each accurate."""
import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture
def cache(tmp_path, monkeypatch):
    monkeypatch.setenv('QRADIENT_AMD_TUNE_CACHE', str(tmp_path / 'tune.json'))
    return tmp_path / 'tune.json'


def test_qaoa_shapes_agree_and_tuned_shape_is_applied(gpu, cache):
    from qradient_amd import autotune
    from qradient_amd.batch import QaoaBatch
    from qradient_amd.optimization_problems import MaxCut
    n, p, B = 14, 3, 24
    rs = np.random.RandomState(9)
    obs = [MaxCut.regular3(n, rs).to_observable() for _ in range(B)]
    betas, gammas = np.pi * rs.rand(B, p), np.pi * rs.rand(B, p)
    ref = [sv.qaoa_grad_run_adjoint(n, obs[b], betas[b], gammas[b]) for b in range(4)]
    qb = QaoaBatch(n, p, B, dtype='complex128')
    for shape in ({'kmax': 0, 'cmin': 3, 'chunk': 0, 'wg_per_cu': 16}, {'kmax': 12, 'cmin': 4, 'chunk': 5, 'wg_per_cu': 8},
                  {'kmax': 11, 'cmin': 3, 'chunk': 16, 'wg_per_cu': 32}, {'kmax': 13, 'cmin': 3, 'chunk': 1,
                                                                          'wg_per_cu': 4}):
        print('qaoa shape', shape, flush=True)
        autotune.apply_qaoa(qb._plan, shape)
        E, g = qb.grad_run(obs, betas, gammas)
        for b in range(4):
            assert abs(E[b] - ref[b][0]) <= 1e-12 and rel_err(g[b], ref[b][1]) <= 1e-12, (shape, b)
    del qb
    best, results = autotune.tune_qaoa(n, p, B, 'complex64', reps=1)
    assert results and best == autotune.lookup(autotune.qaoa_key(n, p, B, 'complex64'))
    tuned = QaoaBatch(n, p, B, dtype='complex64', tune=True)
    assert tuned.shape == best
    E, g = tuned.grad_run(obs, betas, gammas)
    for b in range(4):
        assert rel_err(g[b], ref[b][1]) <= 1e-6


def test_ckpt_tuner_picks_a_candidate(gpu, cache):
    from qradient_amd import autotune
    from qradient_amd.circuit_logic import McClean
    n, L = 18, 8
    k, results = autotune.tune_ckpt(n, L, 'complex64', candidates=[1, 2, 4])
    assert k in (1, 2, 4) and len(results) == 3
    assert autotune.lookup(autotune.ckpt_key(n, L, 'complex64', 'cnot')) == k
    rs = np.random.RandomState(4)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    for kk in (1, 2, 4):
        c = McClean(n, obs, L, axes=axes, angles=angles, dtype='complex64', ckpt_every=kk)
        E, g = c.grad_run()
        c._destroy_plan()
        assert rel_err(g, go) <= 1e-6
