"""Pass launches replayed as a HIP graph: every replay, with new angles and
with other device work between the replays, matches direct launches bitwise
and the oracle within the dtype's bar."""
import ctypes

import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _plan_run(c, graph):
    from qradient_amd import _native
    _native.call('qg_plan_set_graph', c._get_plan(), 1 if graph else 0)
    return c.grad_run()


@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-12), ('complex64', 1e-6)])
def test_mcclean_graph_replays(gpu, dtype, tol):
    from qradient_amd.circuit_logic import McClean
    from qradient_amd.physical_components import State
    n, L = 17, 4
    rs = np.random.RandomState(7)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    g_on = McClean(n, obs, L, dtype=dtype)
    g_off = McClean(n, obs, L, dtype=dtype)
    scratch = State(12, dtype=dtype)
    for rep in range(4):
        axes, angles = sv.mcclean_random_params(L, n, rs)
        for c in (g_on, g_off):
            c.axes, c.angles = axes, angles
        E1, gr1 = _plan_run(g_on, True)
        scratch.xrot(0.3 * rep, rep % 12)          # other device work between replays
        E0, gr0 = _plan_run(g_off, False)
        assert E1 == E0 and np.array_equal(gr1, gr0), rep
        Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
        assert abs(E1 - Eo) <= tol * max(1.0, abs(Eo))
        assert rel_err(gr1, go) <= tol


def test_qaoa_batch_graph_replays(gpu):
    from qradient_amd import _native
    from qradient_amd.batch import QaoaBatch
    from qradient_amd.optimization_problems import MaxCut
    n, p, B = 16, 3, 6
    rs = np.random.RandomState(11)
    obs = [MaxCut.regular3(n, rs).to_observable() for _ in range(B)]
    qb = QaoaBatch(n, p, B, dtype='complex128')
    for rep in range(3):
        betas = rs.uniform(0, np.pi, (B, p))
        gammas = rs.uniform(0, np.pi, (B, p))
        _native.call('qg_plan_set_graph', qb._plan, 1)
        E1, g1 = qb.grad_run(obs, betas, gammas)
        _native.call('qg_plan_set_graph', qb._plan, 0)
        E0, g0 = qb.grad_run(obs, betas, gammas)
        assert np.array_equal(E1, E0) and np.array_equal(g1, g0), rep
        for b in range(B):
            Eo, go = sv.qaoa_grad_run_adjoint(n, obs[b], betas[b], gammas[b])
            assert abs(E1[b] - Eo) <= 1e-12 * max(1.0, abs(Eo))
            assert rel_err(g1[b], go) <= 1e-12
