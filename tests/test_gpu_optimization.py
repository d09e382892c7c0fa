"""Optimiser steps over the GPU grad_run (optimization.py:41-143) against the
same rule driven by the oracle's grad_run, and the batched QAOA optimiser
against per-circuit QaoaOpt runs."""
import numpy as np
import pytest

from oracle import statevector as sv
from qradient_amd.optimization import Adam, McCleanOpt, QaoaBatchOpt, QaoaOpt
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


# Adam divides each gradient component by its own running magnitude, so a
# small component's absolute gradient error (1e-12 of the gradient's norm)
# becomes step_size * error / |g_i| in the parameter: the bar is looser than
# for plain gradient descent, whose update is linear in the gradient.
@pytest.mark.parametrize('rule,tol', [({'name': 'GradientDescent', 'step_size': 0.1}, 1e-11),
                                      ({'name': 'Adam', 'step_size': 0.05}, 1e-9)])
@pytest.mark.parametrize('n,L', [(8, 3), (15, 2)])
def test_mcclean_steps_match_oracle_loop(gpu, n, L, rule, tol):
    from qradient_amd.circuit_logic import McClean
    from qradient_amd.optimization import GradientDescent
    rs = np.random.RandomState(70 + n)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    steps = 5
    circ = McClean(n, obs, L, axes=axes, angles=angles.copy())
    opt = McCleanOpt(circ, rule, max_iter=steps + 1)
    for _ in range(steps):
        opt.step()
    # the same rule on the oracle's gradients
    p = angles.copy()
    ref = (Adam if rule['name'] == 'Adam' else GradientDescent)(p, rule)
    for k in range(steps):
        E, g = sv.mcclean_grad_run(n, axes, p, obs)
        assert abs(opt.cost_history[k] - E) <= 1e-12 * max(1.0, abs(E))
        ref.step(g, E)
        assert rel_err(opt.param_history[k + 1], p) <= tol
    assert circ.angles is opt.optimizer.parameters


def test_qaoa_batch_opt_equals_qaoa_opt(gpu):
    from qradient_amd.batch import QaoaBatch
    from qradient_amd.circuit_logic import Qaoa
    n, p, B, steps = 10, 2, 4, 3
    rs = np.random.RandomState(80)
    obs = [sv.maxcut_observable(n, sv.random_regular3_edges(n, rs)) for _ in range(B)]
    betas, gammas = np.pi * rs.rand(B, p), np.pi * rs.rand(B, p)
    rule = {'name': 'Adam', 'step_size': 0.02}
    bopt = QaoaBatchOpt(QaoaBatch(n, p, B, dtype='complex128'), obs, rule, betas, gammas, max_iter=steps + 1)
    for _ in range(steps):
        bopt.step()
    for b in range(B):
        opt = QaoaOpt(Qaoa(n, obs[b], p), dict(rule), betas[b], gammas[b], max_iter=steps + 1)
        for _ in range(steps):
            opt.step()
        assert rel_err(bopt.param_history[:, b], opt.param_history) <= 1e-12
        assert rel_err(bopt.cost_history[:, b], opt.cost_history) <= 1e-12
