"""Optimisers (reference qradient/optimization.py:3-194) on the CPU, against
closed forms on quadratic costs and against the reference's bookkeeping
(histories, aliasing of the circuit's angles).  No GPU: the circuits here are
quadratic stand-ins with the grad_run signature."""
import numpy as np
import pytest

from qradient_amd.optimization import (Adam, BatchRateDecayOnPlateau, GradientDescent, McCleanOpt, QaoaBatchOpt,
                                       QaoaOpt, RateDecayOnPlateau)


class QuadMcClean:
    """E = 1/2 |angles - c|^2, grad = angles - c."""

    def __init__(self, L, n, seed=0):
        rs = np.random.RandomState(seed)
        self.lnum, self.qnum = L, n
        self.angles = rs.standard_normal((L, n))
        self.c = rs.standard_normal((L, n))

    def grad_run(self):
        d = self.angles - self.c
        return 0.5 * float(np.sum(d * d)), d.copy()


class QuadQaoa:
    def __init__(self, p, seed=0):
        self.lnum = p
        self.c = np.random.RandomState(seed).standard_normal((p, 2))

    def grad_run(self, betas, gammas):
        d = np.stack([betas, gammas], axis=1) - self.c
        return 0.5 * float(np.sum(d * d)), d


class QuadBatch:
    def __init__(self, B, p, seed=0):
        self.p = p
        self.c = np.random.RandomState(seed).standard_normal((B, p, 2))

    def pack(self, obs):
        return list(obs)

    def grad_run(self, obs, betas, gammas):
        d = np.stack([betas, gammas], axis=-1) - self.c
        return 0.5 * np.sum(d * d, axis=(1, 2)), d


def test_gradient_descent_closed_form():
    circ = QuadMcClean(3, 4)
    x0 = circ.angles.copy()
    opt = McCleanOpt(circ, {'name': 'GradientDescent', 'step_size': 0.1}, max_iter=20)
    for _ in range(10):
        opt.step()
    # x_k - c = (1 - lr)^k (x_0 - c)
    assert np.allclose(circ.angles - circ.c, 0.9 ** 10 * (x0 - circ.c), rtol=0, atol=1e-14)
    assert opt.iter == 10 and np.array_equal(opt.param_history[10], circ.angles)
    assert np.array_equal(opt.param_history[0], x0)
    # cost_history[k] is the cost at param_history[k]
    for k in range(10):
        d = opt.param_history[k] - circ.c
        assert abs(opt.cost_history[k] - 0.5 * np.sum(d * d)) < 1e-14


def test_gradient_descent_decay_function():
    circ = QuadMcClean(2, 2, seed=1)
    x0 = circ.angles.copy()
    opt = McCleanOpt(circ, {'name': 'GradientDescent', 'step_size': 0.2,
                            'decay_function': lambda s, k: s / k}, max_iter=5)
    for _ in range(3):
        opt.step()
    f = (1 - 0.2) * (1 - 0.1) * (1 - 0.2 / 3)
    assert np.allclose(circ.angles - circ.c, f * (x0 - circ.c), rtol=0, atol=1e-14)


def test_adam_first_step_is_signed_step_size():
    """m_hat = g and v_hat = g^2 after one step: each parameter moves by
    step_size * g / (|g| + eps)."""
    g = np.array([[0.5, -2.0], [1e-3, 3.0]])
    p = np.zeros_like(g)
    a = Adam(p, {'step_size': 0.01})
    a.step(g)
    assert np.allclose(p, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)
    assert a.parameters is p


def test_adam_converges_and_aliases_angles():
    circ = QuadMcClean(4, 5, seed=2)
    angles = circ.angles
    opt = McCleanOpt(circ, {'name': 'Adam', 'step_size': 0.05}, max_iter=2000)
    for _ in range(1500):
        opt.step()
    # the circuit's angles array is the optimiser's parameter array (optimization.py:59-61)
    assert circ.angles is angles and opt.optimizer.parameters is angles
    assert np.max(np.abs(circ.angles - circ.c)) < 1e-3
    assert opt.cost_history[1499] < 1e-3 * opt.cost_history[0]


def test_rate_decay_on_plateau():
    p = np.zeros(3)
    r = RateDecayOnPlateau(p, {'step_size': 1.0, 'plateau_length': 3, 'decay_rate': 0.5})
    g = np.ones(3)
    costs = [5.0, 4.0, 4.5, 4.6, 4.7, 3.0, 3.5]
    sizes = []
    for c in costs:
        r.step(g, c)
        sizes.append(r.step_size)
    # two improvements, then three worse steps halve the step (counter reset), then a new best
    assert sizes == [1.0, 1.0, 1.0, 1.0, 0.5, 0.5, 0.5]
    assert r.cost == 3.0 and r.plateau_counter == 1
    assert np.allclose(p, -np.sum([1.0, 1.0, 1.0, 1.0, 0.5, 0.5, 0.5]) * g)


def test_unknown_optimizer_and_history_length():
    circ = QuadMcClean(1, 2)
    with pytest.raises(ValueError):
        McCleanOpt(circ, {'name': 'Newton'})
    opt = McCleanOpt(circ, {'name': 'GradientDescent'}, max_iter=3)
    opt.step()
    opt.step()
    # the histories hold max_iter rows, as in the reference: the next step
    # has no row for its parameters
    with pytest.raises(IndexError):
        opt.step()


class TracingMcClean(QuadMcClean):
    """Records which gradient method a step called (optimization.py:69-84)."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.calls = []

    def _traced(name):
        def method(self, **kw):
            self.calls.append((name, kw))
            return self.grad_run()
        return method

    grad_run_with_component_sampling = _traced('grad_run_with_component_sampling')
    sample_grad = _traced('sample_grad')
    sample_grad_with_component_sampling = _traced('sample_grad_with_component_sampling')
    sample_grad_dense = _traced('sample_grad_dense')
    sample_grad_dense_with_component_sampling = _traced('sample_grad_dense_with_component_sampling')


def test_step_dispatches_like_the_reference():
    circ = TracingMcClean(1, 2)
    opt = McCleanOpt(circ, {'name': 'Adam'}, max_iter=10)
    opt.step(component_sampling=True)
    opt.step(shot_num=10, dense_mode=False)
    opt.step(shot_num=5, dense_mode=False, component_sampling=True)
    opt.step(shot_num=3)
    opt.step(shot_num=4, component_sampling=True)
    assert circ.calls == [('grad_run_with_component_sampling', {}), ('sample_grad', {'shot_num': 10}),
                          ('sample_grad_with_component_sampling', {'shot_num': 5}),
                          ('sample_grad_dense', {'shot_num': 3}),
                          ('sample_grad_dense_with_component_sampling', {'shot_num': 4})]
    assert opt.iter == 5


class TracingQaoa(QuadQaoa):
    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.calls = []

    def sample_grad_dense(self, betas, gammas, shot_num=1):
        self.calls.append(('sample_grad_dense', shot_num))
        return self.grad_run(betas, gammas)


def test_qaoa_step_dispatches_like_the_reference():
    circ = TracingQaoa(2)
    qopt = QaoaOpt(circ, {'name': 'Adam'}, np.zeros(2), np.zeros(2), max_iter=3)
    qopt.step(shot_num=10)
    assert circ.calls == [('sample_grad_dense', 10)]
    with pytest.raises(ValueError):
        qopt.step(shot_num=10, dense_mode=False)


def test_mcclean_reset():
    circ = QuadMcClean(2, 3, seed=4)
    x0 = circ.angles.copy()
    opt = McCleanOpt(circ, {'name': 'Adam', 'step_size': 0.1}, max_iter=10)
    for _ in range(4):
        opt.step()
    opt.reset()
    assert opt.iter == 0 and np.array_equal(circ.angles, x0) and np.array_equal(opt.param_history[0], x0)


def test_qaoa_opt_layout_and_reset():
    circ = QuadQaoa(3, seed=5)
    b0, g0 = np.array([0.1, 0.2, 0.3]), np.array([1.0, 1.1, 1.2])
    opt = QaoaOpt(circ, {'name': 'GradientDescent', 'step_size': 0.5}, b0, g0, max_iter=10)
    assert np.array_equal(opt.param_history[0, :, 0], b0) and np.array_equal(opt.param_history[0, :, 1], g0)
    for _ in range(4):
        opt.step()
    x0 = np.stack([b0, g0], axis=1)
    assert np.allclose(opt.param_history[4] - circ.c, 0.5 ** 4 * (x0 - circ.c), rtol=0, atol=1e-14)
    opt.reset(betas=b0 + 1, gammas=g0)
    assert opt.iter == 0 and np.array_equal(opt.param_history[0, :, 0], b0 + 1)
    with pytest.raises(ValueError):
        opt.step(shot_num=5, dense_mode=False)


@pytest.mark.parametrize('rule', [{'name': 'Adam', 'step_size': 0.05},
                                  {'name': 'GradientDescent', 'step_size': 0.3},
                                  {'name': 'RateDecayOnPlateau', 'step_size': 2.5, 'plateau_length': 2}])
def test_batch_opt_equals_single_circuit_trajectories(rule):
    """Each circuit of a batch follows the trajectory QaoaOpt gives it alone
    (step size 2.5 overshoots, so the plateau rule decays)."""
    B, p, steps = 4, 3, 12
    rs = np.random.RandomState(6)
    betas, gammas = rs.rand(B, p), rs.rand(B, p)
    qb = QuadBatch(B, p, seed=7)
    bopt = QaoaBatchOpt(qb, [None] * B, rule, betas, gammas, max_iter=steps + 1)
    for _ in range(steps):
        bopt.step()
    for b in range(B):
        circ = QuadQaoa(p)
        circ.c = qb.c[b]
        opt = QaoaOpt(circ, dict(rule), betas[b], gammas[b], max_iter=steps + 1)
        for _ in range(steps):
            opt.step()
        assert np.array_equal(bopt.param_history[:, b], opt.param_history)
        assert np.array_equal(bopt.cost_history[:, b], opt.cost_history)
    if rule['name'] == 'RateDecayOnPlateau':
        assert isinstance(bopt.optimizer, BatchRateDecayOnPlateau) and np.all(bopt.optimizer.step_size < 2.5)
