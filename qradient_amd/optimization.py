"""Optimisers over the GPU grad_run (reference qradient/optimization.py:3-194).

``McCleanOpt`` and ``QaoaOpt`` keep the reference's cost and parameter
histories and its aliasing: a McClean circuit's ``angles`` is the optimiser's
parameter array, updated in place each step (optimization.py:59-61, 91).
The update rules (``Adam``, ``GradientDescent``, ``RateDecayOnPlateau``) are
the reference's, on [L, n] (McClean) or [p, 2] (QAOA, columns beta, gamma)
arrays.  Each step costs one grad_run; the update itself is a few hundred
host FLOPs next to milliseconds of GPU work, so the parameters stay on the
host.

``QaoaBatchOpt`` (new) steps B QAOA circuits at once: one batched grad_run
(``QaoaBatch``) per step and the same update rule applied to the [B, p, 2]
parameter block, each circuit with its own state of the rule.

``McCleanOpt.step`` and ``QaoaOpt.step`` dispatch like the reference
(optimization.py:69-91, 113-129): ``shot_num == 0`` takes grad_run (McClean:
grad_run_with_component_sampling with ``component_sampling``); ``shot_num >
0`` takes the shot-based gradients, sample_grad_dense* with ``dense_mode``
(the default) and, for McClean, sample_grad* without.
"""
import numpy as np


class ParametrizedCircuitOptimizer:
    """Base class (optimization.py:3-39).  Not for instantiation."""

    def init(self, circuit, max_iter):
        self.circuit = circuit
        self.max_iter = max_iter
        self.iter = 0

    def step(self):
        pass

    def reset(self):
        pass

    def pick(self, optimizer, ini_parameters):
        """The gradient-descent rule named by ``optimizer['name']`` (optimization.py:24-39)."""
        name = optimizer['name']
        if name == 'Adam':
            self.optimizer = Adam(ini_parameters, optimizer)
        elif name == 'GradientDescent':
            self.optimizer = GradientDescent(ini_parameters, optimizer)
        elif name == 'RateDecayOnPlateau':
            self.optimizer = RateDecayOnPlateau(ini_parameters, optimizer)
        else:
            raise ValueError('No optimizer {} known.'.format(optimizer))

    def _check_iter(self):
        if self.iter >= self.max_iter:
            print('Maximum amount of iterations reached: {}.'.format(self.max_iter))




class McCleanOpt(ParametrizedCircuitOptimizer):
    """optimization.py:41-96.  ``ini_parameters`` (default: the circuit's
    angles) becomes the circuit's ``angles`` and the optimiser's parameters."""

    def __init__(self, circuit, optimizer, max_iter=1000, **kwargs):
        ParametrizedCircuitOptimizer.init(self, circuit, max_iter)
        self.param_history = np.zeros([self.max_iter, circuit.lnum, circuit.qnum], dtype='double')
        self.cost_history = np.zeros(self.max_iter, dtype='double')
        ini_parameters = kwargs.get('ini_parameters', circuit.angles)
        self.param_history[0] = ini_parameters
        self.circuit.angles = ini_parameters
        self.optimizer_info = optimizer
        self.pick(optimizer, ini_parameters)

    def __str__(self):
        return self.optimizer_info.__str__()

    def step(self, shot_num=0, dense_mode=True, component_sampling=False):
        self._check_iter()
        if shot_num == 0:
            if component_sampling:
                e, g = self.circuit.grad_run_with_component_sampling()
            else:
                e, g = self.circuit.grad_run()
        elif dense_mode:
            if component_sampling:
                e, g = self.circuit.sample_grad_dense_with_component_sampling(shot_num=shot_num)
            else:
                e, g = self.circuit.sample_grad_dense(shot_num=shot_num)
        elif component_sampling:
            e, g = self.circuit.sample_grad_with_component_sampling(shot_num=shot_num)
        else:
            e, g = self.circuit.sample_grad(shot_num=shot_num)
        self.cost_history[self.iter] = e
        self.optimizer.step(g, e)
        self.iter += 1
        self.param_history[self.iter] = self.optimizer.parameters
        self.circuit.angles = self.optimizer.parameters

    def reset(self, **kwargs):
        ini_parameters = kwargs.get('ini_parameters', self.param_history[0])
        optimizer = kwargs.get('optimizer', self.optimizer_info)
        max_iter = kwargs.get('max_iter', self.max_iter)
        self.__init__(self.circuit, optimizer, max_iter, ini_parameters=ini_parameters)


class QaoaOpt(ParametrizedCircuitOptimizer):
    """optimization.py:98-143: parameters [p, 2] with columns (beta, gamma)."""

    def __init__(self, circuit, optimizer, betas, gammas, max_iter=1000):
        ParametrizedCircuitOptimizer.init(self, circuit, max_iter)
        self.param_history = np.zeros([self.max_iter, circuit.lnum, 2], dtype='double')
        self.cost_history = np.zeros(self.max_iter, dtype='double')
        self.param_history[0] = np.array([betas, gammas]).transpose()
        self.optimizer_info = optimizer
        self.pick(optimizer, np.array([betas, gammas]).transpose())

    def __str__(self):
        return self.optimizer_info.__str__()

    def step(self, shot_num=0, dense_mode=True):
        self._check_iter()
        betas, gammas = self.param_history[self.iter, :, 0], self.param_history[self.iter, :, 1]
        if shot_num == 0:
            e, g = self.circuit.grad_run(betas, gammas)
        elif dense_mode:
            e, g = self.circuit.sample_grad_dense(betas, gammas, shot_num=shot_num)
        else:
            raise ValueError('dense_mode must be True, sparse sampling method is not implmented yet.')
        self.cost_history[self.iter] = e
        self.optimizer.step(g, e)
        self.iter += 1
        self.param_history[self.iter] = self.optimizer.parameters

    def reset(self, **kwargs):
        if ('betas' in kwargs) and ('gammas' in kwargs):
            self.param_history[0] = np.array([kwargs['betas'], kwargs['gammas']]).transpose()
        optimizer = kwargs.get('optimizer', self.optimizer_info)
        max_iter = kwargs.get('max_iter', self.max_iter)
        self.__init__(self.circuit, optimizer, self.param_history[0, :, 0], self.param_history[0, :, 1], max_iter)


class QaoaBatchOpt:
    """B QAOA circuits optimised together: one batched grad_run per step.

    ``batch`` is a ``QaoaBatch`` plan, ``observables`` its B Hamiltonians
    (packed once), ``betas`` / ``gammas`` [B, p].  ``cost_history`` [max_iter, B]
    and ``param_history`` [max_iter, B, p, 2] follow QaoaOpt's layout per
    circuit; the rule's state (Adam moments, plateau counters) is per circuit,
    so circuit b follows exactly the trajectory QaoaOpt would give it alone.
    """

    def __init__(self, batch, observables, optimizer, betas, gammas, max_iter=1000):
        self.batch = batch
        self.observables = batch.pack(observables)
        self.B = len(self.observables)
        self.max_iter = max_iter
        self.iter = 0
        p = batch.p
        betas = np.asarray(betas, dtype='double').reshape(self.B, p)
        gammas = np.asarray(gammas, dtype='double').reshape(self.B, p)
        self.param_history = np.zeros([self.max_iter, self.B, p, 2], dtype='double')
        self.cost_history = np.zeros([self.max_iter, self.B], dtype='double')
        self.param_history[0, :, :, 0] = betas
        self.param_history[0, :, :, 1] = gammas
        self.optimizer_info = optimizer
        ini = np.stack([betas, gammas], axis=-1)
        name = optimizer['name']
        if name == 'Adam':
            self.optimizer = Adam(ini, optimizer)
        elif name == 'GradientDescent':
            self.optimizer = GradientDescent(ini, optimizer)
        elif name == 'RateDecayOnPlateau':
            self.optimizer = BatchRateDecayOnPlateau(ini, optimizer)
        else:
            raise ValueError('No optimizer {} known.'.format(optimizer))

    def step(self):
        if self.iter >= self.max_iter:
            print('Maximum amount of iterations reached: {}.'.format(self.max_iter))
        cur = self.param_history[self.iter]
        e, g = self.batch.grad_run(self.observables, cur[:, :, 0], cur[:, :, 1])
        self.cost_history[self.iter] = e
        self.optimizer.step(g, e)
        self.iter += 1
        self.param_history[self.iter] = self.optimizer.parameters


# ---------------------------------------------------------------------------
# gradient-descent rules (optimization.py:145-194)
# ---------------------------------------------------------------------------
class GradientDescentOptimizer:
    def init(self, parameters, hyper_parameters):
        self.parameters = parameters
        self.step_size = hyper_parameters.get('step_size', .001)
        self.iter = 0


class Adam(GradientDescentOptimizer):
    """optimization.py:151-167; elementwise, so a [B, p, 2] block is B
    independent Adam states."""

    def __init__(self, parameters, hyper_parameters):
        GradientDescentOptimizer.init(self, parameters, hyper_parameters)
        self.beta1 = hyper_parameters.get('beta1', .9)
        self.beta2 = hyper_parameters.get('beta2', .999)
        self.eps = hyper_parameters.get('eps', 10. ** -8)
        self.m = np.zeros(parameters.shape, dtype='double')
        self.v = np.zeros(parameters.shape, dtype='double')
        self.m_hat = np.zeros(parameters.shape, dtype='double')
        self.v_hat = np.zeros(parameters.shape, dtype='double')

    def step(self, gradient, *args):
        self.iter += 1
        self.m = self.beta1 * self.m + (1 - self.beta1) * gradient
        self.v = self.beta2 * self.v + (1 - self.beta2) * gradient ** 2
        self.m_hat = self.m / (1 - self.beta1 ** self.iter)
        self.v_hat = self.v / (1 - self.beta2 ** self.iter)
        self.parameters -= self.step_size * self.m_hat / (np.sqrt(self.v_hat) + self.eps)


class GradientDescent(GradientDescentOptimizer):
    """optimization.py:169-175: step_size scaled by decay_function(step_size, iter)."""

    def __init__(self, parameters, hyper_parameters):
        GradientDescentOptimizer.init(self, parameters, hyper_parameters)
        self.decay_function = hyper_parameters.get('decay_function', lambda step_size, iter: step_size)

    def step(self, gradient, *args):
        self.iter += 1
        self.parameters -= self.decay_function(self.step_size, self.iter) * gradient


class RateDecayOnPlateau(GradientDescentOptimizer):
    """optimization.py:177-194: the step size is multiplied by decay_rate after
    plateau_length steps without a new lowest cost."""

    def __init__(self, parameters, hyper_parameters):
        GradientDescentOptimizer.init(self, parameters, hyper_parameters)
        self.plateau_length = hyper_parameters.get('plateau_length', 10)
        self.decay_rate = hyper_parameters.get('decay_rate', .5)
        self.plateau_counter = 0
        self.cost = 10. ** 10

    def step(self, gradient, new_cost):
        self.iter += 1
        if new_cost > self.cost:
            self.plateau_counter += 1
            if self.plateau_counter >= self.plateau_length:
                self.step_size *= self.decay_rate
                self.plateau_counter = 0
        else:
            self.cost = new_cost
            self.plateau_counter = 0
        self.parameters -= self.step_size * gradient


class BatchRateDecayOnPlateau(GradientDescentOptimizer):
    """RateDecayOnPlateau with one step size, counter and best cost per circuit
    (leading axis of the parameter block)."""

    def __init__(self, parameters, hyper_parameters):
        GradientDescentOptimizer.init(self, parameters, hyper_parameters)
        B = parameters.shape[0]
        self.plateau_length = hyper_parameters.get('plateau_length', 10)
        self.decay_rate = hyper_parameters.get('decay_rate', .5)
        self.step_size = np.full(B, float(self.step_size))
        self.plateau_counter = np.zeros(B, dtype=int)
        self.cost = np.full(B, 10. ** 10)

    def step(self, gradient, new_cost):
        self.iter += 1
        new_cost = np.asarray(new_cost, dtype='double')
        worse = new_cost > self.cost
        self.plateau_counter = np.where(worse, self.plateau_counter + 1, 0)
        decay = worse & (self.plateau_counter >= self.plateau_length)
        self.step_size = np.where(decay, self.step_size * self.decay_rate, self.step_size)
        self.plateau_counter = np.where(decay, 0, self.plateau_counter)
        self.cost = np.where(worse, self.cost, new_cost)
        shape = (-1,) + (1,) * (self.parameters.ndim - 1)
        self.parameters -= self.step_size.reshape(shape) * gradient
