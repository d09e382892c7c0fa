"""Optimisers over the GPU grad_run (reference qradient/optimization.py:3-194).

Interface kept from the reference: the class names, constructor and ``step``
signatures, the ``cost_history`` / ``param_history`` arrays, the
``optimizer_info`` dict, the rule chosen by ``optimizer['name']`` with its
hyper-parameter keys and defaults, the update equations, the gradient method
``step`` selects from ``shot_num`` / ``dense_mode`` / ``component_sampling``
(optimization.py:69-91, 113-129), the aliasing of a McClean circuit's
``angles`` with the rule's parameter array (optimization.py:59-61, 91), and
``ValueError`` for an unknown rule or QAOA's sparse sampling mode.

Structure here: one history recorder (``_History``) shared by every driver;
the drivers only say where the gradient comes from.  The rules are a
registry (``RULES``) of small classes over parameter arrays of any shape, so a
[B, p, 2] block of a batched QAOA run is B independent rule states.  Each
step costs one grad_run on the GPU; the rule's update is a few hundred host
FLOPs next to milliseconds of device work, so the parameters stay on the host.

Deliberate difference: ``RateDecayOnPlateau`` starts from an infinite best
cost (the reference starts from 1e10); the two agree for every cost below
1e10.
"""
import numpy as np

# hyper-parameter defaults of the rules (optimization.py:147, 154-156, 172, 180-181)
DEFAULTS = {'step_size': 1e-3, 'beta1': 0.9, 'beta2': 0.999, 'eps': 1e-8,
            'plateau_length': 10, 'decay_rate': 0.5}


def _hyper(hyper_parameters, key):
    return hyper_parameters.get(key, DEFAULTS[key])


# ---------------------------------------------------------------------------
# update rules (optimization.py:145-194)
# ---------------------------------------------------------------------------
class GradientDescentOptimizer:
    """A rule updating ``parameters`` in place from a gradient (and the cost of
    the point it was taken at).  ``iter`` counts the updates."""

    def init(self, parameters, hyper_parameters):
        self.parameters = parameters
        self.step_size = _hyper(hyper_parameters, 'step_size')
        self.iter = 0

    def step(self, gradient, new_cost=None):
        self.iter += 1
        self.parameters -= self._increment(np.asarray(gradient), new_cost)

    def _increment(self, gradient, new_cost):
        raise NotImplementedError


class Adam(GradientDescentOptimizer):
    """Adam with bias-corrected moments (optimization.py:151-167); elementwise."""

    def __init__(self, parameters, hyper_parameters):
        self.init(parameters, hyper_parameters)
        self.beta1 = _hyper(hyper_parameters, 'beta1')
        self.beta2 = _hyper(hyper_parameters, 'beta2')
        self.eps = _hyper(hyper_parameters, 'eps')
        self.m = np.zeros(np.shape(parameters))
        self.v = np.zeros(np.shape(parameters))
        self.m_hat = np.zeros(np.shape(parameters))
        self.v_hat = np.zeros(np.shape(parameters))

    def _increment(self, g, new_cost):
        k = self.iter
        self.m = self.beta1 * self.m + (1 - self.beta1) * g
        self.v = self.beta2 * self.v + (1 - self.beta2) * g ** 2
        self.m_hat = self.m / (1 - self.beta1 ** k)
        self.v_hat = self.v / (1 - self.beta2 ** k)
        return self.step_size * self.m_hat / (np.sqrt(self.v_hat) + self.eps)


class GradientDescent(GradientDescentOptimizer):
    """Plain descent with the rate ``decay_function(step_size, iter)``
    (optimization.py:169-175; default: a constant rate)."""

    def __init__(self, parameters, hyper_parameters):
        self.init(parameters, hyper_parameters)
        self.decay_function = hyper_parameters.get('decay_function', self._constant_rate)

    @staticmethod
    def _constant_rate(step_size, iteration):
        return step_size

    def _increment(self, g, new_cost):
        return self.decay_function(self.step_size, self.iter) * g


class RateDecayOnPlateau(GradientDescentOptimizer):
    """Descent whose rate is multiplied by ``decay_rate`` once the cost has
    failed ``plateau_length`` times in a row to reach a new best
    (optimization.py:177-194).  Counts and the best cost are arrays over the
    leading axis when ``batched`` (one rule state per circuit)."""

    def __init__(self, parameters, hyper_parameters, batched=False):
        self.init(parameters, hyper_parameters)
        self.plateau_length = _hyper(hyper_parameters, 'plateau_length')
        self.decay_rate = _hyper(hyper_parameters, 'decay_rate')
        shape = (np.shape(parameters)[0],) if batched else ()
        self.batched = batched
        self.step_size = np.full(shape, float(self.step_size)) if batched else self.step_size
        self.plateau_counter = np.zeros(shape, dtype=int) if batched else 0
        self.cost = np.full(shape, np.inf) if batched else np.inf

    def _increment(self, g, new_cost):
        cost = np.asarray(new_cost, dtype=float)
        stalled = cost > self.cost
        count = np.where(stalled, self.plateau_counter + 1, 0)
        decay = stalled & (count >= self.plateau_length)
        step = np.where(decay, self.step_size * self.decay_rate, self.step_size)
        count = np.where(decay, 0, count)
        best = np.where(stalled, self.cost, cost)
        if self.batched:
            self.step_size, self.plateau_counter, self.cost = step, count, best
            return step.reshape((-1,) + (1,) * (g.ndim - 1)) * g
        self.step_size, self.plateau_counter, self.cost = float(step), int(count), float(best)
        return self.step_size * g


class BatchRateDecayOnPlateau(RateDecayOnPlateau):
    """RateDecayOnPlateau with a step size, counter and best cost per circuit
    (the leading axis of the parameter block)."""

    def __init__(self, parameters, hyper_parameters):
        RateDecayOnPlateau.__init__(self, parameters, hyper_parameters, batched=True)


RULES = {'Adam': Adam, 'GradientDescent': GradientDescent, 'RateDecayOnPlateau': RateDecayOnPlateau}


def make_rule(optimizer, parameters, batched=False):
    """The rule named by ``optimizer['name']`` over ``parameters`` (updated in place)."""
    name = optimizer.get('name') if isinstance(optimizer, dict) else None
    if name not in RULES:
        raise ValueError('unknown optimizer {!r}: the name must be one of {}'.format(optimizer, sorted(RULES)))
    if batched and name == 'RateDecayOnPlateau':
        return BatchRateDecayOnPlateau(parameters, optimizer)
    return RULES[name](parameters, optimizer)


# ---------------------------------------------------------------------------
# drivers (optimization.py:3-143)
# ---------------------------------------------------------------------------
class _History:
    """cost_history[k] is the cost at param_history[k]; param_history[0] holds
    the initial point and step k writes point k + 1."""

    def _open(self, shape, max_iter, ini):
        self.max_iter = max_iter
        self.iter = 0
        self.param_history = np.zeros((max_iter,) + tuple(shape))
        self.cost_history = np.zeros(max_iter)
        self.param_history[0] = ini

    def _warn_if_exhausted(self):
        if self.iter >= self.max_iter:
            print('{}: all {} iterations of the history are used.'.format(type(self).__name__, self.max_iter))

    def _record(self, cost, gradient):
        self.cost_history[self.iter] = cost
        self.optimizer.step(gradient, cost)
        self.iter += 1
        self.param_history[self.iter] = self.optimizer.parameters


class ParametrizedCircuitOptimizer(_History):
    """Driver base (optimization.py:3-39): a circuit, its rule and histories."""

    def init(self, circuit, max_iter):
        self.circuit = circuit
        self.max_iter = max_iter
        self.iter = 0

    def pick(self, optimizer, ini_parameters):
        self.optimizer = make_rule(optimizer, ini_parameters)

    def step(self):
        pass

    def reset(self):
        pass

    def __str__(self):
        return str(self.optimizer_info)


class McCleanOpt(ParametrizedCircuitOptimizer):
    """Optimiser of a McClean circuit's [L, n] angles (optimization.py:41-96).
    ``ini_parameters`` (default: the circuit's angles) becomes both the
    circuit's ``angles`` and the rule's parameter array."""

    # (shot-based, dense_mode, component_sampling) -> the circuit's gradient method
    _METHODS = {(False, True, False): 'grad_run', (False, False, False): 'grad_run',
                (False, True, True): 'grad_run_with_component_sampling',
                (False, False, True): 'grad_run_with_component_sampling',
                (True, True, False): 'sample_grad_dense',
                (True, True, True): 'sample_grad_dense_with_component_sampling',
                (True, False, False): 'sample_grad',
                (True, False, True): 'sample_grad_with_component_sampling'}

    def __init__(self, circuit, optimizer, max_iter=1000, **kwargs):
        self.init(circuit, max_iter)
        self._setup(kwargs.get('ini_parameters', circuit.angles), optimizer, max_iter)

    def _setup(self, ini, optimizer, max_iter):
        self._open((self.circuit.lnum, self.circuit.qnum), max_iter, ini)
        self.circuit.angles = ini
        self.optimizer_info = optimizer
        self.pick(optimizer, ini)

    def step(self, shot_num=0, dense_mode=True, component_sampling=False):
        self._warn_if_exhausted()
        method = getattr(self.circuit, self._METHODS[(shot_num != 0, bool(dense_mode), bool(component_sampling))])
        e, g = method(shot_num=shot_num) if shot_num != 0 else method()
        self._record(e, g)
        self.circuit.angles = self.optimizer.parameters

    def reset(self, **kwargs):
        ini = kwargs.get('ini_parameters', self.param_history[0])
        self._setup(ini, kwargs.get('optimizer', self.optimizer_info), kwargs.get('max_iter', self.max_iter))


class QaoaOpt(ParametrizedCircuitOptimizer):
    """Optimiser of a QAOA circuit's [p, 2] parameters, columns (beta, gamma)
    (optimization.py:98-143)."""

    def __init__(self, circuit, optimizer, betas, gammas, max_iter=1000):
        self.init(circuit, max_iter)
        self._setup(betas, gammas, optimizer, max_iter)

    def _setup(self, betas, gammas, optimizer, max_iter):
        ini = np.stack([np.asarray(betas, dtype=float), np.asarray(gammas, dtype=float)], axis=1)
        self._open((self.circuit.lnum, 2), max_iter, ini)
        self.optimizer_info = optimizer
        self.pick(optimizer, ini)

    def step(self, shot_num=0, dense_mode=True):
        self._warn_if_exhausted()
        point = self.param_history[self.iter]
        if shot_num == 0:
            e, g = self.circuit.grad_run(point[:, 0], point[:, 1])
        elif not dense_mode:
            raise ValueError('QaoaOpt.step: only dense_mode=True is implemented for shot-based gradients')
        else:
            e, g = self.circuit.sample_grad_dense(point[:, 0], point[:, 1], shot_num=shot_num)
        self._record(e, g)

    def reset(self, **kwargs):
        first = self.param_history[0]
        betas, gammas = kwargs.get('betas'), kwargs.get('gammas')
        if betas is None or gammas is None:
            betas, gammas = first[:, 0].copy(), first[:, 1].copy()
        self._setup(betas, gammas, kwargs.get('optimizer', self.optimizer_info), kwargs.get('max_iter', self.max_iter))


class QaoaBatchOpt(_History):
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
        p = batch.p
        ini = np.stack([np.asarray(betas, dtype=float).reshape(self.B, p),
                        np.asarray(gammas, dtype=float).reshape(self.B, p)], axis=-1)
        self._open((self.B, p, 2), max_iter, ini)
        self.cost_history = np.zeros((max_iter, self.B))
        self.optimizer_info = optimizer
        self.optimizer = make_rule(optimizer, ini, batched=True)

    def step(self):
        self._warn_if_exhausted()
        point = self.param_history[self.iter]
        e, g = self.batch.grad_run(self.observables, point[..., 0], point[..., 1])
        self._record(e, g)
