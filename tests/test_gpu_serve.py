"""Serving on the GPU: concurrent requests give bitwise the results they give
alone, and agree with the oracle within the complex64 tolerance."""
import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _reqs():
    from qradient_amd.serve import poisson_trace
    return poisson_trace(12, 50.0, seed=5, n_range=(8, 18), L_range=(3, 12))


@pytest.mark.parametrize('policy', ['fifo', 'shortest'])
def test_results_identical_alone_and_in_mix(gpu, policy):
    from qradient_amd.serve import GradServer, Request
    trace = _reqs()
    alone = []
    srv = GradServer(slots=1, budget_bytes=8 << 30)
    for _, n, L, axes, angles, obs in trace:
        alone.append(srv.submit(Request(n, L, axes, angles, obs)).result(timeout=120))
    srv.close()
    srv = GradServer(slots=6, budget_bytes=8 << 30, policy=policy)
    futs = [srv.submit(Request(n, L, axes, angles, obs)) for _, n, L, axes, angles, obs in trace]
    mixed = [f.result(timeout=120) for f in futs]
    st = srv.stats()
    srv.close()
    assert st['peak_bytes'] <= st['budget_bytes']
    for (Ea, ga), (Em, gm) in zip(alone, mixed):
        assert Ea == Em and np.array_equal(ga, gm)
    for (_, n, L, axes, angles, obs), (E, g) in zip(trace[:6], alone[:6]):
        Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
        assert abs(E - Eo) <= 1e-6 * max(1.0, abs(Eo)) and rel_err(g, go) <= 1e-6


def test_checkpoint_mode_requests(gpu):
    """Shapes above history_max_bytes run in checkpoint mode and stay accurate."""
    from qradient_amd.serve import GradServer, Request, poisson_trace
    trace = poisson_trace(3, 10.0, seed=8, n_range=(14, 16), L_range=(10, 16))
    srv = GradServer(slots=3, budget_bytes=4 << 30, history_max_bytes=1 << 20, ckpt_every=4)
    futs = [srv.submit(Request(n, L, axes, angles, obs)) for _, n, L, axes, angles, obs in trace]
    for (_, n, L, axes, angles, obs), f in zip(trace, futs):
        E, g = f.result(timeout=120)
        Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
        assert rel_err(g, go) <= 1e-6
    srv.close()


def test_slots_overlap_on_the_device(gpu):
    """Requests on different slots run their kernels at the same time: the
    device spans (HIP events on each slot's stream, against the server's
    reference event) of some pair of requests overlap."""
    from qradient_amd.serve import GradServer, Request, poisson_trace
    trace = poisson_trace(8, 1e9, seed=11, n_range=(20, 20), L_range=(20, 20))
    srv = GradServer(slots=4, budget_bytes=16 << 30)
    futs = [srv.submit(Request(n, L, axes, angles, obs)) for _, n, L, axes, angles, obs in trace]
    for f in futs:
        f.result(timeout=120)
    reqs = list(srv.completed)
    srv.close()
    spans = sorted((r.info['device_start_ms'], r.info['device_end_ms'], r.info['slot']) for r in reqs)
    print('device spans (ms, slot):', [(round(a, 2), round(b, 2), s) for a, b, s in spans], flush=True)
    assert all(b > a for a, b, _ in spans)
    overlap = [(x, y) for i, x in enumerate(spans) for y in spans[i + 1:] if y[0] < x[1] and x[2] != y[2]]
    assert overlap, spans


def test_serving_trace_full_range_alone_mixed_and_oracle(gpu):
    """The 14-request trace poisson_trace(14, 50, seed 3) over 13..28 qubits and
    9..64 layers (the 24-qubit / 48-layer, 27 / 56 and 28 / 63 requests exceed
    the 4 GiB history limit and take the checkpoint mode): bitwise the same
    alone and in a mix; the requests of at
    most 20 qubits within 1e-6 of the oracle fixture; the largest checked by
    the exact parameter-shift rule with complex128 forward runs."""
    from qradient_amd.circuit_logic import McClean
    from qradient_amd.serve import GradServer, Request, poisson_trace
    from tests.golden import fixtures as fx
    trace = poisson_trace(14, 50.0, seed=3, n_range=(12, 28), L_range=(4, 64))
    ref = fx.load('serve_trace_seed3.npz')
    budget = 48 << 30
    srv = GradServer(slots=1, budget_bytes=budget)
    modes = [srv.mode_of(n, L) for _, n, L, _, _, _ in trace]
    assert any(m > 0 for m in modes) and any(m == 0 for m in modes)
    alone = []
    for k, (_, n, L, axes, angles, obs) in enumerate(trace):
        print('serve alone request', k, 'n', n, 'L', L, 'mode', modes[k], flush=True)
        alone.append(srv.submit(Request(n, L, axes, angles, obs)).result(timeout=120))
    srv.close()
    srv = GradServer(slots=6, budget_bytes=budget)
    futs = [srv.submit(Request(n, L, axes, angles, obs)) for _, n, L, axes, angles, obs in trace]
    mixed = [f.result(timeout=120) for f in futs]
    st = srv.stats()
    srv.close()
    print('mixed run', st, flush=True)
    assert st['peak_bytes'] <= budget
    for (Ea, ga), (Em, gm) in zip(alone, mixed):
        assert Ea == Em and np.array_equal(ga, gm)
    for k in ref['index']:
        k = int(k)
        _, n, L, axes, angles, obs = trace[k]
        assert np.array_equal(ref['axes_%d' % k], axes) and np.array_equal(ref['angles_%d' % k], angles)
        E, g = alone[k]
        Eo, go = float(ref['E_%d' % k]), ref['grad_%d' % k]
        assert abs(E - Eo) <= 1e-6 * max(1.0, abs(Eo)), (k, E, Eo)
        assert rel_err(g, go) <= 1e-6, (k, rel_err(g, go))
    # the largest request: exact +-pi/2 shift rule on two parameters
    big = max(range(len(trace)), key=lambda k: (trace[k][1], trace[k][2]))
    _, n, L, axes, angles, obs = trace[big]
    E, g = alone[big]
    gnorm = float(np.linalg.norm(g))
    c = McClean(n, obs, L, axes=axes, angles=np.array(angles, copy=True), dtype='complex128')
    last = [(L - 1, q) for q in range(n) if axes[L - 1, q] != 2]
    for (i, q) in [last[0], (L // 2, 5)]:
        base = c.angles[i, q]
        vals = []
        for s in (np.pi / 2, -np.pi / 2):
            c.angles[i, q] = base + s
            vals.append(c.run_expec_val())
        c.angles[i, q] = base
        shift = 0.5 * (vals[0] - vals[1])
        print('  shift check n={} L={} ({}, {}): served {:.6e} shift {:.6e}'.format(n, L, i, q, g[i, q], shift),
              flush=True)
        assert abs(shift - g[i, q]) <= 1e-6 * gnorm
    c._destroy_plan()
