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


def test_results_identical_alone_and_in_mix(gpu):
    from qradient_amd.serve import GradServer, Request
    trace = _reqs()
    alone = []
    srv = GradServer(slots=1, budget_bytes=8 << 30)
    for _, n, L, axes, angles, obs in trace:
        alone.append(srv.submit(Request(n, L, axes, angles, obs)).result(timeout=120))
    srv.close()
    srv = GradServer(slots=6, budget_bytes=8 << 30)
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
