"""GPU parity of Qaoa.grad_run (single and batched) against the CPU oracle."""
import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

K33 = [(a, b) for a in range(3) for b in range(3, 6)]


def _graph(n, seed, m=None):
    rs = np.random.RandomState(seed)
    if n % 2 == 0 and n >= 4 and m is None:
        return sv.random_regular3_edges(n, rs)
    allp = [(i, j) for i in range(n) for j in range(i + 1, n)]
    idx = rs.choice(len(allp), size=min(len(allp), m or n), replace=False)
    return sorted(allp[k] for k in idx)


@pytest.mark.parametrize('n,p,seed', [(2, 1, 0), (3, 2, 1), (6, 3, 2), (10, 4, 3), (13, 2, 4), (14, 3, 5),
                                      (16, 2, 6), (20, 2, 7)])
@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-12), ('complex64', 1e-6)])
def test_qaoa_grad_run_matches_oracle(gpu, n, p, seed, dtype, tol):
    from qradient_amd.circuit_logic import Qaoa
    obs = sv.maxcut_observable(n, _graph(n, seed))
    rs = np.random.RandomState(100 + seed)
    b, g = np.pi * rs.rand(p), np.pi * rs.rand(p)
    E, grad = Qaoa(n, obs, p, dtype=dtype).grad_run(b, g)
    Eo, go = sv.qaoa_grad_run(n, obs, b, g)
    assert abs(E - Eo) <= tol * max(1.0, abs(Eo))
    assert grad.shape == (p, 2)
    assert rel_err(grad, go) <= tol, rel_err(grad, go)


def test_qaoa_closed_form_p1(gpu):
    from qradient_amd.circuit_logic import Qaoa
    obs = sv.maxcut_observable(6, K33)
    b, g = 0.45, 0.8
    E, grad = Qaoa(6, obs, 1).grad_run([b], [g])
    m = len(K33)
    assert abs(E - m * np.sin(2 * b) * np.sin(2 * g) * np.cos(2 * g) ** 2) < 1e-12
    assert abs(grad[0, 0] - m * 2 * np.cos(2 * b) * np.sin(2 * g) * np.cos(2 * g) ** 2) < 1e-12


def test_qaoa_non_integer_weights_use_sincos_phase(gpu):
    from qradient_amd.circuit_logic import Qaoa
    n, p = 9, 3
    rs = np.random.RandomState(3)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    obs['z'] = np.array([0.3 if q % 3 == 0 else None for q in range(n)], dtype=object)
    b, g = rs.rand(p), rs.rand(p)
    E, grad = Qaoa(n, obs, p).grad_run(b, g)
    Eo, go = sv.qaoa_grad_run(n, obs, b, g)
    assert abs(E - Eo) < 1e-12 and rel_err(grad, go) < 1e-12


def test_qaoa_run_expec_val(gpu):
    from qradient_amd.circuit_logic import Qaoa
    obs = sv.maxcut_observable(8, _graph(8, 9))
    b, g = [0.3, 1.2], [0.8, -0.4]
    assert abs(Qaoa(8, obs, 2).run_expec_val(b, g) - sv.qaoa_run_expec_val(8, obs, b, g)) < 1e-12


@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-12), ('complex64', 1e-6)])
def test_batched_equals_single_runs(gpu, dtype, tol):
    from qradient_amd.batch import QaoaBatch
    n, p, B = 12, 3, 7
    rs = np.random.RandomState(11)
    graphs = [sv.random_regular3_edges(n, rs) for _ in range(B)]
    obs = [sv.maxcut_observable(n, e) for e in graphs]
    betas, gammas = np.pi * rs.rand(B, p), np.pi * rs.rand(B, p)
    E, grad = QaoaBatch(n, p, B, dtype=dtype).grad_run(obs, betas, gammas)
    for b in range(B):
        Eo, go = sv.qaoa_grad_run(n, obs[b], betas[b], gammas[b])
        assert abs(E[b] - Eo) <= tol * max(1.0, abs(Eo))
        assert rel_err(grad[b], go) <= tol
    # each circuit alone gives bitwise the same result as inside the batch
    E1, g1 = QaoaBatch(n, p, 1, dtype=dtype).grad_run(obs[3:4], betas[3:4], gammas[3:4])
    assert E1[0] == E[3] and np.array_equal(g1[0], grad[3])


@pytest.mark.parametrize('n,p,integral', [(15, 2, False), (18, 3, False), (16, 2, True)])
@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-12), ('complex64', 1e-6)])
def test_qaoa_pass_path_diag_terms(gpu, n, p, integral, dtype, tol):
    """Tile-pass QAOA with one- and two-bit terms (sincos phase for real
    weights, the phase table for integer weights); the terms are split per
    launch by its register bits."""
    from qradient_amd.circuit_logic import Qaoa
    rs = np.random.RandomState(40 + n)
    zz = np.full((n, n), None)
    for i in range(n):
        j = (i + 1 + i % 4) % n
        a, c = min(i, j), max(i, j)
        if a != c:
            zz[a, c] = float(rs.randint(-2, 3) or 1) if integral else float(rs.standard_normal())
    z = np.array([(float(rs.randint(-1, 2) or 2) if integral else 0.5 * rs.standard_normal()) if q % 3 == 1 else None
                  for q in range(n)], dtype=object)
    obs = {'zz': zz, 'z': z}
    b, g = np.pi * rs.rand(p), np.pi * rs.rand(p)
    E, grad = Qaoa(n, obs, p, dtype=dtype).grad_run(b, g)
    Eo, go = sv.qaoa_grad_run(n, obs, b, g)
    assert abs(E - Eo) <= tol * max(1.0, abs(Eo))
    assert rel_err(grad, go) <= tol, rel_err(grad, go)
