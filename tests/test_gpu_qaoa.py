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


def _random_state(n, rs):
    v = rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize('n,p', [(5, 2), (11, 3), (16, 2)])
@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-12), ('complex64', 1e-6)])
def test_qaoa_ini_state(gpu, n, p, dtype, tol):
    """grad_run / run_expec_val from a given initial state (qaoa.py:23-38,
    40-70); the caller's array is left as it was."""
    from qradient_amd.circuit_logic import Qaoa
    rs = np.random.RandomState(60 + n)
    obs = sv.maxcut_observable(n, _graph(n, n))
    b, g = np.pi * rs.rand(p), np.pi * rs.rand(p)
    ini = _random_state(n, rs)
    keep = ini.copy()
    circ = Qaoa(n, obs, p, dtype=dtype)
    E, grad = circ.grad_run(b, g, ini_state=ini)
    Eo, go = sv.qaoa_grad_run(n, obs, b, g, ini_state=keep)
    assert abs(E - Eo) <= tol * max(1.0, abs(Eo))
    assert rel_err(grad, go) <= tol, rel_err(grad, go)
    assert abs(circ.run_expec_val(b, g, ini_state=ini) - Eo) <= tol * max(1.0, abs(Eo))
    assert np.array_equal(ini, keep)
    # the same plan still runs from |+>^n afterwards
    E2, g2 = circ.grad_run(b, g)
    Eo2, go2 = sv.qaoa_grad_run(n, obs, b, g)
    assert abs(E2 - Eo2) <= tol * max(1.0, abs(Eo2)) and rel_err(g2, go2) <= tol


def test_qaoa_ini_state_composes_layers(gpu):
    """Two layers from |+>^n equal one layer from the state after the first:
    the energy and the second layer's gradient agree."""
    from qradient_amd.circuit_logic import Qaoa
    n = 14
    rs = np.random.RandomState(8)
    obs = sv.maxcut_observable(n, _graph(n, 8))
    b, g = np.pi * rs.rand(2), np.pi * rs.rand(2)
    E2, grad2 = Qaoa(n, obs, 2).grad_run(b, g)
    C = sv.classical_ham(n, obs)
    mid = sv.reset_state(n, '+') * np.exp(-1j * g[0] * C)
    sv.xrot_all(mid, b[0])
    E1, grad1 = Qaoa(n, obs, 1).grad_run(b[1:], g[1:], ini_state=mid)
    assert abs(E1 - E2) < 1e-11
    assert rel_err(grad1[0], grad2[1]) < 1e-11


@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-12), ('complex64', 1e-6)])
def test_batched_ini_states(gpu, dtype, tol):
    from qradient_amd.batch import QaoaBatch
    n, p, B = 15, 2, 5
    rs = np.random.RandomState(12)
    obs = [sv.maxcut_observable(n, _graph(n, 20 + b, m=2 * n)) for b in range(B)]
    betas, gammas = np.pi * rs.rand(B, p), np.pi * rs.rand(B, p)
    ini = np.stack([_random_state(n, rs) for _ in range(B)])
    keep = ini.copy()
    batch = QaoaBatch(n, p, B, dtype=dtype)
    E, grad = batch.grad_run(obs, betas, gammas, ini_states=ini)
    Ev = batch.expec_val(obs, betas, gammas, ini_states=ini)
    for k in range(B):
        Eo, go = sv.qaoa_grad_run(n, obs[k], betas[k], gammas[k], ini_state=keep[k])
        assert abs(E[k] - Eo) <= tol * max(1.0, abs(Eo))
        assert abs(Ev[k] - Eo) <= tol * max(1.0, abs(Eo))
        assert rel_err(grad[k], go) <= tol
    assert np.array_equal(ini, keep)
