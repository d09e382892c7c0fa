"""The library against the golden fixtures (tests/golden): McClean 10 qubits
x 10 layers with Z0 Z1 (both ladders), McClean 14 qubits with Z and ZZ terms,
QAOA on a 3-regular graph, and four circuits of the 512-graph batch workload.
Bars: complex128 1e-12, complex64 1e-6 (norm-relative gradients)."""
import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err
from tests.golden import fixtures as fx

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('entangler', ['cnot', 'cz'])
@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-12), ('complex64', 1e-6)])
def test_mcclean_config1(gpu, entangler, dtype, tol):
    from qradient_amd.circuit_logic import McClean
    d = fx.load('mcclean_n10_L10_z0z1.npz')
    c = McClean(10, fx.obs_from(d), 10, axes=d['axes'], angles=d['angles'], dtype=dtype, entangler=entangler)
    E, g = c.grad_run()
    assert abs(E - d['E_' + entangler]) <= tol
    assert rel_err(g, d['grad_' + entangler]) <= tol


@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-12), ('complex64', 1e-6)])
def test_mcclean_n14(gpu, dtype, tol):
    from qradient_amd.circuit_logic import McClean
    d = fx.load('mcclean_n14_L3_zzchain.npz')
    c = McClean(14, fx.obs_from(d), 3, axes=d['axes'], angles=d['angles'], dtype=dtype)
    E, g = c.grad_run()
    assert abs(E - d['E']) <= tol * max(1.0, abs(d['E']))
    assert rel_err(g, d['grad']) <= tol


@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-12), ('complex64', 1e-6)])
def test_qaoa_small(gpu, dtype, tol):
    from qradient_amd.circuit_logic import Qaoa
    d = fx.load('qaoa_n8_p3_reg3.npz')
    E, g = Qaoa(8, fx.maxcut_obs(8, d['edges']), 3, dtype=dtype).grad_run(d['betas'], d['gammas'])
    assert abs(E - d['E']) <= tol * max(1.0, abs(d['E']))
    assert rel_err(g, d['grad']) <= tol


def test_qaoa_batch_workload_head(gpu):
    """The first four circuits of the 512-graph workload, batched, complex128."""
    from qradient_amd.batch import QaoaBatch
    d = fx.load('qaoa_reg3_n20_b512.npz')
    B = 4
    obs = [fx.maxcut_obs(20, d['edges'][b]) for b in range(B)]
    E, g = QaoaBatch(20, 8, B, dtype='complex128').grad_run(obs, d['betas'][:B], d['gammas'][:B])
    for b in range(B):
        Eo, go = sv.qaoa_grad_run_adjoint(20, obs[b], d['betas'][b], d['gammas'][b])
        assert abs(E[b] - Eo) <= 1e-12 * max(1.0, abs(Eo))
        assert rel_err(g[b], go) <= 1e-12


def test_sampled_gradients(gpu):
    """Seeded sampled gradients of the library equal the frozen oracle values
    (same random numbers, complex128 probabilities within 1e-12)."""
    from qradient_amd.circuit_logic import McClean, Qaoa
    d = fx.load('sampled_n5.npz')
    obs = fx.obs_from(d)
    n, L, shots = 5, 2, 40
    circ = McClean(n, obs, L, axes=d['axes'], angles=d['angles'], use_observable_components=True)
    runs = {'component': lambda: circ.grad_run_with_component_sampling(),
            'shift': lambda: circ.sample_grad(shot_num=shots),
            'shift_component': lambda: circ.sample_grad_with_component_sampling(shot_num=shots),
            'dense': lambda: circ.sample_grad_dense(shot_num=shots),
            'dense_component': lambda: circ.sample_grad_dense_with_component_sampling(shot_num=shots),
            'qaoa_dense': lambda: Qaoa(n, fx.maxcut_obs(n, d['edges']), 2).sample_grad_dense(
                d['betas'], d['gammas'], shot_num=shots)}
    for key, fn in runs.items():
        np.random.seed(31)
        E, g = fn()
        assert abs(E - d['E_' + key]) <= 1e-12, key
        assert np.max(np.abs(g - d['grad_' + key])) <= 1e-12, key
