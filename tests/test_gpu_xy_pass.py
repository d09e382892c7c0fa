"""x / y observable terms (observable.py:34-79) on the pass kernels: McClean
grad_run and run_expec_val above 13 qubits, with ini_state, in checkpoint
mode and under HIP graph replay, against the oracle (complex128 within
1e-12, complex64 within 1e-6)."""
import ctypes

import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = {'complex128': 1e-12, 'complex64': 1e-6}


def _obs(n, rs):
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    obs['x'] = np.array([rs.standard_normal() if q % 2 == 0 else None for q in range(n)], dtype=object)
    obs['y'] = np.array([rs.standard_normal() if q % 3 == 1 else None for q in range(n)], dtype=object)
    obs['z'] = np.array([0.3 if q == n - 1 else None for q in range(n)], dtype=object)
    return obs


@pytest.mark.parametrize('n,L,seed,ckpt', [(14, 3, 0, 0), (17, 2, 1, 0), (20, 2, 2, 0), (16, 4, 3, 2)])
@pytest.mark.parametrize('dtype', ['complex128', 'complex64'])
def test_xy_grad_run_on_pass_kernels(gpu, n, L, seed, ckpt, dtype):
    from qradient_amd.circuit_logic import McClean
    rs = np.random.RandomState(seed)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = _obs(n, rs)
    print('case', n, L, seed, ckpt, dtype, flush=True)
    circ = McClean(n, obs, L, axes=axes, angles=angles, dtype=dtype, ckpt_every=ckpt)
    E, g = circ.grad_run()
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    tol = TOL[dtype]
    assert abs(E - Eo) <= tol * max(1.0, abs(Eo)), (E, Eo)
    assert rel_err(g, go) <= tol, rel_err(g, go)
    Ef = circ.run_expec_val()
    assert abs(Ef - Eo) <= tol * max(1.0, abs(Eo))


@pytest.mark.parametrize('dtype', ['complex128', 'complex64'])
def test_xy_with_ini_state_small_n(gpu, dtype):
    """ini_state sends n <= 13 to the pass kernels too."""
    from qradient_amd.circuit_logic import McClean
    n, L = 9, 3
    rs = np.random.RandomState(5)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = _obs(n, rs)
    ini = rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)
    ini /= np.linalg.norm(ini)
    E, g = McClean(n, obs, L, axes=axes, angles=angles, dtype=dtype).grad_run(ini_state=ini)
    Eo, go = sv.mcclean_grad_run(n, axes, angles, obs, ini_state=ini)
    assert abs(E - Eo) <= TOL[dtype] * max(1.0, abs(Eo))
    assert rel_err(g, go) <= TOL[dtype]


def test_xy_graph_replay(gpu):
    from qradient_amd.circuit_logic import McClean
    n, L = 15, 2
    rs = np.random.RandomState(6)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = _obs(n, rs)
    direct = McClean(n, obs, L, axes=axes, angles=angles)
    graph = McClean(n, obs, L, axes=axes, angles=angles, graph=True)
    E0, g0 = direct.grad_run()
    for k in range(3):
        a = angles + 0.1 * k
        direct.angles = a
        graph.angles = a
        E0, g0 = direct.grad_run()
        E1, g1 = graph.grad_run()
        assert E1 == E0 and np.array_equal(g1, g0)
        Eo, go = sv.mcclean_grad_run_adjoint(n, axes, a, obs)
        assert rel_err(g1, go) <= 1e-12


def test_xy_profiled_run(gpu):
    """Per-launch profiling with the elementwise observable launch in the list."""
    from qradient_amd import _native
    from qradient_amd.circuit_logic import McClean
    n, L = 14, 2
    rs = np.random.RandomState(7)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    circ = McClean(n, _obs(n, rs), L, axes=axes, angles=angles)
    circ.grad_run()
    _native.call('qg_plan_set_profiling', circ._plan, 1)
    circ.grad_run()
    cnt = ctypes.c_int()
    _native.call('qg_plan_pass_times', circ._plan, 0, None, None, ctypes.byref(cnt))
    ms = (ctypes.c_float * cnt.value)()
    kinds = (ctypes.c_int * cnt.value)()
    _native.call('qg_plan_pass_times', circ._plan, cnt.value, ms, kinds, ctypes.byref(cnt))
    assert list(kinds).count(3) == 1 and all(m >= 0 for m in ms)
