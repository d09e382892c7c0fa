"""The whole-circuit kernel (n <= 13, one workgroup per circuit) against the
oracle, and against the tile-pass kernels it replaces at those sizes.

Bars as for the pass kernels: complex128 within 1e-12 norm-relative of the
oracle, complex64 plans (FP64 throughout on this path) within 1e-6.
"""
import ctypes

import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _mcclean(n, L, axes, angles, obs, dtype='complex128', entangler='cnot', fused=1):
    from qradient_amd import _native
    from qradient_amd.physical_components import Observable
    ctx = _native.Context.default(0)
    h = ctypes.c_void_p()
    _native.call('qg_mcclean_plan', ctx.handle, n, L, _native.DTYPES[dtype], _native.ENTANGLERS[entangler], 0,
                 ctypes.byref(h))
    try:
        _native.call('qg_plan_set_fused', h, fused)
        arr, k = Observable(obs).c_terms()
        E = ctypes.c_double()
        g = np.empty((L, n))
        ax = np.ascontiguousarray(axes, dtype=np.int32)
        an = np.ascontiguousarray(angles, dtype=np.float64)
        _native.call('qg_mcclean_grad_run', h, _native.iptr(ax), _native.dptr(an), arr, k, None, ctypes.byref(E),
                     _native.dptr(g))
        return E.value, g
    finally:
        _native.load().qg_plan_destroy(h)


def _chain(n, rs):
    return sv.zz_chain_observable(n, rs.standard_normal(n - 1)) if n > 1 else {'z': np.array([0.7])}


@pytest.mark.parametrize('n,L,seed', [(1, 3, 0), (2, 2, 1), (3, 4, 2), (4, 3, 3), (5, 5, 4), (6, 7, 5), (7, 3, 6),
                                      (9, 5, 7), (11, 4, 8), (12, 6, 9), (13, 5, 10)])
@pytest.mark.parametrize('entangler', ['cnot', 'cz'])
def test_fused_mcclean_matches_oracle(gpu, n, L, seed, entangler):
    rs = np.random.RandomState(seed)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = _chain(n, rs)
    E, g = _mcclean(n, L, axes, angles, obs, entangler=entangler)
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs, entangler=entangler)
    assert abs(E - Eo) <= TOL * max(1.0, abs(Eo)), (E, Eo)
    assert rel_err(g, go) <= TOL, rel_err(g, go)


@pytest.mark.parametrize('n,L,seed', [(3, 3, 30), (8, 4, 31), (13, 3, 32)])
def test_fused_mcclean_xy_terms(gpu, n, L, seed):
    """x / y observable terms (observable.py:34-79) go through the fused path only."""
    rs = np.random.RandomState(seed)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = _chain(n, rs)
    obs['x'] = np.array([rs.standard_normal() if q % 2 == 0 else None for q in range(n)], dtype=object)
    obs['y'] = np.array([rs.standard_normal() if q % 3 == 1 else None for q in range(n)], dtype=object)
    E, g = _mcclean(n, L, axes, angles, obs)
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    assert abs(E - Eo) <= TOL * max(1.0, abs(Eo)), (E, Eo)
    assert rel_err(g, go) <= TOL, rel_err(g, go)


@pytest.mark.parametrize('n,L,seed', [(2, 3, 40), (7, 5, 41), (12, 4, 42), (13, 3, 43)])
@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-12), ('complex64', 1e-6)])
def test_fused_equals_pass_kernels(gpu, n, L, seed, dtype, tol):
    rs = np.random.RandomState(seed)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = _chain(n, rs)
    Ef, gf = _mcclean(n, L, axes, angles, obs, dtype=dtype, fused=1)
    Ep, gp = _mcclean(n, L, axes, angles, obs, dtype=dtype, fused=0)
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    assert rel_err(gp, go) <= tol, rel_err(gp, go)
    assert rel_err(gf, go) <= TOL, rel_err(gf, go)
    assert abs(Ef - Ep) <= tol * max(1.0, abs(Eo))


def _qaoa_batch(n, p, B, seed, fused=1):
    from qradient_amd import _native
    from qradient_amd.physical_components import Observable
    rs = np.random.RandomState(seed)
    ctx = _native.Context.default(0)
    h = ctypes.c_void_p()
    _native.call('qg_qaoa_plan', ctx.handle, n, p, _native.DTYPES['complex128'], B, ctypes.byref(h))
    obs_list, terms, off = [], [], [0]
    for b in range(B):
        edges = [(i, (i + 1 + b % 3) % n) for i in range(n)]
        zz = np.full((n, n), None)
        for i, j in edges:
            a, c = min(i, j), max(i, j)
            if a != c:
                zz[a, c] = float(rs.randint(1, 4)) if b % 2 == 0 else float(rs.standard_normal())
        obs = {'zz': zz}
        obs_list.append(obs)
        t = Observable(obs).terms
        terms += t
        off.append(off[-1] + len(t))
    allt, _ = _native.terms_array(terms)
    betas = rs.uniform(0, np.pi, (B, p))
    gammas = rs.uniform(0, np.pi, (B, p))
    E = np.empty(B)
    g = np.empty((B, p, 2))
    try:
        _native.call('qg_plan_set_fused', h, fused)
        _native.call('qg_qaoa_grad_run', h, B, allt, _native.iptr(np.array(off, dtype=np.int32)),
                     _native.dptr(np.ascontiguousarray(betas)), _native.dptr(np.ascontiguousarray(gammas)),
                     _native.dptr(E), _native.dptr(g))
    finally:
        _native.load().qg_plan_destroy(h)
    return obs_list, betas, gammas, E, g


@pytest.mark.parametrize('n,p,B,seed', [(3, 2, 3, 50), (8, 3, 5, 51), (13, 2, 4, 52)])
def test_fused_qaoa_batch_matches_oracle(gpu, n, p, B, seed):
    obs_list, betas, gammas, E, g = _qaoa_batch(n, p, B, seed)
    for b in range(B):
        Eo, go = sv.qaoa_grad_run_adjoint(n, obs_list[b], betas[b], gammas[b])
        assert abs(E[b] - Eo) <= TOL * max(1.0, abs(Eo)), (b, E[b], Eo)
        assert rel_err(g[b], go) <= TOL, (b, rel_err(g[b], go))
    _, _, _, Ep, gp = _qaoa_batch(n, p, B, seed, fused=0)
    assert rel_err(gp, g) <= TOL
    assert np.max(np.abs(Ep - E)) <= TOL * max(1.0, np.max(np.abs(E)))
