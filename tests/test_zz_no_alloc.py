"""The allocation-free McClean grad run against the plan-based entry point.

qg_mcclean_grad_run_noalloc runs the whole-circuit kernel from one static
device array per GPU; qg_mcclean_grad_run with a plan of the same shape runs
the same kernel from plan allocations.  Both are called through the C ABI
(ctypes) and must agree bitwise (same kernel, same tables), and both within
1e-12 of the oracle, on two calls in a row, the second one smaller, so the
second call runs over the first call's leftovers in the shared array.
"""
import ctypes

import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _case(n, L, seed, entangler):
    rs = np.random.RandomState(seed)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    obs['x'] = np.array([0.3] + [None] * (n - 1), dtype=object)  # an x term: the kernel's x / y branch
    return axes, angles, obs


def _terms(obs):
    from qradient_amd.physical_components import Observable
    return Observable(obs).c_terms()


def _planned(ctx, n, L, ent, axes, angles, arr, k):
    from qradient_amd import _native
    plan = ctypes.c_void_p()
    _native.call('qg_mcclean_plan', ctx.handle, n, L, _native.QG_C128, ent, 0, ctypes.byref(plan))
    try:
        E = ctypes.c_double()
        g = np.zeros(L * n)
        _native.call('qg_mcclean_grad_run', plan, _native.iptr(axes), _native.dptr(angles), arr, k, None,
                     ctypes.byref(E), _native.dptr(g))
        return E.value, g.reshape(L, n)
    finally:
        _native.call('qg_plan_destroy', plan)


def _noalloc(ctx, n, L, ent, axes, angles, arr, k):
    from qradient_amd import _native
    E = ctypes.c_double()
    g = np.zeros(L * n)
    _native.call('qg_mcclean_grad_run_noalloc', ctx.handle, n, L, ent, _native.iptr(axes), _native.dptr(angles),
                 arr, k, ctypes.byref(E), _native.dptr(g))
    return E.value, g.reshape(L, n)


@pytest.mark.parametrize('entangler', ['cnot', 'cz'])
def test_noalloc_equals_planned_entry_point(gpu, entangler):
    from qradient_amd import _native
    ctx = _native.Context(0)
    ent = _native.ENTANGLERS[entangler]
    # the second call is smaller: fewer qubits, layers and terms
    for n, L, seed in ((13, 24, 5), (7, 3, 6)):
        axes, angles, obs = _case(n, L, seed, entangler)
        axes = np.ascontiguousarray(axes, dtype=np.int32)
        angles = np.ascontiguousarray(angles, dtype=np.float64)
        arr, k = _terms(obs)
        print('noalloc case n={} L={} seed={} entangler={}'.format(n, L, seed, entangler), flush=True)
        E1, g1 = _noalloc(ctx, n, L, ent, axes, angles, arr, k)
        E0, g0 = _planned(ctx, n, L, ent, axes, angles, arr, k)
        assert E1 == E0
        assert np.array_equal(g1, g0)
        Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs, entangler=entangler)
        assert abs(E1 - Eo) <= 1e-12 * max(abs(Eo), 1.0)
        assert rel_err(g1, go) <= 1e-12


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    from qradient_amd import _native
    ctx = _native.Context(0)
    arr, k = _native.terms_array([(_native.TERM_ZZ, 0, 1, 1.0)])
    E = ctypes.c_double()
    for n, L in ((14, 2), (13, 2000)):   # past the whole-circuit kernel; past the scratch
        axes = np.zeros(L * n, dtype=np.int32)
        angles = np.zeros(L * n)
        g = np.zeros(L * n)
        with pytest.raises(ValueError):
            _native.call('qg_mcclean_grad_run_noalloc', ctx.handle, n, L, 0, _native.iptr(axes),
                         _native.dptr(angles), arr, k, ctypes.byref(E), _native.dptr(g))
