"""qg_plan_reshape: one McClean plan re-laid out for a sequence of circuit
shapes gives bitwise the results of a fresh plan per shape, and once its
tables have grown to the largest shape a second cycle through the shapes
allocates no device memory (the workspace is supplied by the caller)."""
import ctypes

import numpy as np
import pytest

from oracle import statevector as sv

pytestmark = pytest.mark.gpu

SHAPES = [(16, 12), (20, 6), (14, 30), (13, 5), (18, 20), (22, 4), (12, 9), (21, 3)]


def _case(n, L, seed):
    rs = np.random.RandomState(seed)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    terms = [(3, i, i + 1, float(w)) for i, w in enumerate(rs.standard_normal(n - 1))]
    return np.ascontiguousarray(axes, dtype=np.int32), np.ascontiguousarray(angles), terms


def _run(plan, n, L, axes, angles, terms):
    from qradient_amd import _native
    arr, k = _native.terms_array(terms)
    E = ctypes.c_double()
    g = np.empty(L * n)
    _native.call('qg_mcclean_grad_run', plan, _native.iptr(axes), _native.dptr(angles), arr, k, None,
                 ctypes.byref(E), _native.dptr(g))
    return E.value, g


@pytest.mark.parametrize('dtype', ['complex64', 'complex128'])
def test_reshape_equals_fresh_plans_and_stops_allocating(gpu, dtype):
    import torch
    from qradient_amd import _native
    ctx = _native.Context(0)
    dt = _native.DTYPES[dtype]
    cases = {s: _case(s[0], s[1], 100 + i) for i, s in enumerate(SHAPES)}
    fresh = {}
    for (n, L), (axes, angles, terms) in cases.items():
        h = ctypes.c_void_p()
        _native.call('qg_mcclean_plan', ctx.handle, n, L, dt, 0, 0, ctypes.byref(h))
        try:
            fresh[(n, L)] = _run(h, n, L, axes, angles, terms)
        finally:
            _native.call('qg_plan_destroy', h)
    plan = ctypes.c_void_p()
    n0, L0 = SHAPES[0]
    _native.call('qg_mcclean_plan', ctx.handle, n0, L0, dt, 0, 0, ctypes.byref(plan))
    try:
        # the workspace: one caller-owned buffer, large enough for every shape
        need = 0
        for (n, L) in SHAPES:
            _native.call('qg_plan_reshape', plan, n, L, 0)
            b = ctypes.c_size_t()
            _native.call('qg_plan_workspace_size', plan, ctypes.byref(b))
            need = max(need, b.value)
        ws = torch.empty(need, dtype=torch.uint8, device='cuda')
        free_after = []
        for cycle in range(2):
            for (n, L) in SHAPES:
                print('reshape cycle {} shape n={} L={} {}'.format(cycle, n, L, dtype), flush=True)
                _native.call('qg_plan_reshape', plan, n, L, 0)
                _native.call('qg_plan_set_workspace', plan, ctypes.c_void_p(ws.data_ptr()), need)
                axes, angles, terms = cases[(n, L)]
                E, g = _run(plan, n, L, axes, angles, terms)
                E0, g0 = fresh[(n, L)]
                assert E == E0 and np.array_equal(g, g0), (n, L)
            torch.cuda.synchronize()
            free_after.append(torch.cuda.mem_get_info(0)[0])
        assert free_after[1] == free_after[0]
    finally:
        _native.call('qg_plan_destroy', plan)
