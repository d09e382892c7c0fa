"""Per-launch profiling (qg_plan_set_profiling / qg_plan_pass_times), which
bench.py's roofline figure is computed from: every launch of every profiled
run carries HIP events, the times come back run after run when asked for,
and switching profiling on starts a new series."""
import ctypes

import numpy as np
import pytest

from oracle import statevector as sv

pytestmark = pytest.mark.gpu


def _times(native, plan):
    cnt = ctypes.c_int()
    native.call('qg_plan_pass_times', plan, 0, None, None, ctypes.byref(cnt))
    ms = (ctypes.c_float * cnt.value)()
    kinds = (ctypes.c_int * cnt.value)()
    native.call('qg_plan_pass_times', plan, cnt.value, ms, kinds, ctypes.byref(cnt))
    return np.array(ms[:]), np.array(kinds[:])


def test_profiled_runs_accumulate(gpu):
    from qradient_amd import _native
    from qradient_amd.circuit_logic import McClean
    n, L = 16, 3
    rs = np.random.RandomState(3)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    circ = McClean(n, obs, L, axes=axes, angles=angles, dtype='complex64')
    E0, g0 = circ.grad_run()
    _native.call('qg_plan_set_profiling', circ._plan, 1)
    for _ in range(3):
        E, g = circ.grad_run()
        assert E == E0 and np.array_equal(g, g0)   # profiling changes no result
    ms, kinds = _times(_native, circ._plan)
    per_run = len(kinds) // 3
    assert per_run > 0 and len(kinds) == 3 * per_run
    # the same launch kinds run after run: forward passes, the observable, adjoint passes
    assert np.array_equal(kinds[:per_run], kinds[per_run:2 * per_run])
    assert set(kinds.tolist()) <= {0, 1, 2} and 1 in kinds
    assert np.all(ms > 0)
    # switching profiling on again starts a new series
    _native.call('qg_plan_set_profiling', circ._plan, 1)
    circ.grad_run()
    ms2, kinds2 = _times(_native, circ._plan)
    assert np.array_equal(kinds2, kinds[:per_run])
    _native.call('qg_plan_set_profiling', circ._plan, 0)
