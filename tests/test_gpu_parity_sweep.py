"""Seeded sweep over the McClean plan options against the oracle: qubit and
layer counts, dtype, entangler, memory mode (history or checkpoints every k
layers), tile shape (local bits, contiguous low bits) and complex64 tile
precision, drawn together so that combinations the targeted tests do not pair
up are covered too.  Bars: complex128 1e-12, complex64 1e-6 (norm-relative)."""
import ctypes

import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _cases(count=24, seed=2718):
    rs = np.random.RandomState(seed)
    out = []
    for k in range(count):
        n = int(rs.randint(14, 21))
        L = int(rs.randint(1, 6))
        dtype = ['complex128', 'complex64'][k % 2]
        ent = ['cnot', 'cz'][rs.randint(2)]
        ckpt = [0, 0, 1, 2, 3][rs.randint(5)]
        kmax = int(rs.randint(10, 14))
        cmin = int(rs.randint(2, 5))
        lds64 = int(rs.randint(0, 3)) if dtype == 'complex64' else -1
        out.append((k, n, L, dtype, ent, ckpt, kmax, cmin, lds64))
    return out


@pytest.mark.parametrize('k,n,L,dtype,ent,ckpt,kmax,cmin,lds64', _cases())
def test_plan_options_against_oracle(gpu, k, n, L, dtype, ent, ckpt, kmax, cmin, lds64):
    from qradient_amd import _native
    from qradient_amd.circuit_logic import McClean
    rs = np.random.RandomState(500 + k)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    obs['z'] = np.array([rs.standard_normal() if q % 3 == 1 else None for q in range(n)], dtype=object)
    print('sweep case', k, dict(n=n, L=L, dtype=dtype, ent=ent, ckpt=ckpt, kmax=kmax, cmin=cmin, lds64=lds64),
          flush=True)
    c = McClean(n, obs, L, axes=axes, angles=angles, dtype=dtype, entangler=ent, ckpt_every=ckpt)
    h = c._get_plan()
    _native.call('qg_plan_set_fused', h, 0)  # the pass kernels at every size
    _native.call('qg_plan_tune', h, kmax, cmin, 4, 4)
    if lds64 >= 0:
        _native.call('qg_plan_set_lds64', h, lds64)
    mode = ctypes.c_int()
    _native.call('qg_plan_ckpt_every', h, ctypes.byref(mode))
    assert (mode.value == 0) == (ckpt == 0)
    E, g = c.grad_run()
    c._destroy_plan()
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs, entangler=ent)
    tol = 1e-12 if dtype == 'complex128' else 1e-6
    assert abs(E - Eo) <= tol * max(1.0, abs(Eo)), (E, Eo)
    assert rel_err(g, go) <= tol, rel_err(g, go)
