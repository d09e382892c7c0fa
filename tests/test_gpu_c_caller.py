"""The C ABI from a plain C program (tests/c_caller/c_caller.c, linked against
libqradient_amd.so): a McClean grad run called directly and through the
serving loop gives bitwise the same E and gradient, and both agree with the
oracle (complex128 within 1e-12, complex64 within 1e-6, norm-relative)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'c_caller', 'c_caller')


@pytest.mark.parametrize('n,L,dtype,tol', [(10, 4, 0, 1e-12), (16, 5, 0, 1e-12), (16, 5, 1, 1e-6)])
def test_c_program_direct_and_served(gpu, tmp_path, n, L, dtype, tol):
    if not os.path.exists(BIN):
        pytest.fail('tests/c_caller/c_caller is not built: run __graft_entry__.build()')
    rs = np.random.RandomState(300 + n + dtype)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    w = rs.standard_normal(n - 1)
    path = tmp_path / 'circuit.txt'
    with open(path, 'w') as f:
        f.write('{} {}\n'.format(n, L))
        f.write(' '.join(str(int(a)) for a in axes.ravel()) + '\n')
        f.write(' '.join(repr(float(a)) for a in angles.ravel()) + '\n')
        f.write(' '.join(repr(float(x)) for x in w) + '\n')
    print('c_caller n={} L={} dtype={}'.format(n, L, dtype), flush=True)
    r = subprocess.run([BIN, str(path), str(dtype)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = {line.split()[0]: line.split()[1:] for line in r.stdout.splitlines()}
    direct = np.array([float(x) for x in rows['direct']])
    served = np.array([float(x) for x in rows['served']])
    assert np.array_equal(direct, served)
    E, g = direct[0], direct[1:].reshape(L, n)
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, sv.zz_chain_observable(n, w))
    assert abs(E - Eo) <= tol * max(1.0, abs(Eo)) and rel_err(g, go) <= tol, (E, Eo, rel_err(g, go))
