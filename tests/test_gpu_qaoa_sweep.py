"""Seeded sweep over the batched QAOA plan options against the oracle: qubit
count, layers, batch size, chunk (circuits resident at once, so a batch may
span several chunks), tile shape (local bits, contiguous low bits),
workgroups per CU, the classical Hamiltonian's weights (MaxCut's unit
weights and small integers take the phase table, real weights the sincos
phase; with and without one-bit terms) and the initial states (|+>^n or
given), drawn together so that combinations the targeted tests do not pair
up are covered too.  Bars per circuit: complex128 1e-12, complex64 1e-6
(norm-relative gradient; |dE| relative to max(1, |E|)).

Given initial states are Haar-random for complex128 and random product
states for complex64.  From a Haar-random state the gradient is small against
the Hamiltonian's scale (norm 0.004 for a 15-qubit MaxCut case), and
complex64 storage alone moves it by more than the relative bar: rounding the
initial state to complex64 once, then running the FP64 oracle, is already
4.9e-7 from the FP64 result in that case, and the GPU path stores the state
once per pass (1.7e-6).  A random product state keeps the gradient of order
one while still exercising an arbitrary initial state."""
import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

WEIGHTS = ['maxcut', 'integer', 'real', 'integer_z']


def _cases(count=24, seed=1618):
    rs = np.random.RandomState(seed)
    out = []
    for k in range(count):
        n = int(rs.randint(14, 20))
        p = int(rs.randint(1, 4))
        dtype = ['complex128', 'complex64'][k % 2]
        B = int(rs.randint(1, 5))
        chunk = [0, 0, 1, 2][rs.randint(4)]
        kmax = [0, 9, 10, 11, 12, 13][rs.randint(6)]
        cmin = int(rs.randint(2, 5))
        wgpc = [2, 8, 16][rs.randint(3)]
        weights = WEIGHTS[rs.randint(len(WEIGHTS))]
        ini = bool(rs.randint(2))
        out.append((k, n, p, dtype, B, chunk, kmax, cmin, wgpc, weights, ini))
    return out


def _edges(n, rs):
    """A ring plus n // 2 chords, a < b, no repeats."""
    e = {(min(i, (i + 1) % n), max(i, (i + 1) % n)) for i in range(n)}
    while len(e) < n + n // 2:
        a, b = sorted(rs.choice(n, 2, replace=False))
        e.add((int(a), int(b)))
    return sorted(e)


def _observable(n, weights, rs):
    edges = _edges(n, rs)
    if weights == 'maxcut':
        return sv.maxcut_observable(n, edges)
    zz = np.full((n, n), None)
    for a, b in edges:
        zz[a, b] = float(rs.standard_normal()) if weights == 'real' else float(rs.choice([-2, -1, 1, 2]))
    obs = {'zz': zz}
    if weights in ('real', 'integer_z'):
        obs['z'] = np.array([(0.5 * rs.standard_normal() if weights == 'real' else float(rs.choice([-1, 1, 3])))
                             if q % 3 == 2 else None for q in range(n)], dtype=object)
    return obs


def _state(n, rs, haar):
    if haar:
        v = rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)
        return v / np.linalg.norm(v)
    v = np.ones(1, dtype=complex)
    for _ in range(n):
        a, ph = 0.5 * np.pi * rs.rand(), 2 * np.pi * rs.rand()
        v = np.kron(v, np.array([np.cos(a), np.exp(1j * ph) * np.sin(a)]))
    return v


@pytest.mark.parametrize('k,n,p,dtype,B,chunk,kmax,cmin,wgpc,weights,ini', _cases())
def test_batch_options_against_oracle(gpu, k, n, p, dtype, B, chunk, kmax, cmin, wgpc, weights, ini):
    from qradient_amd import _native
    from qradient_amd.batch import QaoaBatch
    rs = np.random.RandomState(900 + k)
    obs = [_observable(n, weights, rs) for _ in range(B)]
    betas, gammas = np.pi * rs.rand(B, p), np.pi * rs.rand(B, p)
    states = np.stack([_state(n, rs, dtype == 'complex128') for _ in range(B)]) if ini else None
    print('qaoa sweep case', k, dict(n=n, p=p, dtype=dtype, B=B, chunk=chunk, kmax=kmax, cmin=cmin, wgpc=wgpc,
                                      weights=weights, ini=ini), flush=True)
    batch = QaoaBatch(n, p, B, dtype=dtype)
    _native.call('qg_qaoa_tune', batch._plan, kmax, cmin, chunk, wgpc)
    E, grad = batch.grad_run(obs, betas, gammas, ini_states=states)
    tol = 1e-12 if dtype == 'complex128' else 1e-6
    for b in range(B):
        ini_b = None if states is None else states[b].copy()
        Eo, go = sv.qaoa_grad_run(n, obs[b], betas[b], gammas[b], ini_state=ini_b)
        assert abs(E[b] - Eo) <= tol * max(1.0, abs(Eo)), (b, E[b], Eo)
        assert rel_err(grad[b], go) <= tol, (b, rel_err(grad[b], go))
