"""The C restatement of the reference algorithm (oracle/c/oracle_ref.c), which
bench.py times as the CPU baseline, against the numpy oracle
(oracle/statevector.py, itself pinned in test_oracle_pins.py): McClean with
both observable kinds and the truncated-layer sample the benchmark times, and
QAOA.  Bar: 1e-12 norm-relative (complex128 on both sides)."""
import numpy as np
import pytest

from oracle import c_oracle
from oracle import statevector as sv
from tests.conftest import rel_err


@pytest.fixture(scope='module', autouse=True)
def built():
    c_oracle.build()


@pytest.mark.parametrize('n,L,seed', [(2, 3, 0), (5, 4, 1), (9, 6, 2), (12, 3, 3)])
def test_mcclean_c_equals_numpy(n, L, seed):
    rs = np.random.RandomState(seed)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    obs['z'] = np.array([0.5 if q % 3 == 0 else None for q in range(n)], dtype=object)
    E, g = c_oracle.mcclean_grad_run(n, axes, angles, obs)
    Eo, go = sv.mcclean_grad_run(n, axes, angles, obs)
    assert abs(E - Eo) <= 1e-12 * max(1.0, abs(Eo))
    assert rel_err(g, go) <= 1e-12


def test_mcclean_c_truncated_sample_is_the_shorter_circuit():
    """layers=Ls runs the first Ls layers as a circuit of its own: the bounded
    sample of the CPU baseline is a real grad_run of Ls layers."""
    n, L, Ls = 8, 6, 3
    rs = np.random.RandomState(4)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    E, g = c_oracle.mcclean_grad_run(n, axes, angles, obs, layers=Ls)
    Eo, go = sv.mcclean_grad_run(n, axes[:Ls], angles[:Ls], obs)
    assert g.shape == (Ls, n)
    assert abs(E - Eo) <= 1e-12 and rel_err(g, go) <= 1e-12


@pytest.mark.parametrize('n,p,seed', [(4, 2, 5), (8, 3, 6), (10, 5, 7)])
def test_qaoa_c_equals_numpy(n, p, seed):
    rs = np.random.RandomState(seed)
    edges = sv.random_regular3_edges(n, rs) if n % 2 == 0 else [(0, 1), (1, 2), (2, 3), (0, 3)]
    obs = sv.maxcut_observable(n, edges)
    betas, gammas = np.pi * rs.rand(p), np.pi * rs.rand(p)
    E, g = c_oracle.qaoa_grad_run(n, obs, betas, gammas)
    Eo, go = sv.qaoa_grad_run(n, obs, betas, gammas)
    assert abs(E - Eo) <= 1e-12 * max(1.0, abs(Eo))
    assert rel_err(g, go) <= 1e-12


@pytest.mark.parametrize('n,L,seed', [(6, 5, 11), (12, 6, 12), (16, 4, 13)])
def test_extended_precision_build_is_the_arbiter(n, L, seed):
    """The x87 build of the same algorithm agrees with both FP64 formulations
    to well within 1e-12 at small depth, and is closer to the matrix-free
    adjoint (fewer roundings per gradient) than the FP64 history order is."""
    rs = np.random.RandomState(seed)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    Ex, gx = c_oracle.mcclean_grad_run(n, axes, angles, obs, extended=True)
    E, g = c_oracle.mcclean_grad_run(n, axes, angles, obs)
    Ea, ga = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    assert abs(Ex - E) <= 1e-14 and abs(Ex - Ea) <= 1e-14
    assert rel_err(g, gx) <= 1e-13 and rel_err(ga, gx) <= 1e-13


def test_extended_precision_qaoa():
    n, p = 8, 3
    rs = np.random.RandomState(14)
    obs = sv.maxcut_observable(n, sv.random_regular3_edges(n, rs))
    betas, gammas = np.pi * rs.rand(p), np.pi * rs.rand(p)
    Ex, gx = c_oracle.qaoa_grad_run(n, obs, betas, gammas, extended=True)
    Eo, go = sv.qaoa_grad_run(n, obs, betas, gammas)
    assert abs(Ex - Eo) <= 1e-13 and rel_err(go, gx) <= 1e-13
