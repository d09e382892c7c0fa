"""GPU parity of McClean.grad_run / run_expec_val against the CPU oracle.

Bar: complex128 (the plain method) within 1e-12 norm-relative of the oracle's
complex128 gradient and |dE| <= 1e-12 * max(|E|, sum|w|); complex64 (FP64
arithmetic, complex64 state) within 1e-6 norm-relative.
"""
import numpy as np
import pytest

from oracle import statevector as sv
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

C128_TOL = 1e-12
C64_TOL = 1e-6


def _case(n, L, seed, obs_kind='chain'):
    rs = np.random.RandomState(seed)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    if obs_kind == 'chain':
        obs = sv.zz_chain_observable(n, rs.standard_normal(max(n - 1, 0))) if n > 1 else {'z': np.array([0.7])}
    elif obs_kind == 'z0z1':
        zz = np.full((n, n), None)
        zz[0, 1] = 1.0
        obs = {'zz': zz}
    elif obs_kind == 'zmix':
        obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
        obs['z'] = np.array([rs.standard_normal() if q % 2 == 0 else None for q in range(n)], dtype=object)
    return axes, angles, obs


def _run(n, L, axes, angles, obs, dtype='complex128', entangler='cnot', ini_state=None):
    from qradient_amd.circuit_logic import McClean
    c = McClean(n, obs, L, axes=axes, angles=angles, dtype=dtype, entangler=entangler)
    return c.grad_run(ini_state=ini_state)


@pytest.mark.parametrize('n,L,seed', [(1, 3, 0), (2, 2, 1), (3, 4, 2), (4, 3, 3), (5, 5, 4), (8, 6, 5),
                                      (10, 10, 6), (12, 5, 7), (13, 4, 8), (14, 4, 9), (16, 3, 10),
                                      (17, 2, 11), (20, 2, 12)])
@pytest.mark.parametrize('entangler', ['cnot', 'cz'])
def test_grad_run_c128_matches_oracle(gpu, n, L, seed, entangler):
    axes, angles, obs = _case(n, L, seed)
    E, g = _run(n, L, axes, angles, obs, entangler=entangler)
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs, entangler=entangler)
    wsum = sum(abs(float(w)) for w in obs.get('zz', np.array([])).ravel() if w is not None) + 1.0
    assert abs(E - Eo) <= C128_TOL * max(abs(Eo), wsum), (E, Eo)
    assert g.shape == (L, n)
    assert rel_err(g, go) <= C128_TOL, rel_err(g, go)


@pytest.mark.parametrize('n,L,seed', [(6, 4, 20), (10, 10, 21), (14, 6, 22), (18, 4, 23)])
def test_grad_run_c64_within_tolerance(gpu, n, L, seed):
    axes, angles, obs = _case(n, L, seed)
    E, g = _run(n, L, axes, angles, obs, dtype='complex64')
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    assert abs(E - Eo) <= C64_TOL * max(abs(Eo), 1.0)
    assert rel_err(g, go) <= C64_TOL, rel_err(g, go)


def test_config1_z0z1(gpu):
    """McClean n=10, L=10, observable Z0Z1, complex128: E and all 100 gradients."""
    axes, angles, obs = _case(10, 10, 1234, 'z0z1')
    E, g = _run(10, 10, axes, angles, obs)
    Eo, go = sv.mcclean_grad_run(10, axes, angles, obs)
    assert abs(E - Eo) <= C128_TOL
    assert rel_err(g, go) <= C128_TOL


def test_z_terms_and_ini_state(gpu):
    n, L = 7, 3
    axes, angles, obs = _case(n, L, 31, 'zmix')
    rs = np.random.RandomState(5)
    ini = rs.standard_normal(2 ** n) + 1j * rs.standard_normal(2 ** n)
    ini /= np.linalg.norm(ini)
    keep = ini.copy()
    E, g = _run(n, L, axes, angles, obs, ini_state=ini)
    Eo, go = sv.mcclean_grad_run(n, axes, angles, obs, ini_state=ini)
    assert np.array_equal(ini, keep)
    assert abs(E - Eo) <= 1e-12 * max(1.0, abs(Eo))
    assert rel_err(g, go) <= C128_TOL


def test_run_expec_val(gpu):
    from qradient_amd.circuit_logic import McClean
    n, L = 9, 4
    axes, angles, obs = _case(n, L, 41)
    c = McClean(n, obs, L, axes=axes, angles=angles)
    assert abs(c.run_expec_val() - sv.mcclean_run_expec_val(n, axes, angles, obs)) <= 1e-12


def test_repeated_calls_are_bitwise_identical(gpu):
    n, L = 15, 3
    axes, angles, obs = _case(n, L, 51)
    from qradient_amd.circuit_logic import McClean
    c = McClean(n, obs, L, axes=axes, angles=angles)
    E1, g1 = c.grad_run()
    E2, g2 = c.grad_run()
    assert E1 == E2 and np.array_equal(g1, g2)


def test_invalid_axis_raises(gpu):
    from qradient_amd.circuit_logic import McClean
    axes, angles, obs = _case(4, 2, 3)
    axes[1, 2] = 3
    c = McClean(4, obs, 2, axes=axes, angles=angles)
    with pytest.raises(ValueError):
        c.grad_run()


@pytest.mark.parametrize('n,L,k', [(5, 6, 1), (10, 7, 3), (15, 5, 2), (18, 6, 4), (13, 4, 4)])
@pytest.mark.parametrize('entangler', ['cnot', 'cz'])
def test_checkpoint_uncompute_mode_c128(gpu, n, L, k, entangler):
    """psi uncomputed in the adjoint sweep, restarted from checkpoints every k layers."""
    from qradient_amd.circuit_logic import McClean
    axes, angles, obs = _case(n, L, 60 + n)
    c = McClean(n, obs, L, axes=axes, angles=angles, entangler=entangler, ckpt_every=k)
    E, g = c.grad_run()
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs, entangler=entangler)
    assert abs(E - Eo) <= C128_TOL * max(1.0, abs(Eo))
    assert rel_err(g, go) <= C128_TOL


@pytest.mark.parametrize('n,L,k', [(12, 12, 4), (17, 8, 3)])
def test_checkpoint_uncompute_mode_c64(gpu, n, L, k):
    from qradient_amd.circuit_logic import McClean
    axes, angles, obs = _case(n, L, 70 + n)
    c = McClean(n, obs, L, axes=axes, angles=angles, dtype='complex64', ckpt_every=k)
    E, g = c.grad_run()
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    assert abs(E - Eo) <= C64_TOL * max(1.0, abs(Eo))
    assert rel_err(g, go) <= C64_TOL


def test_tune_applies_cached_shape(gpu, tmp_path, monkeypatch):
    """tune=True times the tile shapes once, caches the best, and the tuned
    plan keeps the complex128 parity bar."""
    from qradient_amd import autotune
    from qradient_amd.circuit_logic import McClean
    monkeypatch.setenv('QRADIENT_AMD_TUNE_CACHE', str(tmp_path / 'tune.json'))
    n, L = 17, 3
    axes, angles, obs = _case(n, L, 81)
    c = McClean(n, obs, L, axes=axes, angles=angles, tune=True)
    E, g = c.grad_run()
    shape = autotune.cached('mcclean', n, L, 'complex128', 'cnot')
    assert shape is not None and shape['kmax'] in (10, 11, 12, 13)
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    assert abs(E - Eo) <= C128_TOL * max(1.0, abs(Eo))
    assert rel_err(g, go) <= C128_TOL


def _lds64(c, mode=None):
    import ctypes
    from qradient_amd import _native
    h = c._get_plan()
    if mode is not None:
        _native.call('qg_plan_set_lds64', h, mode)
    v = ctypes.c_int()
    _native.call('qg_plan_lds64', h, ctypes.byref(v))
    return v.value


@pytest.mark.parametrize('n,L,k,mode', [(14, 6, 0, 1), (17, 5, 2, 1), (20, 3, 0, 1), (14, 6, 0, 2), (17, 5, 2, 2),
                                         (19, 4, 0, 2)])
def test_fp64_tiles_c64_matches_oracle(gpu, n, L, k, mode):
    """complex64 passes with FP64 LDS tiles (mode 1: every pass, mode 2: the
    forward passes), history and checkpoint modes."""
    from qradient_amd.circuit_logic import McClean
    axes, angles, obs = _case(n, L, 90 + n)
    c = McClean(n, obs, L, axes=axes, angles=angles, dtype='complex64', ckpt_every=k)
    assert _lds64(c, mode) == mode
    E, g = c.grad_run()
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    assert abs(E - Eo) <= C64_TOL * max(1.0, abs(Eo))
    assert rel_err(g, go) <= C64_TOL


def test_fp64_tiles_chosen_for_long_circuits(gpu):
    from qradient_amd.circuit_logic import McClean
    axes, angles, obs = _case(16, 101, 7)
    long = McClean(16, obs, 101, axes=axes, angles=angles, dtype='complex64')   # 2 passes x 101 layers
    assert _lds64(long) == 2  # FP64 tiles in the forward passes
    short = McClean(16, obs, 99, axes=axes[:99], angles=angles[:99], dtype='complex64')
    assert _lds64(short) == 0
    c128 = McClean(16, obs, 101, axes=axes, angles=angles)
    assert _lds64(c128) == 0


def test_fp64_tiles_reduce_c64_error(gpu):
    """On a long circuit the FP64 tiles (one rounding per pass) stay closer to
    the complex128 oracle than FP32 tiles (a rounding per transpose too), and
    FP64 tiles in the forward passes alone recover most of that."""
    from qradient_amd.circuit_logic import McClean
    n, L = 16, 90
    axes, angles, obs = _case(n, L, 95)
    c = McClean(n, obs, L, axes=axes, angles=angles, dtype='complex64')
    _, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    errs = {}
    for mode in (0, 1, 2):
        _lds64(c, mode)
        errs[mode] = rel_err(c.grad_run()[1], go)
    print('  FP32 tiles {0:.3e}, FP64 tiles {1:.3e}, FP64 forward tiles {2:.3e}'.format(*[errs[m] for m in (0, 1, 2)]),
          flush=True)
    assert max(errs.values()) <= C64_TOL
    assert errs[1] < errs[0] and errs[2] < errs[0], errs


@pytest.mark.parametrize('n,L,entangler', [(15, 4, 'cnot'), (18, 3, 'cz'), (20, 3, 'cnot')])
def test_wide_tiles_c64_matches_oracle(gpu, n, L, entangler):
    """14 local bits per pass (1024-thread workgroups) for complex64 history
    plans; the observable then takes a gate-free pass of its own."""
    import ctypes
    from qradient_amd import _native
    from qradient_amd.circuit_logic import McClean
    axes, angles, obs = _case(n, L, 110 + n)
    c = McClean(n, obs, L, axes=axes, angles=angles, dtype='complex64', entangler=entangler)
    _native.call('qg_plan_tune', c._get_plan(), 14, 3, 4, 4)
    E, g = c.grad_run()
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs, entangler=entangler)
    assert abs(E - Eo) <= C64_TOL * max(1.0, abs(Eo))
    assert rel_err(g, go) <= C64_TOL
    E2, g2 = c.grad_run()
    assert E2 == E and np.array_equal(g2, g)
    assert abs(c.run_expec_val() - Eo) <= C64_TOL * max(1.0, abs(Eo))
    c.grad_run()
    psi = c.final_state()
    ref = sv.mcclean_forward(n, axes, angles, entangler=entangler)
    assert np.max(np.abs(psi - ref)) <= 1e-5


@pytest.mark.parametrize('n,L', [(17, 3), (22, 2)])
@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-12), ('complex64', 1e-6)])
def test_z_terms_on_pass_path(gpu, n, L, dtype, tol):
    """One-bit terms with the observable on the last (high-bit) pass."""
    axes, angles, obs = _case(n, L, 120 + n, 'zmix')
    E, g = _run(n, L, axes, angles, obs, dtype=dtype)
    Eo, go = sv.mcclean_grad_run_adjoint(n, axes, angles, obs)
    assert abs(E - Eo) <= tol * max(1.0, abs(Eo))
    assert rel_err(g, go) <= tol


@pytest.mark.parametrize('n,lds64', [(28, -1), (29, 1)])
def test_states_of_4_gib_and_more(gpu, n, lds64):
    """64-bit addressing (states of 4 GiB and more): complex128 at n = 28 and
    both dtypes at n = 29, checked against each other (the oracle is too slow
    at this size); complex64 at n = 29 with FP64 tiles."""
    from qradient_amd.circuit_logic import McClean
    L = 1
    axes, angles, obs = _case(n, L, 130 + n)
    c128 = McClean(n, obs, L, axes=axes, angles=angles)
    E1, g1 = c128.grad_run()
    c128._destroy_plan()
    c64 = McClean(n, obs, L, axes=axes, angles=angles, dtype='complex64')
    _lds64(c64, lds64)
    E2, g2 = c64.grad_run()
    c64._destroy_plan()
    assert abs(E1 - E2) <= C64_TOL * max(1.0, abs(E1))
    assert rel_err(g2, g1) <= C64_TOL
    assert np.all(np.isfinite(g1)) and np.linalg.norm(g1) > 0


def test_largest_complex64_state(gpu):
    """31 qubits in complex64 (16 GiB per state, 2^31 amplitudes: the largest
    state whose plan fits the device's budget), 2 layers.  No oracle and no
    complex128 run fit at this size, so the checks are what the mathematics
    fixes: the exact +-pi/2 parameter-shift rule (mc_clean.py:139-155) on one
    parameter in each pass's bit range, from complex64 forward runs, and the
    analytic zeros of the last layer's Z-axis gradients under a diagonal
    observable."""
    import gc
    import torch
    from qradient_amd.circuit_logic import McClean
    gc.collect()               # plans of earlier tests' circuits
    torch.cuda.empty_cache()   # blocks torch keeps reserved
    n, L = 31, 2
    axes, angles, obs = _case(n, L, 131)
    axes[L - 1, :4] = 2  # Z-axis rotations on the last layer: zero gradients
    c = McClean(n, obs, L, axes=axes, angles=angles.copy(), dtype='complex64')
    print('largest state: n=31 L=2 complex64 grad_run', flush=True)
    E, g = c.grad_run()
    gmax = float(np.max(np.abs(g)))
    assert gmax > 0 and np.all(np.isfinite(g))
    assert np.max(np.abs(g[L - 1][axes[L - 1] == 2])) <= 1e-6 * gmax
    # complex64 expectations carry about 1e-7 of the observable's weight
    scale = max(gmax, sum(abs(float(w)) for w in obs['zz'].ravel() if w is not None))
    for (i, q) in [(0, 30), (0, 15), (0, 2), (1, 20)]:   # bits 0, 15, 28 and 10
        base = c.angles[i, q]
        vals = []
        for s in (np.pi / 2, -np.pi / 2):
            c.angles[i, q] = base + s
            vals.append(c.run_expec_val())
        c.angles[i, q] = base
        shift = 0.5 * (vals[0] - vals[1])
        print('  shift check ({}, {}): adjoint {:.6e} shift {:.6e}'.format(i, q, g[i, q], shift), flush=True)
        assert abs(shift - g[i, q]) <= 1e-5 * scale, (i, q, shift, g[i, q])
    c._destroy_plan()


def test_plan_mode_and_bits_do_not_depend_on_free_memory(gpu):
    """The planner picks history or checkpoints from the circuit shape and the
    device's total memory.  A 26-qubit, 64-layer complex64 circuit keeps its
    per-pass history (3 passes x 64 layers + 4 states of 512 MiB, 98 GiB) on an
    idle device; a plan made while an unrelated allocation leaves only that
    much plus 10 GiB free must choose the same mode and give bitwise the same
    result (a planner reading free memory would switch to checkpoints)."""
    import ctypes
    import torch
    from qradient_amd import _native
    from qradient_amd.circuit_logic import McClean
    n, L = 26, 64
    axes, angles, obs = _case(n, L, 41)
    hist_bytes = (3 * L + 4) * (2 ** n) * 8

    def run():
        c = McClean(n, obs, L, axes=axes, angles=angles, dtype='complex64')
        E, g = c.grad_run()
        k = ctypes.c_int32()
        _native.call('qg_plan_ckpt_every', c._get_plan(), ctypes.byref(k))
        c._destroy_plan()
        return E, g, k.value

    E0, g0, k0 = run()
    assert k0 == 0  # history
    free, total = torch.cuda.mem_get_info(0)
    hold_bytes = free - hist_bytes - (10 << 30)
    hold = torch.empty(hold_bytes // 4, dtype=torch.float32, device='cuda')
    try:
        free_now, _ = torch.cuda.mem_get_info(0)
        assert 0.85 * free_now < hist_bytes  # the old free-memory rule would checkpoint here
        E1, g1, k1 = run()
    finally:
        del hold
        torch.cuda.empty_cache()
    assert k0 == k1
    assert E0 == E1 and np.array_equal(g0, g1)
