"""The golden fixtures against the oracle and the input generators (CPU).

The fixtures were written by tests/golden/make_golden.py from the oracle; these
tests recompute them, with the matrix-free restatement and with the
reference's own formulation (CSR operators, stored history), so a drift of
either shows up here.
"""
import numpy as np

from oracle import reference_algorithm as ra
from oracle import statevector as sv
from tests.golden import fixtures as fx

TOL = 1e-13


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def test_mcclean_config1_fixture():
    d = fx.load('mcclean_n10_L10_z0z1.npz')
    obs = fx.obs_from(d)
    for ent in ('cnot', 'cz'):
        E, g = sv.mcclean_grad_run_adjoint(10, d['axes'], d['angles'], obs, entangler=ent)
        assert abs(E - d['E_' + ent]) <= TOL and _rel(g, d['grad_' + ent]) <= TOL
    ref = ra.McCleanRef(10, obs, d['axes'], d['angles'])
    ref.build()
    E, g = ref.grad_run()
    assert abs(E - d['E_cnot']) <= TOL and _rel(g, d['grad_cnot']) <= TOL


def test_mcclean_config1_inputs_follow_the_reference_draw_order():
    d = fx.load('mcclean_n10_L10_z0z1.npz')
    axes, angles = sv.mcclean_random_params(10, 10, np.random.RandomState(1234))
    assert np.array_equal(axes, d['axes']) and np.array_equal(angles, d['angles'])


def test_mcclean_n14_fixture():
    d = fx.load('mcclean_n14_L3_zzchain.npz')
    E, g = sv.mcclean_grad_run_adjoint(14, d['axes'], d['angles'], fx.obs_from(d))
    assert abs(E - d['E']) <= TOL * max(1.0, abs(d['E'])) and _rel(g, d['grad']) <= TOL


def test_qaoa_small_fixture():
    d = fx.load('qaoa_n8_p3_reg3.npz')
    obs = fx.maxcut_obs(8, d['edges'])
    E, g = sv.qaoa_grad_run_adjoint(8, obs, d['betas'], d['gammas'])
    assert abs(E - d['E']) <= TOL * max(1.0, abs(d['E'])) and _rel(g, d['grad']) <= TOL


def test_qaoa_batch_workload_fixture():
    """The 512-graph workload: reproduced by the library's seeded generator,
    and every graph is simple and 3-regular with edges a < b."""
    from qradient_amd.optimization_problems import MaxCut
    d = fx.load('qaoa_reg3_n20_b512.npz')
    rs = np.random.RandomState(2024)
    edges = np.array([MaxCut.regular3(20, rs).edge_set for _ in range(512)])
    assert np.array_equal(edges, d['edges'])
    assert np.array_equal(np.pi * rs.rand(512, 8), d['gammas'])
    assert np.array_equal(np.pi * rs.rand(512, 8), d['betas'])
    for e in d['edges']:
        assert np.all(e[:, 0] < e[:, 1]) and len({tuple(x) for x in e}) == 30
        assert np.array_equal(np.bincount(e.ravel(), minlength=20), np.full(20, 3))


def _sampled_inputs():
    d = fx.load('sampled_n5.npz')
    return d, fx.obs_from(d)


def test_sampled_gradients_fixture():
    """The oracle's seeded sampled gradients (mc_clean.py:80-350, qaoa.py:83-158)
    recomputed from the same seed equal the frozen values."""
    from oracle import reference_algorithm as ra
    d, obs = _sampled_inputs()
    n, shots = 5, 40
    axes, angles = d['axes'], d['angles']

    def seeded(fn):
        np.random.seed(31)
        return fn()

    def component():
        dist = sv.weight_distribution(obs)
        c = np.random.choice(np.arange(dist.size), p=dist)
        return sv.mcclean_grad_run_component(n, axes, angles, obs, c)

    runs = {'component': component,
            'shift': lambda: sv.mcclean_sample_grad(n, axes, angles, obs, shots),
            'shift_component': lambda: sv.mcclean_sample_grad(n, axes, angles, obs, shots, component_sampling=True),
            'dense': lambda: ra.mcclean_sample_grad_dense(n, obs, axes, angles, shots),
            'dense_component': lambda: ra.mcclean_sample_grad_dense(n, obs, axes, angles, shots,
                                                                    component_sampling=True),
            'qaoa_dense': lambda: ra.qaoa_sample_grad_dense(n, fx.maxcut_obs(n, d['edges']), d['betas'],
                                                            d['gammas'], shots)}
    for key, fn in runs.items():
        E, g = seeded(fn)
        assert abs(E - d['E_' + key]) <= 1e-13, key
        assert np.max(np.abs(g - d['grad_' + key])) <= 1e-13, key


def test_config2_fixtures_are_consistent():
    """The 24-qubit, 40-layer fixtures (too slow to recompute here): their
    inputs are the seed-2024 workload drawn in the reference's order, the two
    FP64 formulations agree to the recorded 1.7e-12, and the extended-precision
    values sit within that of both (each FP64 formulation rounds differently)."""
    from tests.golden.make_golden import mcclean_cfg2_workload
    d = fx.load('mcclean_n24_L40_zzchain.npz')
    ext = fx.load('mcclean_n24_L40_zzchain_ext.npz')
    n, L, axes, angles, w = mcclean_cfg2_workload()
    assert np.array_equal(d['axes'], axes) and np.array_equal(d['angles'], angles)
    assert np.array_equal(d['weights'], w)
    assert d['grad'].shape == ext['grad'].shape == (L, n)
    assert abs(_rel(d['grad_adjoint'], d['grad']) - float(d['formulation_rel_err'])) <= 1e-15
    assert 1e-13 < float(d['formulation_rel_err']) <= 1e-11
    assert _rel(d['grad'], ext['grad']) <= 1e-11 and _rel(d['grad_adjoint'], ext['grad']) <= 1e-11
    assert abs(float(d['E']) - float(ext['E'])) <= 1e-12 and abs(float(d['E_adjoint']) - float(ext['E'])) <= 1e-12


def test_qaoa_batch_expected_fixture_spot_check():
    """Two of the 512 oracle results of the batched QAOA workload, recomputed
    with the matrix-free oracle (the whole set takes minutes in C)."""
    w = fx.load('qaoa_reg3_n20_b512.npz')
    ref = fx.load('qaoa_reg3_n20_b512_expected.npz')
    assert ref['E'].shape == (512,) and ref['grad'].shape == (512, 8, 2)
    for b in (0, 511):
        E, g = sv.qaoa_grad_run_adjoint(20, fx.maxcut_obs(20, w['edges'][b]), w['betas'][b], w['gammas'][b])
        assert abs(E - ref['E'][b]) <= 1e-12 * max(1.0, abs(E)) and _rel(g, ref['grad'][b]) <= 1e-12
