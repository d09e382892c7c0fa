"""Writes the golden fixtures of this directory (run from the repo root:
``python tests/golden/make_golden.py``).

The reference cannot run at its snapshot and ships no vectors for this path,
so the expected values come from the oracle (oracle/statevector.py, complex128),
which tests/test_oracle_pins.py pins against closed forms, dense matrices and
finite differences.  The fixtures freeze the workloads and their values, so a
later change of the oracle, the input generators or the library shows up as a
mismatch (tests/test_golden.py on the CPU, tests/test_gpu_golden.py on the GPU).

Files (numpy .npz, no pickled objects):
  mcclean_n10_L10_z0z1.npz   McClean 10 qubits, 10 layers, observable Z0 Z1, seed 1234
                             (axes drawn before angles, mc_clean.py:13-14); E and the
                             100 gradients for the CNOT ladder and the CZ ladder
  mcclean_n14_L3_zzchain.npz McClean 14 qubits, 3 layers, N(0,1) ZZ chain plus Z terms
  qaoa_n8_p3_reg3.npz        QAOA 8 qubits, p = 3, a seeded 3-regular MaxCut graph
  qaoa_reg3_n20_b512.npz     the batched QAOA workload: 512 seeded 3-regular graphs on
                             20 vertices (edges a < b), betas and gammas in [0, pi)
  sampled_n5.npz             the sampled gradients (mc_clean.py:80-350, qaoa.py:83-158) of a
                             McClean 5-qubit, 2-layer circuit (positive ZZ-chain weights) and a
                             QAOA 5-qubit, p = 2 MaxCut circuit, each from numpy's global
                             RandomState seeded 31 just before the call: E and grad of
                             grad_run_with_component_sampling, sample_grad(_with_component_
                             sampling) and sample_grad_dense(_with_component_sampling) with 40
                             shots, and Qaoa.sample_grad_dense with 40 shots
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', '..'))

from oracle import statevector as sv  # noqa: E402


def z0z1(n):
    zz = np.full((n, n), None)
    zz[0, 1] = 1.0
    return {'zz': zz}


def obs_arrays(obs, n):
    """An observable dict as plain arrays (NaN where the reference has None)."""
    out = {}
    for key in ('z', 'zz'):
        if key in obs:
            out['obs_' + key] = np.array([[np.nan if v is None else float(v) for v in row]
                                          for row in np.atleast_2d(obs[key])]).reshape(np.shape(obs[key]))
    return out


def mcclean_cfg1():
    n, L = 10, 10
    rs = np.random.RandomState(1234)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = z0z1(n)
    res = {'axes': axes.astype(np.int32), 'angles': angles}
    for ent in ('cnot', 'cz'):
        E, g = sv.mcclean_grad_run(n, axes, angles, obs, entangler=ent)
        res['E_' + ent] = np.float64(E)
        res['grad_' + ent] = g
    res.update(obs_arrays(obs, n))
    np.savez(os.path.join(HERE, 'mcclean_n10_L10_z0z1.npz'), **res)


def mcclean_n14():
    n, L = 14, 3
    rs = np.random.RandomState(77)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, rs.standard_normal(n - 1))
    obs['z'] = np.array([0.25 * (q + 1) if q % 4 == 0 else None for q in range(n)], dtype=object)
    E, g = sv.mcclean_grad_run(n, axes, angles, obs)
    np.savez(os.path.join(HERE, 'mcclean_n14_L3_zzchain.npz'), axes=axes.astype(np.int32), angles=angles,
             E=np.float64(E), grad=g, **obs_arrays(obs, n))


def qaoa_small():
    n, p = 8, 3
    rs = np.random.RandomState(5)
    edges = np.array(sv.random_regular3_edges(n, rs), dtype=np.int32)
    betas, gammas = np.pi * rs.rand(p), np.pi * rs.rand(p)
    obs = sv.maxcut_observable(n, edges)
    E, g = sv.qaoa_grad_run(n, obs, betas, gammas)
    np.savez(os.path.join(HERE, 'qaoa_n8_p3_reg3.npz'), edges=edges, betas=betas, gammas=gammas,
             E=np.float64(E), grad=g)


def qaoa_batch_workload():
    from qradient_amd.optimization_problems import MaxCut
    n, p, B = 20, 8, 512
    rs = np.random.RandomState(2024)
    edges = np.array([MaxCut.regular3(n, rs).edge_set for _ in range(B)], dtype=np.int32)
    gammas = np.pi * rs.rand(B, p)
    betas = np.pi * rs.rand(B, p)
    np.savez(os.path.join(HERE, 'qaoa_reg3_n20_b512.npz'), edges=edges, betas=betas, gammas=gammas)


def sampled_n5():
    from oracle import reference_algorithm as ra
    n, L, shots, seed = 5, 2, 40, 31
    rs = np.random.RandomState(8)
    axes, angles = sv.mcclean_random_params(L, n, rs)
    obs = sv.zz_chain_observable(n, np.abs(rs.standard_normal(n - 1)) + 0.1)
    res = {'axes': axes.astype(np.int32), 'angles': angles}
    res.update(obs_arrays(obs, n))

    def seeded(key, fn):
        np.random.seed(seed)
        E, g = fn()
        res['E_' + key], res['grad_' + key] = np.float64(E), g

    def component_grad_run():
        dist = sv.weight_distribution(obs)
        c = np.random.choice(np.arange(dist.size), p=dist)
        return sv.mcclean_grad_run_component(n, axes, angles, obs, c)

    seeded('component', component_grad_run)
    seeded('shift', lambda: sv.mcclean_sample_grad(n, axes, angles, obs, shots))
    seeded('shift_component', lambda: sv.mcclean_sample_grad(n, axes, angles, obs, shots, component_sampling=True))
    seeded('dense', lambda: ra.mcclean_sample_grad_dense(n, obs, axes, angles, shots))
    seeded('dense_component', lambda: ra.mcclean_sample_grad_dense(n, obs, axes, angles, shots,
                                                                    component_sampling=True))
    edges = np.array([(0, 1), (1, 2), (2, 3), (3, 4), (0, 4), (1, 3)], dtype=np.int32)
    betas, gammas = np.pi * rs.rand(2), np.pi * rs.rand(2)
    res.update(edges=edges, betas=betas, gammas=gammas)
    seeded('qaoa_dense', lambda: ra.qaoa_sample_grad_dense(n, sv.maxcut_observable(n, edges), betas, gammas, shots))
    np.savez(os.path.join(HERE, 'sampled_n5.npz'), **res)


if __name__ == '__main__':
    mcclean_cfg1()
    mcclean_n14()
    qaoa_small()
    qaoa_batch_workload()
    sampled_n5()
    print('fixtures written to', HERE)
