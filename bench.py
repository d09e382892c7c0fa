"""Benchmark: McClean grad_run, 24 qubits, 40 layers (960 parameters).

Workload: seeded random Pauli axes and uniform angles in [0, 2 pi), observable
sum_i w_i Z_i Z_{i+1} with w ~ N(0, 1), CNOT-ladder entangler.  One "step" is
one full grad_run (forward sweep, lambda = O psi, adjoint sweep; E plus all 960
gradients), inputs resident on the device.  ``value`` is the complex64-state
(FP64 arithmetic and reductions) throughput in circuits/s; the plain
complex128 method is timed on the same circuit and reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  ``roofline`` is measured live: HIP events around every
pass launch of the timed steps (on the library's stream) give the average
duration of the dominant kernel; its algorithmic bytes per launch come from
the pass's state reads and writes.  ``cpu_baseline`` times the C restatement
of the reference algorithm (oracle/c/oracle_ref.c) on one core over a bounded
sample (the first layer of the same circuit) and scales it to the full
circuit.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_QUBITS, N_LAYERS, SEED = 24, 40, 2024
HBM_PEAK_GBS = 8000.0


def workload():
    rs = np.random.RandomState(SEED)
    axes = (3 * rs.rand(N_LAYERS, N_QUBITS)).astype(int)
    angles = 2 * np.pi * rs.rand(N_LAYERS, N_QUBITS)
    w = rs.standard_normal(N_QUBITS - 1)
    zz = np.full((N_QUBITS, N_QUBITS), None)
    for i in range(N_QUBITS - 1):
        zz[i, i + 1] = float(w[i])
    return axes, angles, {'zz': zz}


def pass_times(circ, native):
    cnt = ctypes.c_int()
    native.call('qg_plan_pass_times', circ._plan, 0, None, None, ctypes.byref(cnt))
    ms = (ctypes.c_float * cnt.value)()
    kinds = (ctypes.c_int * cnt.value)()
    native.call('qg_plan_pass_times', circ._plan, cnt.value, ms, kinds, ctypes.byref(cnt))
    return np.array(ms[:]), np.array(kinds[:])


def time_steps(circ, steps, warmup, sync, profile):
    from qradient_amd import _native
    for _ in range(warmup):
        circ.grad_run()
    if profile:
        _native.call('qg_plan_set_profiling', circ._plan, 1)
    per = []
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        E, g = circ.grad_run()
    sync()
    dt = time.perf_counter() - t0
    if profile:
        # every launch of the timed steps carried HIP events on the library's
        # stream; their times are read now, outside the timed region
        per.append(pass_times(circ, _native))
        _native.call('qg_plan_set_profiling', circ._plan, 0)
    return dt, E, g, per


def roofline(per, state_bytes, steps):
    """Dominant pass class: reverse two-state passes (read A, read B, write B).
    ``per``: (ms, kinds) of every launch of the ``steps`` timed steps."""
    ms = np.concatenate([p[0] for p in per])
    kinds = np.concatenate([p[1] for p in per])
    alg = {0: 2 * state_bytes, 1: 3 * state_bytes, 2: 3 * state_bytes}
    totals = {k: ms[kinds == k].sum() for k in set(kinds.tolist())}
    dom = max(totals, key=totals.get)
    avg_ms = float(ms[kinds == dom].mean())
    achieved = alg[dom] / (avg_ms * 1e-3) / 1e9
    return dom, avg_ms, achieved, alg[dom], {int(k): float(v) / steps for k, v in totals.items()}


def cpu_baseline(axes, angles, obs):
    from threadpoolctl import threadpool_limits
    from oracle import c_oracle
    layers = 1
    with threadpool_limits(1):
        t0 = time.perf_counter()
        c_oracle.mcclean_grad_run(N_QUBITS, axes, angles, obs, layers=layers)
        dt = time.perf_counter() - t0
    per_circuit = dt * N_LAYERS / layers
    return {'value': 1.0 / per_circuit, 'unit': 'circuits/s', 'cores': 1, 'kind': 'port',
            'sample': 'first {} of {} layers of the same 24-qubit circuit (forward + adjoint sweep of that '
                      'layer, reference algorithm with stored history), C restatement oracle/c/oracle_ref.c, '
                      '1 thread; {:.1f} s measured, scaled x{} to the full circuit'.format(layers, N_LAYERS, dt,
                                                                                        N_LAYERS // layers)}


def _line(metric, value, unit, steps, warmup, ms, dtype, config, **extra):
    out = {'metric': metric, 'value': value, 'unit': unit, 'n_gpus': 1, 'steps': steps, 'warmup': warmup,
           'ms_per_step': ms, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': dtype,
           'data': 'synthetic (seeded)', 'config': config}
    out.update(extra)
    return out


def bench_qaoa(args, sync):
    """512 seeded random 3-regular MaxCut graphs, 20 qubits, p = 8: one batched grad_run per step."""
    from qradient_amd.batch import QaoaBatch
    from qradient_amd.optimization_problems import MaxCut
    n, p, B = 20, 8, 512
    # the workload's graphs and angles, committed as a fixture (written by
    # tests/golden/make_golden.py with the library's seeded generator)
    wl = np.load(os.path.join(ROOT, 'tests', 'golden', 'qaoa_reg3_n20_b512.npz'), allow_pickle=False)
    obs = [MaxCut(n, edge_set=wl['edges'][b]).to_observable() for b in range(B)]
    gammas = wl['gammas']
    betas = wl['betas']
    res = {}
    for dtype in ('complex64', 'complex128'):
        qb = QaoaBatch(n, p, B, dtype=dtype)
        packed = qb.pack(obs)
        steps = args.steps if dtype == 'complex64' else max(2, args.steps // 2)
        for _ in range(args.warmup):
            qb.grad_run(packed, betas, gammas)
        sync()
        t0 = time.perf_counter()
        dev = 0.0
        for _ in range(steps):
            E, g = qb.grad_run(packed, betas, gammas)
            dev += qb.last_device_ms()
        sync()
        dt = time.perf_counter() - t0
        res[dtype] = dict(ms=dt / steps * 1e3, dev_ms=dev / steps, E=E, g=g, steps=steps)
        del qb
    c64, c128 = res['complex64'], res['complex128']
    rel = np.linalg.norm((c64['g'] - c128['g']).reshape(B, -1), axis=1) / np.linalg.norm(
        c128['g'].reshape(B, -1), axis=1)
    sbytes = (2 ** n) * 8
    b_alg = (6 * p + 3) * sbytes * B
    return _line('QAOA grad_run circuits/s (512 x MaxCut 3-regular, 20 qubits, p=8, complex64 state + FP64 '
                 'reductions)', B * 1e3 / c64['ms'], 'circuits/s', c64['steps'], args.warmup, c64['ms'],
                 'c64 (FP64 arithmetic and reductions)',
                 {'workload': 'qaoa_maxcut_reg3_n20_p8_batch512', 'qubits': n, 'layers': p, 'batch': B,
                  'seed': SEED, 'graphs': 'tests/golden/qaoa_reg3_n20_b512.npz'},
                 device_ms_per_step=c64['dev_ms'],
                 c128_plain={'circuits_per_s': B * 1e3 / c128['ms'], 'ms_per_step': c128['ms']},
                 c64_vs_c128_grad_rel_err_max=float(rel.max()),
                 effective_gbs=b_alg / (c64['dev_ms'] * 1e-3) / 1e9,
                 b_alg_bytes=b_alg)


def bench_serve(args, sync):
    """Poisson stream of McClean requests (n ~ U{12..28}, L ~ U{4..64}) through GradServer.

    First the saturation throughput is measured on a burst (every request
    queued at once), then a Poisson trace at 0.7 x that rate is replayed and
    its latency percentiles are reported.
    """
    from qradient_amd import serve
    nburst = max(40, 8 * args.steps)
    burst = serve.poisson_trace(nburst, 1e9, SEED + 1, (12, 28), (4, 64))
    sat = serve.replay(burst, slots=8, progress=100)['circuits_per_s']
    rate = 0.7 * sat
    nreq = max(100, 20 * args.steps)
    trace = serve.poisson_trace(nreq, rate, SEED, (12, 28), (4, 64))
    t0 = time.perf_counter()
    st = serve.replay(trace, slots=8, progress=100)
    dt = time.perf_counter() - t0
    return _line('serving grad_run circuits/s with p50/p99 latency (McClean n~U{12..28}, L~U{4..64}, Poisson '
                 'arrivals at 0.7 x the saturation throughput)', st['circuits_per_s'], 'circuits/s', nreq, 0,
                 dt / nreq * 1e3, 'c64 (FP64 arithmetic and reductions)',
                 {'workload': 'serve_mcclean_poisson', 'requests': nreq, 'rate_per_s': rate, 'slots': 8,
                  'seed': SEED}, saturation_circuits_per_s=sat,
                 latency_ms={'p50': st['p50_ms'], 'p99': st['p99_ms'], 'mean': st['mean_ms']},
                 peak_bytes=st['peak_bytes'], budget_bytes=st['budget_bytes'])


def bench_large(args, sync):
    """McClean 30 qubits, 100 layers: complex64 with checkpoints every k layers vs complex128."""
    from qradient_amd.circuit_logic import McClean
    n, L = 30, 100
    rs = np.random.RandomState(SEED)
    axes = (3 * rs.rand(L, n)).astype(int)
    angles = 2 * np.pi * rs.rand(L, n)
    zz = np.full((n, n), None)
    for i, w in enumerate(rs.standard_normal(n - 1)):
        zz[i, i + 1] = float(w)
    obs = {'zz': zz}
    from qradient_amd import _native
    res = {}
    for dtype in ('complex64', 'complex128'):
        # the library picks the checkpoint period: the smallest that fits HBM;
        # one untimed run first (plan, workspace allocation, first touch)
        circ = McClean(n, obs, L, axes=axes, angles=angles, dtype=dtype)
        circ.grad_run()
        sync()
        t0 = time.perf_counter()
        E, g = circ.grad_run()
        sync()
        dt = time.perf_counter() - t0
        k = ctypes.c_int32()
        _native.call('qg_plan_ckpt_every', circ._plan, ctypes.byref(k))
        t64 = ctypes.c_int32()
        _native.call('qg_plan_lds64', circ._plan, ctypes.byref(t64))
        res[dtype] = dict(ms=dt * 1e3, E=E, g=g, k=int(k.value), lds64=int(t64.value))
        circ._destroy_plan()
    c64, c128 = res['complex64'], res['complex128']
    rel = float(np.linalg.norm(c64['g'] - c128['g']) / np.linalg.norm(c128['g']))
    sb = (2 ** n) * 8
    b_alg = (6 * L + 3 + 2 * -(-L // c64['k'])) * sb
    return _line('grad_run circuits/s (McClean 30 qubits, 100 layers, complex64 state + FP64 reductions, '
                 'checkpoints every %d layers)' % c64['k'], 1e3 / c64['ms'], 'circuits/s', 1, 0, c64['ms'],
                 'c64 (FP64 arithmetic and reductions)',
                 {'workload': 'mcclean_grad_run_n30_L100_zzchain', 'qubits': n, 'layers': L, 'seed': SEED,
                  'ckpt_every': c64['k'], 'fp64_lds_tiles': bool(c64['lds64'])},
                 c128_plain={'ms_per_step': c128['ms'], 'ckpt_every': c128['k']},
                 c64_vs_c128_grad_rel_err=rel, effective_gbs=b_alg / (c64['ms'] * 1e-3) / 1e9, b_alg_bytes=b_alg)


def bench_sample(args, sync):
    """McClean.sample_grad (shot-based parameter shift): 24 qubits, 8 layers, complex64, 100 shots per term.

    One step is one sample_grad call: the unshifted circuit, then 2 * 8 * 24 = 384 shifted circuits, each
    restarted from the stored state after the layer before its shifted one and measured for the 23 ZZ terms
    on the device; the shots are drawn on the host.  Algorithmic bytes per call: 2 S per layer of every
    circuit run (one read and one write of the state per layer, the ideal single pass) and S per 16 terms
    measured, with S = 2^24 * 8 bytes."""
    from qradient_amd.circuit_logic import McClean
    n, L, shots = 24, 8, 100
    rs = np.random.RandomState(SEED)
    axes = (3 * rs.rand(L, n)).astype(int)
    angles = 2 * np.pi * rs.rand(L, n)
    zz = np.full((n, n), None)
    for i, w in enumerate(rs.standard_normal(n - 1)):
        zz[i, i + 1] = float(w)
    obs = {'zz': zz}
    circ = McClean(n, obs, L, axes=axes, angles=angles, dtype='complex64')
    for _ in range(max(1, args.warmup)):
        circ.sample_grad(shot_num=shots)
    steps = max(1, args.steps // 2)
    sync()
    t0 = time.perf_counter()
    dev = 0.0
    for _ in range(steps):
        np.random.seed(SEED)
        E, g = circ.sample_grad(shot_num=shots)
        dev += circ.last_device_ms()
    sync()
    dt = (time.perf_counter() - t0) / steps
    dev /= steps
    S = (2 ** n) * 8
    layers = L + sum(2 * n * (L - i) for i in range(L))
    launches_per_shift = -(-(n - 1) // 16)
    b_alg = 2 * layers * S + 2 * n * L * launches_per_shift * S
    achieved = b_alg / (dev * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        # one layer of one shifted circuit in the numpy oracle, scaled by the layers of the call
        from oracle import statevector as sv
        v = sv.mcclean_forward(n, axes[:1], angles[:1])
        t1 = time.perf_counter()
        sv.entangle(v, 'cnot')
        for q in range(n):
            sv.rot(v, axes[1, q], angles[1, q], q)
        t_layer = time.perf_counter() - t1
        cpu = {'value': 1.0 / (t_layer * layers), 'unit': 'sample_grad calls/s', 'cores': 1, 'kind': 'port',
               'sample': 'one layer of one shifted 24-qubit circuit in the numpy oracle (oracle/statevector.py), '
                         '%.2f s, scaled x%d to the layers of one call; term measurement and shots excluded'
                         % (t_layer, layers)}
    return _line('sample_grad calls/s (McClean 24 qubits, 8 layers, complex64, parameter shift with 100 shots '
                 'per term)', 1.0 / dt, 'sample_grad calls/s', steps, args.warmup, dt * 1e3,
                 'c64 (FP64 arithmetic and reductions)',
                 {'workload': 'mcclean_sample_grad_n24_L8_zzchain_shots100', 'qubits': n, 'layers': L,
                  'shifted_circuits': 2 * n * L, 'shots': shots, 'seed': SEED},
                 device_ms_per_step=dev,
                 roofline={'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                           'frac': achieved / HBM_PEAK_GBS, 'traffic': None,
                           'kernel': 'whole call (forward passes of every circuit, term measurements)',
                           'algorithmic_bytes_per_step': b_alg},
                 cpu_baseline=cpu, E=float(E))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--workload', default='mcclean24',
                    choices=['mcclean24', 'qaoa512', 'serve', 'mcclean30', 'sample24'],
                    help='mcclean24 (default, the headline line); qaoa512: 512 batched QAOA MaxCut grad_runs '
                         '(20 qubits, p=8); serve: Poisson request stream through the serving loop; mcclean30: '
                         'McClean 30 qubits, 100 layers, complex64 checkpoints vs complex128; sample24: '
                         'shot-based parameter-shift gradient, 24 qubits, 8 layers')
    ap.add_argument('--probe', action='store_true',
                    help='short run for hardware-counter collection: two complex64 grad_runs of the workload '
                         'and one plain 1 GiB device copy (the calibration kernel for FETCH_SIZE / WRITE_SIZE)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU: this build has no multi-GPU path (every configuration is a '
              'single-GPU workload); run with --gpus 1', file=sys.stderr)
        sys.exit(2)

    import torch
    from qradient_amd.circuit_logic import McClean

    def sync():
        torch.cuda.synchronize()

    if args.workload != 'mcclean24':
        print(json.dumps({'qaoa512': bench_qaoa, 'serve': bench_serve, 'mcclean30': bench_large,
                          'sample24': bench_sample}[args.workload](args, sync)))
        return
    axes, angles, obs = workload()
    if args.probe:
        circ = McClean(N_QUBITS, obs, N_LAYERS, axes=axes, angles=angles, dtype='complex64')
        for _ in range(2):
            circ.grad_run()
        circ._destroy_plan()
        nbytes = 1 << 30
        src = torch.ones(nbytes // 4, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        sync()
        dst.copy_(src)
        sync()
        print(json.dumps({'probe': 'mcclean_grad_run_n24_L40 complex64 x2, then copy', 'copy_bytes': nbytes}))
        return
    res = {}
    grads = {}
    for dtype in ('complex64', 'complex128'):
        circ = McClean(N_QUBITS, obs, N_LAYERS, axes=axes, angles=angles, dtype=dtype)
        steps = args.steps if dtype == 'complex64' else max(2, args.steps // 2)
        dt, E, g, per = time_steps(circ, steps, args.warmup, sync, profile=True)
        sb = (2 ** N_QUBITS) * (8 if dtype == 'complex64' else 16)
        res[dtype] = dict(ms=dt / steps * 1e3, E=E, roof=roofline(per, sb, steps), steps=steps)
        grads[dtype] = g
        circ._destroy_plan()
    c64, c128 = res['complex64'], res['complex128']
    rel = float(np.linalg.norm(grads['complex64'] - grads['complex128']) / np.linalg.norm(grads['complex128']))
    dom, avg_ms, achieved, alg_bytes, totals = c64['roof']
    traffic = None
    tfile = os.path.join(ROOT, 'profiles', 'traffic_mcclean_n24_L40.json')
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get('c64_reverse_pass_bytes')
    out = {
        'metric': 'grad_run circuits/s (McClean 24 qubits, 40 layers, complex64 state + FP64 reductions)',
        'value': 1e3 / c64['ms'],
        'unit': 'circuits/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': c64['ms'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'c64 (FP64 arithmetic and reductions)',
        'data': 'synthetic (seeded random McClean circuit and N(0,1) ZZ-chain weights)',
        'config': {'workload': 'mcclean_grad_run_n24_L40_zzchain', 'qubits': N_QUBITS, 'layers': N_LAYERS,
                   'parameters': N_QUBITS * N_LAYERS, 'entangler': 'cnot', 'seed': SEED},
        'c128_plain': {'circuits_per_s': 1e3 / c128['ms'], 'ms_per_step': c128['ms'], 'steps': c128['steps'],
                       'roofline_achieved_gbs': c128['roof'][2]},
        'speedup_c64_vs_c128': c128['ms'] / c64['ms'],
        'c64_vs_c128_grad_rel_err': rel,
        'E': c64['E'],
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK_GBS, 'traffic': traffic,
                     'kernel': {0: 'forward pass', 1: 'adjoint two-state pass', 2: 'forward + observable pass'}[dom],
                     'avg_launch_ms': avg_ms, 'algorithmic_bytes_per_launch': alg_bytes,
                     # measured HBM bytes of the kernel over its live launch time
                     'traffic_gbs': (traffic / (avg_ms * 1e-3) / 1e9) if traffic else None,
                     'ms_per_step_by_kind': totals},
    }
    if not args.no_cpu_baseline:
        out['cpu_baseline'] = cpu_baseline(axes, angles, obs)
    print(json.dumps(out))


if __name__ == '__main__':
    main()
