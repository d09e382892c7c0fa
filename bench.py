"""Benchmark: McClean grad_run, 24 qubits, 40 layers (960 parameters).

Workload: seeded random Pauli axes and uniform angles in [0, 2 pi), observable
sum_i w_i Z_i Z_{i+1} with w ~ N(0, 1), CNOT-ladder entangler.  One "step" is
one full grad_run (forward sweep, lambda = O psi, adjoint sweep; E plus all 960
gradients), inputs resident on the device.  ``value`` is the complex64-state
(FP64 arithmetic and reductions) throughput in circuits/s; the plain
complex128 method is timed on the same circuit and reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ...]

Prints one JSON line.  ``roofline`` follows the algorithmic byte count of the
adjoint method, each statevector counted once per layer: per layer the
forward sweep reads and writes psi (2 S) and the reverse sweep reads and
writes psi and lambda (4 S), plus S for the initial state and 2 S for
lambda = O psi with E, so B_alg = (6 L + 3) S for the whole run (S = 2^n
amplitudes x bytes per amplitude).  A launch that covers one of the P passes
of a layer is credited 1/P of its layer's bytes.  ``roofline.achieved`` is
the dominant kernel's algorithmic bytes per launch over its average launch
time (HIP events on the library's stream around every launch of the last
timed step: events between all launches would add about 0.8 ms per step);
``roofline.whole_step`` is B_alg over ms_per_step.  ``hbm`` holds the
bytes the kernel actually moves per launch (rocprofv3 FETCH_SIZE +
WRITE_SIZE counters of the same workload, from the committed profile named in
``hbm.traffic_source``: not measured in this run) and the rate they give.
``copy_ceiling`` is measured in the run: a streaming copy between two 1 GiB
buffers (HBM) and between two buffers of one state's size (which the 256 MB
memory-side cache can hold).  ``cpu_baseline`` times the C restatement of the
reference algorithm (oracle/c/oracle_ref.c) on one core over a bounded sample
(the first 3 layers of the same circuit, median of 3 runs) and scales it to
the full circuit.
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
    """Times `steps` grad runs.  With `profile`, the last timed step records a
    HIP event on the library's stream around each of its launches (an event
    per launch adds about 3 us between dependent kernels, so only one step of
    the timed region carries them); their times are read after the region."""
    from qradient_amd import _native
    for _ in range(warmup):
        circ.grad_run()
    per = []
    sync()
    t0 = time.perf_counter()
    for k in range(steps):
        if profile and k == steps - 1:
            _native.call('qg_plan_set_profiling', circ._plan, 1)
        E, g = circ.grad_run()
    sync()
    dt = time.perf_counter() - t0
    if profile:
        per.append(pass_times(circ, _native))
        _native.call('qg_plan_set_profiling', circ._plan, 0)
    return dt, E, g, per


def roofline(per, state_bytes, steps, layers):
    """Dominant launch class of the timed steps and its algorithmic rate.

    ``per``: (ms, kinds) of every launch of the ``steps`` profiled steps; kinds
    0 forward pass, 1 adjoint two-state pass, 2 last forward pass fused with
    lambda = O psi and E.  With P passes per layer, a forward launch is
    credited 2 S / P, an adjoint launch 4 S / P and the fused one 2 S / P + 2 S.
    """
    ms = np.concatenate([p[0] for p in per])
    kinds = np.concatenate([p[1] for p in per])
    P = int(round(float(np.sum(kinds == 1)) / (steps * layers)))
    S = state_bytes
    alg = {0: 2 * S / P, 1: 4 * S / P, 2: 2 * S / P + 2 * S}
    totals = {k: ms[kinds == k].sum() for k in set(kinds.tolist())}
    dom = max((k for k in totals if k in alg), key=totals.get)
    avg_ms = float(ms[kinds == dom].mean())
    achieved = alg[dom] / (avg_ms * 1e-3) / 1e9
    return dict(kind=dom, avg_ms=avg_ms, achieved=achieved, alg_bytes=alg[dom], passes_per_layer=P,
                ms_by_kind={int(k): float(v) / steps for k, v in totals.items()})


# Launch shapes of the streaming copy tried for the ceiling: workgroups per CU
# and ordinary / non-temporal accesses.  On MI355X one 256-thread workgroup per
# CU streams fastest (5.5-6.1 TB/s on 1-4 GiB buffers against 4.4-4.8 at 8 per
# CU): a grid-stride loop over fewer, longer-lived workgroups keeps the whole
# machine's accesses in one contiguous window.
COPY_SHAPES = [(gpc, nt) for gpc in (1, 2, 8) for nt in (0, 1)]


def _best_copy(ctx, a, b, nbytes, reps=4):
    from qradient_amd import _native
    best = None
    for gpc, nt in COPY_SHAPES:
        ms = ctypes.c_double()
        for r in range(reps):
            src, dst = (a, b) if r % 2 == 0 else (b, a)
            _native.call('qg_stream_copy', ctx.handle, ctypes.c_void_p(dst.data_ptr()),
                         ctypes.c_void_p(src.data_ptr()), nbytes, nt, gpc, 2, ctypes.byref(ms))
            gbs = 2 * nbytes / (ms.value * 1e-3) / 1e9
            if best is None or gbs > best[0]:
                best = (gbs, {'grid_per_cu': gpc, 'nontemporal': nt})
    return best


def copy_ceiling(state_bytes, sync):
    """Streaming-copy rates on this device (qg_stream_copy, the best of the
    launch shapes in COPY_SHAPES): between two 1 GiB buffers (HBM-bound), and
    between two buffers of one state's size ping-ponged (the memory-side cache
    holds such a pair at 24 qubits in complex64)."""
    import torch
    from qradient_amd import _native
    ctx = _native.Context.default(0)
    out = {}
    for name, nbytes in (('hbm_1gib', 1 << 30), ('state_pingpong', int(state_bytes))):
        a = torch.empty(nbytes // 4, dtype=torch.float32, device='cuda')
        b = torch.empty_like(a)
        a.fill_(1.0)
        sync()
        gbs, shape = _best_copy(ctx, a, b, nbytes)
        out[name + '_gbs'] = gbs
        out[name + '_shape'] = shape
        del a, b
    torch.cuda.empty_cache()
    return out


def cpu_baseline(axes, angles, obs, layers=3, runs=3):
    """The C restatement of the reference algorithm on one core: the first
    ``layers`` layers of the circuit (forward with stored history, then the
    adjoint sweep of those layers), median of ``runs`` runs, scaled to the
    circuit's layers."""
    import os as _os
    from threadpoolctl import threadpool_limits
    from oracle import c_oracle
    times = []
    with threadpool_limits(1):
        for _ in range(runs):
            t0 = time.perf_counter()
            c_oracle.mcclean_grad_run(N_QUBITS, axes, angles, obs, layers=layers)
            times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    per_circuit = med * N_LAYERS / layers
    return {'value': 1.0 / per_circuit, 'unit': 'circuits/s', 'cores': 1, 'kind': 'port',
            'sample': 'first {} of {} layers of the same 24-qubit circuit (forward with stored history + adjoint '
                      'sweep of those layers, the reference algorithm), C restatement oracle/c/oracle_ref.c on 1 '
                      'thread; median of {} runs {:.2f} s (runs: {}), scaled x{:.2f} to the full circuit; host has '
                      '{} cores'.format(layers, N_LAYERS, runs, med, ', '.join('%.2f' % t for t in times),
                                        N_LAYERS / layers, len(_os.sched_getaffinity(0)))}


def _line(metric, value, unit, steps, warmup, ms, dtype, config, **extra):
    out = {'metric': metric, 'value': value, 'unit': unit, 'n_gpus': 1, 'steps': steps, 'warmup': warmup,
           'ms_per_step': ms, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': dtype,
           'data': 'synthetic (seeded)', 'config': config}
    out.update(extra)
    return out


def _qaoa_cpu_job(job):
    from oracle import c_oracle
    from oracle import statevector as sv
    n, edges, betas, gammas = job
    c_oracle.qaoa_grad_run(n, sv.maxcut_observable(n, edges), betas, gammas)
    return 0


def qaoa_cpu_baseline(wl, n, count=32):
    """The C restatement of the reference algorithm (oracle/c/oracle_ref.c,
    qaoa.py:40-70 with stored history) over the first ``count`` circuits of
    the workload: independent circuits, one per process, single-threaded each,
    on up to 16 host cores.  Runs before any device work (the workers are
    forked)."""
    import multiprocessing as mp
    from threadpoolctl import threadpool_limits
    from oracle import c_oracle
    c_oracle.load()  # loaded once, inherited by the workers
    P = min(16, len(os.sched_getaffinity(0)))
    jobs = [(n, wl['edges'][b], wl['betas'][b], wl['gammas'][b]) for b in range(count)]
    with threadpool_limits(1):
        with mp.get_context('fork').Pool(P) as pool:
            t0 = time.perf_counter()
            pool.map(_qaoa_cpu_job, jobs, chunksize=1)
            dt = time.perf_counter() - t0
    return {'value': count / dt, 'unit': 'circuits/s', 'cores': P, 'kind': 'port',
            'sample': 'first {} of the 512 circuits (20 qubits, p = 8), C restatement oracle/c/oracle_ref.c, one '
                      'circuit per process on {} processes of 1 thread; {:.2f} s wall; host has {} cores'.format(
                          count, P, dt, len(os.sched_getaffinity(0)))}


def bench_qaoa(args, sync):
    """512 seeded random 3-regular MaxCut graphs, 20 qubits, p = 8: one batched grad_run per step."""
    from qradient_amd.batch import QaoaBatch
    from qradient_amd.optimization_problems import MaxCut
    n, p, B = 20, 8, 512
    # the workload's graphs and angles, committed as a fixture (written by
    # tests/golden/make_golden.py with the library's seeded generator)
    wl = np.load(os.path.join(ROOT, 'tests', 'golden', 'qaoa_reg3_n20_b512.npz'), allow_pickle=False)
    cpu = None if args.no_cpu_baseline else qaoa_cpu_baseline(wl, n)
    obs = [MaxCut(n, edge_set=wl['edges'][b]).to_observable() for b in range(B)]
    gammas = wl['gammas']
    betas = wl['betas']
    res = {}
    shapes = {}
    for dtype in ('complex64', 'complex128'):
        # the autotuner's batch shape (tile, chunk, workgroups per CU): cached,
        # else the committed MI355X default, else timed now
        qb = QaoaBatch(n, p, B, dtype=dtype, tune=True)
        shapes[dtype] = qb.shape
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
                  'seed': SEED, 'graphs': 'tests/golden/qaoa_reg3_n20_b512.npz', 'tuned_shape': shapes['complex64']},
                 device_ms_per_step=c64['dev_ms'],
                 c128_plain={'circuits_per_s': B * 1e3 / c128['ms'], 'ms_per_step': c128['ms'],
                             'tuned_shape': shapes['complex128']},
                 c64_vs_c128_grad_rel_err_max=float(rel.max()),
                 grad_rel_err_vs_oracle=qaoa_oracle_errors(res),
                 effective_gbs=b_alg / (c64['dev_ms'] * 1e-3) / 1e9,
                 b_alg_bytes=b_alg, hbm=qaoa_traffic(c64['dev_ms']), **({'cpu_baseline': cpu} if cpu else {}))


def qaoa_traffic(dev_ms):
    """Bytes the batch's passes move (committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    counts of this workload, not measured in this run) over the device time."""
    path = os.path.join(ROOT, 'profiles', 'traffic_qaoa_n20_p8_b512.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    gbs = t['c64_batch_bytes'] / (dev_ms * 1e-3) / 1e9
    return {'traffic_bytes_per_batch': t['c64_batch_bytes'], 'sustained_gbs': gbs, 'sustained_frac': gbs / HBM_PEAK_GBS,
            'traffic_source': 'profiles/traffic_qaoa_n20_p8_b512.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of '
                              'this workload, scaled by a copy-kernel calibration; committed, not measured in this run'}


def qaoa_oracle_errors(res):
    """Worst norm-relative gradient error over the 512 circuits of each dtype
    against the oracle fixture (tests/golden/qaoa_reg3_n20_b512_expected.npz)."""
    path = os.path.join(ROOT, 'tests', 'golden', 'qaoa_reg3_n20_b512_expected.npz')
    if not os.path.exists(path):
        return None
    ref = np.load(path, allow_pickle=False)['grad']
    out = {}
    for dt, r in res.items():
        d = (r['g'] - ref).reshape(ref.shape[0], -1)
        out[dt] = float(np.max(np.linalg.norm(d, axis=1) / np.linalg.norm(ref.reshape(ref.shape[0], -1), axis=1)))
    return out


# Serving configuration: 4 slots (one per hardware queue the runtime gives a
# process by default, so a small request's kernels do not wait behind a large
# request's in one queue), shortest-first admission, a request overtaken 32
# times served next.  Measured at 18.5 requests/s against 8 slots in arrival
# order: p50 86 ms vs 124 ms, p99 1.92 s vs 2.35 s, the same throughput.
SERVE = dict(slots=4, policy='shortest', backfill_limit=32)
SERVE_FIFO8 = dict(slots=8, policy='fifo', backfill_limit=8)


def bench_serve(args, sync):
    """Poisson stream of McClean requests (n ~ U{12..28}, L ~ U{4..64}) through GradServer.

    First the saturation throughput is measured on a burst (every request
    queued at once), then a Poisson trace at 0.7 x that rate is replayed and
    its latency percentiles are reported; the same trace is replayed once more
    with 8 slots in arrival order, for comparison.
    """
    from qradient_amd import serve
    nburst = max(40, 8 * args.steps)
    burst = serve.poisson_trace(nburst, 1e9, SEED + 1, (12, 28), (4, 64))
    sat = serve.replay(burst, progress=100, **SERVE)['circuits_per_s']
    rate = 0.7 * sat
    nreq = max(100, 20 * args.steps)
    trace = serve.poisson_trace(nreq, rate, SEED, (12, 28), (4, 64))
    t0 = time.perf_counter()
    st = serve.replay(trace, progress=100, **SERVE)
    dt = time.perf_counter() - t0
    fifo = serve.replay(trace, progress=100, **SERVE_FIFO8)
    return _line('serving grad_run circuits/s with p50/p99 latency (McClean n~U{12..28}, L~U{4..64}, Poisson '
                 'arrivals at 0.7 x the saturation throughput)', st['circuits_per_s'], 'circuits/s', nreq, 0,
                 dt / nreq * 1e3, 'c64 (FP64 arithmetic and reductions)',
                 dict({'workload': 'serve_mcclean_poisson', 'requests': nreq, 'rate_per_s': rate, 'seed': SEED},
                      **SERVE), saturation_circuits_per_s=sat,
                 latency_ms={'p50': st['p50_ms'], 'p99': st['p99_ms'], 'mean': st['mean_ms']},
                 peak_bytes=st['peak_bytes'], budget_bytes=st['budget_bytes'],
                 same_trace_8_slots_arrival_order={'circuits_per_s': fifo['circuits_per_s'], 'p50_ms': fifo['p50_ms'],
                                                   'p99_ms': fifo['p99_ms'], 'mean_ms': fifo['mean_ms']})


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
        # the autotuner's checkpoint period for this shape (cached, else the
        # committed MI355X default, else timed now); one untimed run first
        # (plan, workspace allocation, first touch)
        from qradient_amd import autotune
        k_tuned = autotune.lookup(autotune.ckpt_key(n, L, dtype, 'cnot'))
        if k_tuned is None:
            k_tuned, _ = autotune.tune_ckpt(n, L, dtype)
        circ = McClean(n, obs, L, axes=axes, angles=angles, dtype=dtype, ckpt_every=int(k_tuned))
        # the tuned tile shape of this shape (cached or committed), else the library default
        shape = autotune.cached('mcclean', n, L, dtype, 'cnot')
        if shape is not None:
            autotune.apply(circ._get_plan(), shape)
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
        res[dtype] = dict(ms=dt * 1e3, E=E, g=g, k=int(k.value), lds64=int(t64.value), shape=shape)
        circ._destroy_plan()
    c64, c128 = res['complex64'], res['complex128']
    rel = float(np.linalg.norm(c64['g'] - c128['g']) / np.linalg.norm(c128['g']))
    sb = (2 ** n) * 8
    b_alg = (6 * L + 3 + 2 * -(-L // c64['k'])) * sb
    b_alg128 = (6 * L + 3 + 2 * -(-L // c128['k'])) * 2 * sb
    return _line('grad_run circuits/s (McClean 30 qubits, 100 layers, complex64 state + FP64 reductions, '
                 'checkpoints every %d layers)' % c64['k'], 1e3 / c64['ms'], 'circuits/s', 1, 0, c64['ms'],
                 'c64 (FP64 arithmetic and reductions)',
                 {'workload': 'mcclean_grad_run_n30_L100_zzchain', 'qubits': n, 'layers': L, 'seed': SEED,
                  'ckpt_every': c64['k'], 'ckpt_every_tuned': True, 'lds64_mode': c64['lds64'],
                  'tile_shape': c64['shape']},
                 c128_plain={'ms_per_step': c128['ms'], 'ckpt_every': c128['k'], 'tile_shape': c128['shape'],
                             'effective_gbs': b_alg128 / (c128['ms'] * 1e-3) / 1e9, 'b_alg_bytes': b_alg128},
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
                    choices=['mcclean24', 'qaoa512', 'serve', 'mcclean30', 'sample24', 'stream'],
                    help='mcclean24 (default, the headline line); qaoa512: 512 batched QAOA MaxCut grad_runs '
                         '(20 qubits, p=8); serve: Poisson request stream through the serving loop; mcclean30: '
                         'McClean 30 qubits, 100 layers, complex64 checkpoints vs complex128; sample24: '
                         'shot-based parameter-shift gradient, 24 qubits, 8 layers; stream: the streaming-copy '
                         'ceiling')
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
                          'sample24': bench_sample, 'stream': bench_stream}[args.workload](args, sync)))
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
        res[dtype] = dict(ms=dt / steps * 1e3, E=E, roof=roofline(per, sb, 1, N_LAYERS), steps=steps,
                          mode=plan_mode(circ), sbytes=sb)
        grads[dtype] = g
        circ._destroy_plan()
    c64, c128 = res['complex64'], res['complex128']
    rel = float(np.linalg.norm(grads['complex64'] - grads['complex128']) / np.linalg.norm(grads['complex128']))
    rf = c64['roof']
    S = c64['sbytes']
    b_alg = (6 * N_LAYERS + 3) * S
    traffic, tsrc = None, None
    tfile = os.path.join(ROOT, 'profiles', 'traffic_mcclean_n24_L40.json')
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        traffic = tj.get({0: 'c64_forward_pass_bytes', 1: 'c64_reverse_pass_bytes'}.get(rf['kind'], ''))
        tsrc = ('profiles/traffic_mcclean_n24_L40.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this workload, '
                'scaled by a copy-kernel calibration; committed with the kernels it measured, not measured in '
                'this run')
    oracle_err = fixture_errors(grads)
    ceiling = copy_ceiling(S, sync)
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
                   'parameters': N_QUBITS * N_LAYERS, 'entangler': 'cnot', 'seed': SEED,
                   'memory_mode': c64['mode']},
        'c128_plain': {'circuits_per_s': 1e3 / c128['ms'], 'ms_per_step': c128['ms'], 'steps': c128['steps'],
                       'memory_mode': c128['mode'],
                       'roofline_achieved_gbs': c128['roof']['achieved'],
                       'whole_step_gbs': (6 * N_LAYERS + 3) * c128['sbytes'] / (c128['ms'] * 1e-3) / 1e9},
        'speedup_c64_vs_c128': c128['ms'] / c64['ms'],
        'c64_vs_c128_grad_rel_err': rel,
        'grad_rel_err_vs_oracle': oracle_err,
        'E': c64['E'],
        'roofline': {'bound': 'hbm', 'achieved': rf['achieved'], 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': rf['achieved'] / HBM_PEAK_GBS, 'traffic': traffic,
                     'kernel': {0: 'forward pass', 1: 'adjoint two-state pass', 2: 'forward + observable pass'}[
                         rf['kind']],
                     'avg_launch_ms': rf['avg_ms'], 'algorithmic_bytes_per_launch': rf['alg_bytes'],
                     'passes_per_layer': rf['passes_per_layer'],
                     'ms_per_step_by_kind': rf['ms_by_kind'],
                     'whole_step': {'b_alg_bytes': b_alg, 'gbs': b_alg / (c64['ms'] * 1e-3) / 1e9,
                                    'frac': b_alg / (c64['ms'] * 1e-3) / 1e9 / HBM_PEAK_GBS}},
        'hbm': {'traffic_bytes_per_launch': traffic,
                'sustained_gbs': (traffic / (rf['avg_ms'] * 1e-3) / 1e9) if traffic else None,
                'sustained_frac': (traffic / (rf['avg_ms'] * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
                'traffic_source': tsrc},
        'copy_ceiling': ceiling,
    }
    if not args.no_cpu_baseline:
        out['cpu_baseline'] = cpu_baseline(axes, angles, obs)
    print(json.dumps(out))


def plan_mode(circ):
    """'history' or 'checkpoints every k layers' of a McClean plan."""
    from qradient_amd import _native
    k = ctypes.c_int32()
    _native.call('qg_plan_ckpt_every', circ._get_plan(), ctypes.byref(k))
    return 'history' if k.value == 0 else 'checkpoints every %d layers' % k.value


def fixture_errors(grads):
    """Norm-relative gradient error of each dtype against the oracle fixtures of
    this workload, when present: the C restatement of the reference algorithm
    run in x87 extended precision (mcclean_n24_L40_zzchain_ext.npz, the
    arbiter), and the same restatement in FP64 (mcclean_n24_L40_zzchain.npz,
    itself about 1.7e-12 from the extended value at this depth)."""
    out = {}
    for key, name in (('extended', 'mcclean_n24_L40_zzchain_ext.npz'), ('fp64_reference', 'mcclean_n24_L40_zzchain.npz')):
        path = os.path.join(ROOT, 'tests', 'golden', name)
        if os.path.exists(path):
            ref = np.load(path, allow_pickle=False)['grad']
            out[key] = {dt: float(np.linalg.norm(g - ref) / np.linalg.norm(ref)) for dt, g in grads.items()}
    return out or None


def bench_stream(args, sync):
    """Streaming-copy bandwidth (qg_stream_copy): the copy ceiling of this device.
    Sizes from 16 MiB to 4 GiB per buffer, 1 / 2 / 8 workgroups per CU, ordinary
    and non-temporal accesses, buffers ping-ponged (the memory-side cache can
    serve pairs up to its 256 MB)."""
    import torch
    from qradient_amd import _native
    ctx = _native.Context.default(0)
    rows = []
    big = 4 << 30
    a = torch.empty(big // 4, dtype=torch.float32, device='cuda')
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.fill_(2.0)
    sync()
    for gpc, nt in COPY_SHAPES:
        for mib in (16, 64, 128, 256, 1024, 4096):
            nbytes = mib << 20
            ms = ctypes.c_double()
            best = None
            for r in range(max(4, args.steps)):
                src, dst = (a, b) if r % 2 == 0 else (b, a)
                _native.call('qg_stream_copy', ctx.handle, ctypes.c_void_p(dst.data_ptr()),
                             ctypes.c_void_p(src.data_ptr()), nbytes, nt, gpc, 2, ctypes.byref(ms))
                best = ms.value if best is None else min(best, ms.value)
            rows.append({'mib': mib, 'grid_per_cu': gpc, 'nontemporal': nt, 'ms': best,
                         'gbs': 2 * nbytes / (best * 1e-3) / 1e9})
    top = max((r for r in rows if r['mib'] == 1024), key=lambda r: r['gbs'])
    hbm = top['gbs']
    return _line('streaming copy GB/s (read + write, 1 GiB buffers, best launch shape)', hbm, 'GB/s',
                 max(4, args.steps), 0, None, 'u8',
                 {'workload': 'stream_copy', 'grid_per_cu': top['grid_per_cu'], 'nontemporal': top['nontemporal'],
                  'bytes_per_lane': 16},
                 sweep=rows, roofline={'bound': 'hbm', 'achieved': hbm, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                                       'frac': hbm / HBM_PEAK_GBS, 'traffic': None, 'kernel': 'k_stream_copy'})


if __name__ == '__main__':
    main()
