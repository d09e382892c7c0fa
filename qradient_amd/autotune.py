"""Autotuner for grad_run plans.

- McClean tile shape: ``kmax`` (local bits per tile, the fusion depth: gates
  fused into one pass), ``cmin`` (contiguous low bits kept in high-bit passes
  for coalescing) and the log2 amplitudes per thread of the complex64
  forward / history-adjoint passes (``r_fwd``, ``r_rev``).
- McClean checkpoint period (``tune_ckpt``): the per-pass history or a
  checkpoint every k layers, for circuits whose history does not fit.
- QAOA batch shape (``tune_qaoa``): tile shape, circuits resident at once
  (``chunk``) and workgroups per CU of a batched launch (``wg_per_cu``).

The tuner times each candidate on the device and caches the fastest per
workload key in a JSON file (``$QRADIENT_AMD_TUNE_CACHE``, default
``~/.cache/qradient_amd/tune.json``).  Lookups fall back to
``tune_defaults.json`` beside this file: shapes measured on MI355X and
committed with the package, so a fresh machine starts from them.
"""
import ctypes
import json
import os
import time

import numpy as np

from . import _native

DEFAULTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tune_defaults.json')

DEFAULT = {'kmax': 13, 'cmin': 3, 'r_fwd': 4, 'r_rev': 4}


def _passes(n, kmax, cmin):
    return 1 if n <= kmax else 1 + -(-(n - kmax) // (kmax - cmin))


def default_shape(n, dtype):
    """The library's untuned shape: the fewest passes per layer 13 local bits
    allow, with the smallest tile (>= 11 bits for complex64, 12 for
    complex128) that keeps that count; 13 bits for complex64 states of 4 GiB
    and more (n >= 29)."""
    if dtype == 'complex64' and n >= 29:
        return dict(DEFAULT, kmax=13)
    best = _passes(n, 13, 3)
    for kmax in range(11 if dtype == 'complex64' else 12, 13):
        if _passes(n, kmax, 3) == best:
            return dict(DEFAULT, kmax=kmax)
    return dict(DEFAULT)


def cache_path():
    return os.environ.get('QRADIENT_AMD_TUNE_CACHE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'qradient_amd', 'tune.json'))


def _read(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _load_cache():
    return _read(cache_path())


def _save_cache(c):
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + '.tmp'
        with open(tmp, 'w') as f:
            json.dump(c, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def key(kind, n, L, dtype, entangler):
    return '{}:n{}:L{}:{}:{}'.format(kind, n, L, dtype, entangler)


def lookup(k):
    """The cached entry for key k, else the committed default, else None."""
    v = _load_cache().get(k)
    return v if v is not None else _read(DEFAULTS_PATH).get(k)


def _store(k, v):
    c = _load_cache()
    c[k] = v
    _save_cache(c)


def cached(kind, n, L, dtype, entangler):
    return lookup(key(kind, n, L, dtype, entangler))


def candidates(dtype):
    out = []
    for kmax in ((14, 13, 12, 11, 10) if dtype == 'complex64' else (13, 12, 11, 10)):
        for cmin in (2, 3, 4, 5, 6):
            rs = (4,)
            for r_fwd in rs:
                for r_rev in rs:
                    out.append({'kmax': kmax, 'cmin': cmin, 'r_fwd': r_fwd, 'r_rev': r_rev})
    return out


def apply(plan, shape):
    _native.call('qg_plan_tune', plan, int(shape['kmax']), int(shape['cmin']), int(shape['r_fwd']),
                 int(shape['r_rev']))


def tune_mcclean(n, L, dtype='complex64', entangler='cnot', shapes=None, reps=2, seed=0, device=0, store=True):
    """Time every candidate shape on a seeded circuit; return (best, [(shape, ms)])."""
    ctx = _native.Context.default(device)
    rs = np.random.RandomState(seed)
    axes = np.ascontiguousarray((3 * rs.rand(L, n)).astype(np.int32))
    angles = np.ascontiguousarray(2 * np.pi * rs.rand(L, n))
    terms = [(_native.TERM_ZZ, i, i + 1, float(w)) for i, w in enumerate(rs.standard_normal(n - 1))]
    arr, k = _native.terms_array(terms)
    results = []
    for shape in (shapes or candidates(dtype)):
        h = ctypes.c_void_p()
        _native.call('qg_mcclean_plan', ctx.handle, n, L, _native.DTYPES[dtype], _native.ENTANGLERS[entangler], 0,
                     ctypes.byref(h))
        try:
            try:
                apply(h, shape)
            except ValueError:
                continue  # shape changes the pass count of this plan's layout
            E = ctypes.c_double()
            grad = np.empty((L, n))
            best = None
            for _ in range(reps + 1):
                _native.call('qg_mcclean_grad_run', h, _native.iptr(axes), _native.dptr(angles), arr, k, None,
                             ctypes.byref(E), _native.dptr(grad))
                ms = ctypes.c_double()
                _native.call('qg_plan_last_ms', h, ctypes.byref(ms))
                best = ms.value if best is None else min(best, ms.value)
            results.append((shape, best))
        finally:
            _native.load().qg_plan_destroy(h)
    if not results:
        return default_shape(n, dtype), results
    best = min(results, key=lambda r: r[1])[0]
    if store:
        _store(key('mcclean', n, L, dtype, entangler), best)
    return best, results


# ---------------------------------------------------------------------------
# checkpoint period of McClean circuits
# ---------------------------------------------------------------------------
def ckpt_key(n, L, dtype, entangler):
    return key('mcclean_ckpt', n, L, dtype, entangler)


def ckpt_candidates(ctx, n, L, dtype):
    """0 (the per-pass history) when it fits the planner's budget, and the
    feasible periods k = kmin, 2 kmin, 4 kmin, ... (kmin: the smallest period
    whose checkpoints fit)."""
    out = []
    b = ctypes.c_size_t()
    k = ctypes.c_int32()
    _native.call('qg_mcclean_workspace_bytes', ctx.handle, n, L, _native.DTYPES[dtype], 0, ctypes.byref(b),
                 ctypes.byref(k))
    if k.value == 0:
        out.append(0)
        return out  # the history fits: nothing to trade
    # kmin, 2 kmin, 4 kmin: longer periods save memory but add uncompute
    # drift (complex64), so they are only taken when measurably faster
    kk = k.value
    for _ in range(3):
        out.append(min(kk, L))
        if kk >= L:
            break
        kk *= 2
    return out


def tune_ckpt(n, L, dtype='complex64', entangler='cnot', seed=0, device=0, store=True, candidates=None):
    """Times one grad_run per checkpoint period; returns (best k, [(k, ms)])."""
    ctx = _native.Context.default(device)
    rs = np.random.RandomState(seed)
    axes = np.ascontiguousarray((3 * rs.rand(L, n)).astype(np.int32))
    angles = np.ascontiguousarray(2 * np.pi * rs.rand(L, n))
    terms = [(_native.TERM_ZZ, i, i + 1, float(w)) for i, w in enumerate(rs.standard_normal(n - 1))]
    arr, nt = _native.terms_array(terms)
    results = []
    for k in (candidates or ckpt_candidates(ctx, n, L, dtype)):
        h = ctypes.c_void_p()
        _native.call('qg_mcclean_plan', ctx.handle, n, L, _native.DTYPES[dtype], _native.ENTANGLERS[entangler],
                     int(k) if k > 0 else _native.QG_HISTORY, ctypes.byref(h))
        try:
            E = ctypes.c_double()
            grad = np.empty((L, n))
            best = None
            for _ in range(2):  # the first run allocates and touches the workspace
                _native.call('qg_mcclean_grad_run', h, _native.iptr(axes), _native.dptr(angles), arr, nt, None,
                             ctypes.byref(E), _native.dptr(grad))
                ms = ctypes.c_double()
                _native.call('qg_plan_last_ms', h, ctypes.byref(ms))
                best = ms.value if best is None else min(best, ms.value)
            results.append((int(k), best))
        finally:
            _native.load().qg_plan_destroy(h)
    # the shortest period within 2% of the fastest (the most accurate)
    fastest = min(ms for _, ms in results)
    best = min(k for k, ms in results if ms <= 1.02 * fastest)
    if store:
        _store(ckpt_key(n, L, dtype, entangler), best)
    return best, results


# ---------------------------------------------------------------------------
# QAOA batch shape
# ---------------------------------------------------------------------------
QAOA_DEFAULT = {'kmax': 0, 'cmin': 3, 'chunk': 0, 'wg_per_cu': 16}


def qaoa_key(n, p, batch, dtype):
    return 'qaoa:n{}:p{}:B{}:{}'.format(n, p, batch, dtype)


def apply_qaoa(plan, shape):
    _native.call('qg_qaoa_tune', plan, int(shape['kmax']), int(shape['cmin']), int(shape['chunk']),
                 int(shape['wg_per_cu']))


def tune_qaoa(n, p, batch, dtype='complex64', seed=0, device=0, store=True, reps=2):
    """Staged search over the batch shape of a QAOA plan on seeded MaxCut
    graphs of the workload's size: the tile shape first (kmax, cmin), then the
    chunk (circuits resident at once), then workgroups per CU; each candidate
    timed by wall clock over whole batched grad runs (host preparation of the
    descriptors overlaps the device and counts).  Returns (best, [(shape, ms)])."""
    from .batch import QaoaBatch
    from .optimization_problems import MaxCut
    rs = np.random.RandomState(seed)
    obs = [MaxCut.regular3(n, rs).to_observable() for _ in range(batch)]
    betas, gammas = np.pi * rs.rand(batch, p), np.pi * rs.rand(batch, p)
    qb = QaoaBatch(n, p, batch, dtype=dtype, device=device)
    packed = qb.pack(obs)
    results = []

    def timed(shape):
        apply_qaoa(qb._plan, shape)
        qb.grad_run(packed, betas, gammas)
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            qb.grad_run(packed, betas, gammas)
            dt = (time.perf_counter() - t0) * 1e3
            best = dt if best is None else min(best, dt)
        results.append((dict(shape), best))
        return best

    best = dict(QAOA_DEFAULT)
    best_ms = timed(best)
    stages = [[{'kmax': k, 'cmin': c} for k in (0, 11, 12, 13) for c in (3, 4) if k == 0 or c < k],
              [{'chunk': ch} for ch in (0, 256, 128, 64, 32) if ch < batch or ch == 0],
              [{'wg_per_cu': w} for w in (8, 16, 32)]]
    for stage in stages:
        cur = dict(best)
        for upd in stage:
            cand = dict(cur, **upd)
            if cand == best:
                continue
            try:
                ms = timed(cand)
            except (ValueError, NotImplementedError, MemoryError):
                continue
            if ms < best_ms:
                best, best_ms = cand, ms
    apply_qaoa(qb._plan, best)
    del qb
    if store:
        _store(qaoa_key(n, p, batch, dtype), best)
    return best, results
