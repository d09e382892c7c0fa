"""Autotuner for the pass tile shape of grad_run plans.

A plan's passes are shaped by ``kmax`` (local bits per tile, the fusion
depth: gates fused into one pass), ``cmin`` (contiguous low bits kept in
high-bit passes for coalescing) and the log2 amplitudes per thread of the
complex64 forward / history-adjoint passes (``r_fwd``, ``r_rev``).  The tuner
times one grad_run per candidate on the device (hipEvents around the launch
sequence, no host time) and caches the fastest shape per
(kind, n, L, dtype, entangler, device name) in a JSON file
(``$QRADIENT_AMD_TUNE_CACHE``, default ``~/.cache/qradient_amd/tune.json``).
"""
import ctypes
import json
import os

import numpy as np

from . import _native

DEFAULT = {'kmax': 13, 'cmin': 3, 'r_fwd': 4, 'r_rev': 4}


def _passes(n, kmax, cmin):
    return 1 if n <= kmax else 1 + -(-(n - kmax) // (kmax - cmin))


def default_shape(n, dtype):
    """The library's untuned shape: the fewest passes per layer 13 local bits
    allow, with the smallest tile (>= 11 bits for complex64, 12 for
    complex128) that keeps that count."""
    best = _passes(n, 13, 3)
    for kmax in range(11 if dtype == 'complex64' else 12, 13):
        if _passes(n, kmax, 3) == best:
            return dict(DEFAULT, kmax=kmax)
    return dict(DEFAULT)


def cache_path():
    return os.environ.get('QRADIENT_AMD_TUNE_CACHE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'qradient_amd', 'tune.json'))


def _load_cache():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


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


def cached(kind, n, L, dtype, entangler):
    return _load_cache().get(key(kind, n, L, dtype, entangler))


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
        c = _load_cache()
        c[key('mcclean', n, L, dtype, entangler)] = best
        _save_cache(c)
    return best, results
