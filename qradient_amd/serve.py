"""Continuous-batching grad_run server (Python face of the library's server).

The scheduling runs in the library (qg_server_create / qg_submit / qg_poll /
qg_wait_any, declared in include/qradient_amd.h), so callers in any language
get the same server: one device arena allocated at creation, worker slots
with their own HIP streams and one re-laid-out plan each (no device
allocation on a shape change once the plans have grown), admission by the
workspace footprint the library's layout rule gives, bounded backfill past a
head-of-queue request that does not fit, and a memory mode fixed by the
request's shape, so a result is bitwise the same alone and in any mix.

``GradServer.submit`` returns a Future resolving to (E, grad[L, n]); a poller
thread collects finished requests from the library.  ``simulate=True`` runs
the library's scheduler without a device (requests return E = n, grad = L):
the admission logic is testable on hosts without a GPU.
"""
import ctypes
import threading
import time
from concurrent.futures import Future

import numpy as np

from . import _native
from .physical_components import Observable


class Request:
    """One McClean grad_run: axes/angles [L, n] and an observable."""

    def __init__(self, n, L, axes, angles, observable, tag=None):
        self.n, self.L = int(n), int(L)
        self.axes = np.ascontiguousarray(axes, dtype=np.int32).reshape(self.L, self.n)
        self.angles = np.ascontiguousarray(angles, dtype=np.float64).reshape(self.L, self.n)
        self.observable = observable if isinstance(observable, Observable) else Observable(observable)
        self.tag = tag
        self.arrival = self.finish = None
        self.info = None      # the library's qg_request_info, as a dict, once finished
        self.future = Future()


def _info_dict(info):
    return {name: getattr(info, name) for name, _ in _native.RequestInfo._fields_}


class GradServer:
    """``budget_bytes``: the device arena (default 80% of the free memory);
    ``history_max_bytes`` / ``ckpt_every``: the shape-to-memory-mode rule;
    ``policy``: 'fifo' admits the oldest queued request that fits, and
    ``backfill_limit`` bounds how often a head-of-queue request that does not
    fit may be overtaken; 'shortest' admits the cheapest request that fits
    (estimated work (6 L + 3) 2^n), and a request overtaken ``backfill_limit``
    times by later arrivals is served next."""

    def __init__(self, device=0, slots=4, budget_bytes=None, dtype='complex64', history_max_bytes=4 << 30,
                 ckpt_every=8, backfill_limit=8, simulate=False, policy='fifo'):
        if dtype not in _native.DTYPES:
            raise ValueError('dtype must be complex128 or complex64')
        if policy not in ('fifo', 'shortest'):
            raise ValueError("policy must be 'fifo' or 'shortest'")
        if budget_bytes is None:
            if simulate:
                budget_bytes = 1 << 40
            else:
                import torch
                free, _ = torch.cuda.mem_get_info(device)
                budget_bytes = int(0.8 * free)
        self.dtype = dtype
        self.budget = int(budget_bytes)
        self.nslots = int(slots)
        h = ctypes.c_void_p()
        _native.call('qg_server_create', int(device), self.nslots, _native.DTYPES[dtype], self.budget,
                     int(history_max_bytes), int(ckpt_every), int(backfill_limit),
                     (_native.QG_SERVER_SIMULATE if simulate else 0) |
                     (_native.QG_SERVER_SHORTEST_FIRST if policy == 'shortest' else 0), ctypes.byref(h))
        self._h = h
        self._lock = threading.Lock()
        self._pending = {}        # ticket -> Request
        self._stop = False
        self.completed = []       # requests in completion order
        self._poller = threading.Thread(target=self._poll_loop, daemon=True)
        self._poller.start()

    # ---- sizing (the library's rule) ---------------------------------------------
    def footprint(self, n, L):
        b = ctypes.c_size_t()
        k = ctypes.c_int32()
        _native.call('qg_server_footprint', self._h, int(n), int(L), ctypes.byref(b), ctypes.byref(k))
        return int(b.value)

    def mode_of(self, n, L):
        """0 (per-pass history) or the checkpoint period of a request shape."""
        b = ctypes.c_size_t()
        k = ctypes.c_int32()
        _native.call('qg_server_footprint', self._h, int(n), int(L), ctypes.byref(b), ctypes.byref(k))
        return int(k.value)

    # ---- requests ----------------------------------------------------------------
    def submit(self, req):
        """Queue a request; returns a Future resolving to (E, grad[L, n])."""
        req.arrival = time.perf_counter()
        arr, k = req.observable.c_terms()
        t = ctypes.c_int64()
        lib = _native.load()
        with self._lock:
            status = lib.qg_submit(self._h, req.n, req.L, _native.iptr(req.axes), _native.dptr(req.angles), arr, k,
                                   ctypes.byref(t))
            if status == _native.QG_OK:
                self._pending[t.value] = req
        if status != _native.QG_OK:
            try:
                _native.check(status)
            except Exception as e:  # noqa: BLE001 - reported through the future
                req.future.set_exception(e)
        return req.future

    def _poll_loop(self):
        lib = _native.load()
        t = ctypes.c_int64()
        info = _native.RequestInfo()
        while True:
            status = lib.qg_wait_any(self._h, 100, ctypes.byref(t))
            if status == _native.QG_ERR_PENDING:
                with self._lock:
                    if self._stop and not self._pending:
                        return
                continue
            with self._lock:
                req = self._pending.pop(t.value)
            E = ctypes.c_double()
            grad = np.empty((req.L, req.n))
            status = lib.qg_poll(self._h, t.value, -1, ctypes.byref(E), _native.dptr(grad), ctypes.byref(info))
            req.info = _info_dict(info)
            req.finish = req.arrival + req.info['latency_ms'] * 1e-3
            self.completed.append(req)
            if status == _native.QG_OK:
                req.future.set_result((float(E.value), grad))
            else:
                try:
                    _native.check(status)
                except Exception as e:  # noqa: BLE001 - reported through the future
                    req.future.set_exception(e)

    def close(self, wait=True):
        if wait:
            with self._lock:
                futs = [r.future for r in self._pending.values()]
            for f in futs:
                f.exception()
        with self._lock:
            self._stop = True
        self._poller.join()
        h, self._h = self._h, None
        if h:
            _native.call('qg_server_destroy', h)

    def counters(self):
        c = _native.ServerCounters()
        _native.call('qg_server_stats', self._h, ctypes.byref(c))
        return {name: int(getattr(c, name)) for name, _ in _native.ServerCounters._fields_}

    def stats(self):
        done = [r for r in self.completed if r.info is not None]
        if not done:
            return {}
        lat = np.array([r.info['latency_ms'] for r in done])
        t0, t1 = min(r.arrival for r in done), max(r.finish for r in done)
        c = self.counters() if self._h else {'arena_peak': None}
        return {'requests': int(lat.size), 'p50_ms': float(np.percentile(lat, 50)),
                'p99_ms': float(np.percentile(lat, 99)), 'mean_ms': float(lat.mean()),
                'circuits_per_s': float(lat.size / (t1 - t0)) if t1 > t0 else float('nan'),
                'peak_bytes': c['arena_peak'], 'budget_bytes': self.budget}


def poisson_trace(num, rate, seed, n_range=(12, 28), L_range=(4, 64)):
    """Seeded request trace: exponential gaps at `rate`/s, n ~ U{n_range}, L ~ U{L_range},
    McClean axes/angles and a random Z_i Z_{i+1} observable."""
    rs = np.random.RandomState(seed)
    t, out = 0.0, []
    for _ in range(num):
        t += rs.exponential(1.0 / rate)
        n = int(rs.randint(n_range[0], n_range[1] + 1))
        L = int(rs.randint(L_range[0], L_range[1] + 1))
        axes = (3 * rs.rand(L, n)).astype(int)
        angles = 2 * np.pi * rs.rand(L, n)
        zz = np.full((n, n), None)
        for i, w in enumerate(rs.standard_normal(n - 1)):
            zz[i, i + 1] = float(w)
        out.append((t, n, L, axes, angles, {'zz': zz}))
    return out


def replay(trace, slots=8, budget_bytes=None, dtype='complex64', device=0, time_scale=1.0, progress=0, **kw):
    """Submit a trace at its arrival times (scaled) and return the server stats.
    ``progress`` > 0 prints a line to stderr every that many submitted requests
    and every that many results (long traces)."""
    import sys
    srv = GradServer(device=device, slots=slots, budget_bytes=budget_bytes, dtype=dtype, **kw)
    t0 = time.perf_counter()
    futs = []
    for k, (ta, n, L, axes, angles, obs) in enumerate(trace):
        wait = t0 + ta * time_scale - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        futs.append(srv.submit(Request(n, L, axes, angles, obs)))
        if progress and (k + 1) % progress == 0:
            print('replay: %d of %d submitted, %.1f s' % (k + 1, len(trace), time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
    for k, f in enumerate(futs):
        f.result()
        if progress and (k + 1) % progress == 0:
            print('replay: %d of %d done, %.1f s' % (k + 1, len(futs), time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
    st = srv.stats()
    srv.close()
    return st


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description='Poisson-arrival grad_run serving benchmark')
    ap.add_argument('--requests', type=int, default=200)
    ap.add_argument('--rate', type=float, default=20.0, help='arrivals per second')
    ap.add_argument('--slots', type=int, default=4)
    ap.add_argument('--nmin', type=int, default=12)
    ap.add_argument('--nmax', type=int, default=28)
    ap.add_argument('--lmin', type=int, default=4)
    ap.add_argument('--lmax', type=int, default=64)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--policy', choices=['fifo', 'shortest'], default='shortest')
    ap.add_argument('--backfill-limit', type=int, default=32)
    a = ap.parse_args(argv)
    trace = poisson_trace(a.requests, a.rate, a.seed, (a.nmin, a.nmax), (a.lmin, a.lmax))
    st = replay(trace, slots=a.slots, policy=a.policy, backfill_limit=a.backfill_limit)
    st.update({'rate_per_s': a.rate, 'n_range': [a.nmin, a.nmax], 'L_range': [a.lmin, a.lmax],
               'dtype': 'complex64 state, FP64 arithmetic and reductions', 'slots': a.slots})
    print(json.dumps(st))


if __name__ == '__main__':
    main()
