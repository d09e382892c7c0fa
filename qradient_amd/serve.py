"""Continuous-batching grad_run server.

Requests (McClean circuits of any qubit count and depth) are admitted while
device memory and worker slots allow and run concurrently: each slot owns a
library context with its own HIP stream, so kernels of different requests
overlap on the GPU (ctypes releases the GIL during the calls).

Admission is by HBM footprint: the statevector workspace of the request's
plan.  A request waits until its footprint fits the budget left by the other
slots; queued requests that fit may overtake a larger head-of-queue request at
most ``backfill_limit`` times, after which capacity is held for the head (no
starvation).  Each slot keeps the plan (device workspace) of its last shape
resident, so back-to-back requests of one shape reuse their states; the
memory a slot holds counts against the budget until it switches shape.

The memory mode of a request depends only on its shape (full history below
``history_max_bytes``, otherwise checkpoints every ``ckpt_every`` layers), so
its result is bitwise the same whether it runs alone or in a mix.
"""
import collections
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
        self.arrival = self.start = self.finish = None
        self.footprint = None
        self.mode = None
        self.future = Future()


def passes_per_layer(n, kmax=13, cmin=3):
    return 1 if n <= kmax else 1 + -(-(n - kmax) // (kmax - cmin))


class LibraryExecutor:
    """Runs requests through libqradient_amd, one context (stream) per slot."""

    def __init__(self, slots, device=0, dtype='complex64'):
        self.dtype = dtype
        self.ctx = [_native.Context(device) for _ in range(slots)]
        self.plan = [None] * slots     # (shape key, handle)

    def state_bytes(self, n):
        return (2 ** n) * (8 if self.dtype == 'complex64' else 16)

    def release(self, si):
        if self.plan[si] is not None:
            _native.load().qg_plan_destroy(self.plan[si][1])
            self.plan[si] = None

    def run(self, si, req):
        key = (req.n, req.L, req.mode)
        if self.plan[si] is None or self.plan[si][0] != key:
            self.release(si)
            h = ctypes.c_void_p()
            # mode 0 = the history, forced so the numerics depend on the shape only
            _native.call('qg_mcclean_plan', self.ctx[si].handle, req.n, req.L, _native.DTYPES[self.dtype],
                         _native.QG_ENT_CNOT, req.mode if req.mode > 0 else _native.QG_HISTORY, ctypes.byref(h))
            self.plan[si] = (key, h)
        arr, k = req.observable.c_terms()
        E = ctypes.c_double()
        grad = np.empty((req.L, req.n))
        _native.call('qg_mcclean_grad_run', self.plan[si][1], _native.iptr(req.axes), _native.dptr(req.angles), arr,
                     k, None, ctypes.byref(E), _native.dptr(grad))
        return float(E.value), grad

    def close(self):
        for si in range(len(self.plan)):
            self.release(si)


class GradServer:
    def __init__(self, device=0, slots=4, budget_bytes=None, dtype='complex64', history_max_bytes=4 << 30,
                 ckpt_every=8, backfill_limit=8, executor=None):
        if dtype not in _native.DTYPES:
            raise ValueError('dtype must be complex128 or complex64')
        self.dtype = dtype
        self.history_max_bytes = int(history_max_bytes)
        self.ckpt_every = int(ckpt_every)
        self.backfill_limit = int(backfill_limit)
        self.nslots = int(slots)
        self.executor = executor if executor is not None else LibraryExecutor(self.nslots, device, dtype)
        if budget_bytes is None:
            import torch
            free, _ = torch.cuda.mem_get_info(device)
            budget_bytes = int(0.8 * free)
        self.budget = int(budget_bytes)
        self._held = [0] * self.nslots          # bytes a slot's resident plan occupies
        self._shape = [None] * self.nslots      # shape key of that plan
        self._busy = [False] * self.nslots
        self._queue = collections.deque()
        self._cv = threading.Condition()
        self._closed = False
        self._bypassed = 0
        self.completed = []
        self.peak_bytes = 0
        self._workers = []
        self._dispatcher = threading.Thread(target=self._dispatch, daemon=True)
        self._dispatcher.start()

    # ---- sizing ---------------------------------------------------------------
    def _state_bytes(self, n):
        return (2 ** n) * (8 if self.dtype == 'complex64' else 16)

    def mode_of(self, n, L):
        """0 (full history) when it fits history_max_bytes, else the checkpoint period."""
        P = passes_per_layer(n)
        return 0 if (P * L + 6) * self._state_bytes(n) <= self.history_max_bytes else self.ckpt_every

    def footprint(self, n, L):
        """Workspace bytes of the plan for a request shape (the library's layout)."""
        S = self._state_bytes(n)
        mode = self.mode_of(n, L)
        if mode == 0:
            return (passes_per_layer(n) * L + 4) * S
        nck = sum(1 for i in range(L) if (i + 1) % mode == 0 or i == L - 1)
        return (2 + nck + 4) * S

    # ---- public API -------------------------------------------------------------
    def submit(self, req):
        """Queue a request; returns a Future resolving to (E, grad[L, n])."""
        req.arrival = time.perf_counter()
        req.mode = self.mode_of(req.n, req.L)
        req.footprint = self.footprint(req.n, req.L)
        if req.footprint > self.budget:
            req.future.set_exception(MemoryError('request needs {} bytes, the budget is {}'.format(
                req.footprint, self.budget)))
            return req.future
        with self._cv:
            if self._closed:
                raise RuntimeError('server is closed')
            self._queue.append(req)
            self._cv.notify_all()
        return req.future

    def close(self, wait=True):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            self._dispatcher.join()
            for t in list(self._workers):
                t.join()
        self.executor.close()

    def stats(self):
        done = [r for r in self.completed if r.finish is not None]
        if not done:
            return {}
        lat = np.array([r.finish - r.arrival for r in done])
        t0, t1 = min(r.arrival for r in done), max(r.finish for r in done)
        return {'requests': int(lat.size), 'p50_ms': float(np.percentile(lat, 50) * 1e3),
                'p99_ms': float(np.percentile(lat, 99) * 1e3), 'mean_ms': float(lat.mean() * 1e3),
                'circuits_per_s': float(lat.size / (t1 - t0)) if t1 > t0 else float('nan'),
                'peak_bytes': int(self.peak_bytes), 'budget_bytes': self.budget}

    # ---- scheduling (lock held) ---------------------------------------------------
    def _slot_for(self, req):
        """A free slot for req whose admission keeps the held memory within budget, or None."""
        key = (req.n, req.L, req.mode)
        best = None
        for si in range(self.nslots):
            if self._busy[si]:
                continue
            held_others = sum(self._held) - self._held[si]
            if held_others + req.footprint > self.budget:
                continue
            if self._shape[si] == key:
                return si            # reuse the resident plan of this shape
            if best is None or self._held[si] > self._held[best]:
                best = si            # otherwise free the largest idle plan
        return best

    def _pick(self):
        for i, r in enumerate(self._queue):
            si = self._slot_for(r)
            if si is not None:
                self._bypassed = self._bypassed + 1 if i > 0 else 0
                return i, si
            if i == 0 and self._bypassed >= self.backfill_limit:
                return None          # hold capacity for the head of the queue
        return None

    def _dispatch(self):
        while True:
            with self._cv:
                while True:
                    if self._closed and not self._queue:
                        return
                    pick = self._pick() if not all(self._busy) else None
                    if pick is not None:
                        break
                    self._cv.wait()
                i, si = pick
                req = self._queue[i]
                del self._queue[i]
                self._busy[si] = True
                self._shape[si] = (req.n, req.L, req.mode)
                self._held[si] = req.footprint
                self.peak_bytes = max(self.peak_bytes, sum(self._held))
            t = threading.Thread(target=self._run, args=(si, req), daemon=True)
            self._workers = [x for x in self._workers if x.is_alive()] + [t]
            t.start()

    def _run(self, si, req):
        try:
            req.start = time.perf_counter()
            out = self.executor.run(si, req)
            req.finish = time.perf_counter()
            req.future.set_result(out)
        except BaseException as e:
            req.finish = time.perf_counter()
            self.executor.release(si)
            with self._cv:
                self._held[si] = 0
                self._shape[si] = None
            req.future.set_exception(e)
        finally:
            with self._cv:
                self._busy[si] = False
                self.completed.append(req)
                self._cv.notify_all()


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
    ap.add_argument('--slots', type=int, default=8)
    ap.add_argument('--nmin', type=int, default=12)
    ap.add_argument('--nmax', type=int, default=28)
    ap.add_argument('--lmin', type=int, default=4)
    ap.add_argument('--lmax', type=int, default=64)
    ap.add_argument('--seed', type=int, default=0)
    a = ap.parse_args(argv)
    trace = poisson_trace(a.requests, a.rate, a.seed, (a.nmin, a.nmax), (a.lmin, a.lmax))
    st = replay(trace, slots=a.slots)
    st.update({'rate_per_s': a.rate, 'n_range': [a.nmin, a.nmax], 'L_range': [a.lmin, a.lmax],
               'dtype': 'complex64 state, FP64 arithmetic and reductions', 'slots': a.slots})
    print(json.dumps(st))


if __name__ == '__main__':
    main()
