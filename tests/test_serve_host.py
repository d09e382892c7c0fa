"""Server scheduling on the CPU with a fake executor (no GPU): the HBM budget
is never exceeded, large requests are not starved, errors reach the caller."""
import threading
import time

import numpy as np
import pytest

from qradient_amd.serve import GradServer, Request, poisson_trace


class FakeExecutor:
    def __init__(self, delay=0.002):
        self.delay = delay
        self.running = 0
        self.max_running = 0
        self.lock = threading.Lock()

    def run(self, si, req):
        with self.lock:
            self.running += 1
            self.max_running = max(self.max_running, self.running)
        time.sleep(self.delay * (1 + req.n / 10))
        with self.lock:
            self.running -= 1
        if req.tag == 'boom':
            raise ValueError('Invalid axis 7')
        return float(req.n), np.full((req.L, req.n), req.L, dtype=float)

    def release(self, si):
        pass

    def close(self):
        pass


def _req(n, L, tag=None):
    rs = np.random.RandomState(n * 100 + L)
    zz = np.full((n, n), None)
    for i in range(n - 1):
        zz[i, i + 1] = 1.0
    return Request(n, L, (3 * rs.rand(L, n)).astype(int), rs.rand(L, n), {'zz': zz}, tag=tag)


def test_budget_is_respected_and_all_complete():
    ex = FakeExecutor()
    srv = GradServer(slots=4, budget_bytes=3 << 30, executor=ex, history_max_bytes=1 << 30)
    futs = [srv.submit(_req(n, L)) for n, L in [(20, 8), (22, 16), (18, 4), (24, 12), (21, 30), (12, 5)] * 4]
    res = [f.result(timeout=30) for f in futs]
    srv.close()
    assert all(r[1].shape == (f_req_L, f_req_n) for r, (f_req_n, f_req_L) in
               zip(res, [(20, 8), (22, 16), (18, 4), (24, 12), (21, 30), (12, 5)] * 4))
    st = srv.stats()
    assert st['requests'] == 24 and st['peak_bytes'] <= srv.budget
    assert ex.max_running >= 2          # requests did run concurrently


def test_large_request_not_starved():
    ex = FakeExecutor(delay=0.003)
    fp_big = None
    srv = GradServer(slots=4, budget_bytes=2 << 30, executor=ex, history_max_bytes=1 << 30, backfill_limit=3)
    fp_big = srv.footprint(25, 8)
    assert fp_big <= srv.budget
    small = [srv.submit(_req(12, 4)) for _ in range(3)]
    big = srv.submit(_req(25, 8, tag='big'))
    more = [srv.submit(_req(12, 4)) for _ in range(40)]
    big.result(timeout=30)
    order = [r.tag for r in srv.completed]
    assert order.index('big') < 3 + 3 + 5   # overtaken at most backfill_limit times after the head blocked
    for f in small + more:
        f.result(timeout=30)
    srv.close()


def test_errors_propagate_and_oversize_rejected():
    srv = GradServer(slots=2, budget_bytes=1 << 28, executor=FakeExecutor())
    with pytest.raises(ValueError):
        srv.submit(_req(10, 3, tag='boom')).result(timeout=10)
    with pytest.raises(MemoryError):
        srv.submit(_req(27, 40)).result(timeout=10)
    assert srv.submit(_req(10, 3)).result(timeout=10)[0] == 10.0
    srv.close()


def test_mode_is_a_function_of_shape():
    srv = GradServer(slots=1, budget_bytes=1 << 40, executor=FakeExecutor(), history_max_bytes=4 << 30)
    assert srv.mode_of(16, 64) == 0
    assert srv.mode_of(28, 64) == 8
    assert srv.footprint(28, 64) == (2 + 8 + 4) * (2 ** 28) * 8
    srv.close()


def test_poisson_trace_is_seeded():
    a, b = poisson_trace(50, 10.0, 3), poisson_trace(50, 10.0, 3)
    assert [x[:3] for x in a] == [x[:3] for x in b]
    assert all(12 <= x[1] <= 28 and 4 <= x[2] <= 64 for x in a)
    assert all(a[i][0] < a[i + 1][0] for i in range(49))
