"""The library's serving scheduler on the CPU (QG_SERVER_SIMULATE: the real
admission and scheduling code, no device; a request returns E = n and
grad = L): the arena budget is never exceeded, requests run concurrently,
a large request is not starved, errors reach the caller, and the memory mode
and footprint are the library's layout rule for the shape."""
import numpy as np
import pytest

from qradient_amd.serve import GradServer, Request, poisson_trace


def _req(n, L, axes_value=None, tag=None):
    rs = np.random.RandomState(n * 100 + L)
    zz = np.full((n, n), None)
    for i in range(n - 1):
        zz[i, i + 1] = 1.0
    axes = (3 * rs.rand(L, n)).astype(int)
    if axes_value is not None:
        axes[0, 0] = axes_value
    return Request(n, L, axes, rs.rand(L, n), {'zz': zz}, tag=tag)


@pytest.fixture(autouse=True)
def lib():
    from qradient_amd import _native
    import os
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip('libqradient_amd.so not built')
    _native.load()


def test_budget_is_respected_requests_overlap_and_all_complete():
    shapes = [(20, 8), (22, 16), (18, 4), (24, 12), (21, 30), (12, 5)] * 4
    srv = GradServer(slots=4, budget_bytes=3 << 30, history_max_bytes=1 << 30, simulate=True)
    futs = [srv.submit(_req(n, L)) for n, L in shapes]
    res = [f.result(timeout=30) for f in futs]
    st = srv.stats()
    reqs = list(srv.completed)
    srv.close()
    for (n, L), (E, g) in zip(shapes, res):
        assert E == float(n) and g.shape == (L, n) and np.all(g == L)
    assert st['requests'] == 24 and st['peak_bytes'] <= srv.budget
    # at least two requests held the device at the same time
    spans = sorted((r.info['device_start_ms'], r.info['device_end_ms']) for r in reqs)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    assert len({r.info['slot'] for r in reqs}) > 1


def test_large_request_not_starved():
    srv = GradServer(slots=4, budget_bytes=2 << 30, history_max_bytes=1 << 30, backfill_limit=3, simulate=True)
    fp_small, fp_big = srv.footprint(20, 4), srv.footprint(25, 8)
    assert 3 * fp_small + fp_big > srv.budget >= fp_big   # the big one waits while small ones hold memory
    small = [srv.submit(_req(20, 4)) for _ in range(3)]
    big = srv.submit(_req(25, 8, tag='big'))
    more = [srv.submit(_req(20, 4)) for _ in range(40)]
    big.result(timeout=30)
    for f in small + more:
        f.result(timeout=30)
    order = [r.tag for r in srv.completed]
    srv.close()
    # overtaken at most backfill_limit times once it headed the queue
    assert order.index('big') <= 3 + 3 + 4


def test_errors_propagate_and_oversize_rejected():
    srv = GradServer(slots=2, budget_bytes=1 << 28, simulate=True)
    with pytest.raises(ValueError):
        srv.submit(_req(10, 3, axes_value=7)).result(timeout=10)
    with pytest.raises(MemoryError):
        srv.submit(_req(27, 40)).result(timeout=10)
    assert srv.submit(_req(10, 3)).result(timeout=10)[0] == 10.0
    srv.close()


def test_mode_and_footprint_are_the_library_layout():
    srv = GradServer(slots=1, budget_bytes=1 << 40, history_max_bytes=4 << 30, ckpt_every=8, simulate=True)
    S = lambda n: (2 ** n) * 8  # noqa: E731
    assert srv.mode_of(16, 64) == 0
    # 16 qubits: 2 passes per layer of 11 local bits; history of every pass + 4 states
    assert srv.footprint(16, 64) == (2 * 64 + 4) * S(16)
    # 28 qubits, 64 layers: checkpoints every 8 layers (8 of them) + 2 work + 4 states
    assert srv.mode_of(28, 64) == 8
    assert srv.footprint(28, 64) == (2 + 8 + 4) * S(28)
    # whole-circuit kernel: no workspace
    assert srv.footprint(13, 40) == 0
    srv.close()


def test_poisson_trace_is_seeded():
    a, b = poisson_trace(50, 10.0, 3), poisson_trace(50, 10.0, 3)
    assert [x[:3] for x in a] == [x[:3] for x in b]
    assert all(12 <= x[1] <= 28 and 4 <= x[2] <= 64 for x in a)
    assert all(a[i][0] < a[i + 1][0] for i in range(49))


def _order(srv, reqs):
    """Completion order when reqs[0] is running (its simulated run takes about
    1 us per MiB of workspace: tens of ms) while the others are queued."""
    import time
    futs = [srv.submit(reqs[0])]
    t_end = time.time() + 10
    while srv.counters()['running'] < 1 and not futs[0].done():
        assert time.time() < t_end
        time.sleep(1e-4)
    futs += [srv.submit(r) for r in reqs[1:]]
    for f in futs:
        f.result(timeout=30)
    order = [r.tag for r in srv.completed]
    srv.close()
    return order


def test_shortest_first_runs_the_cheapest_queued_request():
    """One slot: while the first request runs the others queue; they start by
    estimated work (6 L + 3) 2^n, not by arrival."""
    srv = GradServer(slots=1, budget_bytes=1 << 40, history_max_bytes=1 << 40, simulate=True, policy='shortest')
    reqs = [_req(26, 16, tag='first'), _req(24, 20, tag='big'), _req(16, 4, tag='small'), _req(20, 6, tag='medium')]
    assert _order(srv, reqs) == ['first', 'small', 'medium', 'big']


def test_shortest_first_ages_an_overtaken_request():
    """A request overtaken backfill_limit times by later arrivals is served next."""
    srv = GradServer(slots=1, budget_bytes=1 << 40, history_max_bytes=1 << 40, backfill_limit=2, simulate=True,
                     policy='shortest')
    reqs = [_req(26, 16, tag='first'), _req(24, 20, tag='big')] + [_req(16, 4, tag='s%d' % k) for k in range(5)]
    assert _order(srv, reqs) == ['first', 's0', 's1', 'big', 's2', 's3', 's4']


def test_fifo_keeps_arrival_order_on_one_slot():
    srv = GradServer(slots=1, budget_bytes=1 << 40, history_max_bytes=1 << 40, simulate=True)
    reqs = [_req(26, 16, tag='first'), _req(24, 20, tag='big'), _req(16, 4, tag='small'), _req(20, 6, tag='medium')]
    assert _order(srv, reqs) == ['first', 'big', 'small', 'medium']
    with pytest.raises(ValueError):
        GradServer(simulate=True, policy='random')


def test_invalid_requests_fail_at_submit_through_the_c_abi():
    """qg_submit validates before queueing (the reference's ValueErrors):
    zz terms must be upper-triangular, angles finite, the qubit count in
    [1, 32]; a valid request right after still runs."""
    import ctypes
    from qradient_amd import _native
    srv = GradServer(slots=1, budget_bytes=1 << 30, simulate=True)
    lib = _native.load()
    n, L = 6, 2
    axes = np.zeros((L, n), dtype=np.int32)
    angles = np.zeros((L, n))
    t = ctypes.c_int64()

    def submit(n_, axes_, angles_, terms):
        arr, k = _native.terms_array(terms)
        return lib.qg_submit(srv._h, n_, L, _native.iptr(axes_), _native.dptr(angles_), arr, k, ctypes.byref(t))

    ok_terms = [(_native.TERM_ZZ, 0, 1, 1.0)]
    assert submit(n, axes, angles, [(_native.TERM_ZZ, 2, 1, 1.0)]) == _native.QG_ERR_INVALID   # lower triangle
    assert submit(n, axes, angles, [(_native.TERM_ZZ, 0, n, 1.0)]) == _native.QG_ERR_INVALID   # j out of range
    assert submit(n, axes, angles, [(_native.TERM_Z, 0, 0, float('nan'))]) == _native.QG_ERR_INVALID
    bad = angles.copy()
    bad[1, 3] = float('inf')
    assert submit(n, axes, bad, ok_terms) == _native.QG_ERR_INVALID
    assert submit(0, axes, angles, ok_terms) == _native.QG_ERR_INVALID
    assert submit(33, np.zeros((L, 33), dtype=np.int32), np.zeros((L, 33)), ok_terms) == _native.QG_ERR_INVALID
    assert srv.submit(_req(n, L)).result(timeout=10)[0] == float(n)
    assert srv.counters()['queued'] == 0
    srv.close()
