"""The server's admission and batching policy on a modelled device
(tc_server_simulate runs the same scheduler code as the live server, with a
simulated clock: a batch of P points takes launch_s + per_point_s * P, results
are released hold_s after they finish).  CPU only."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def tc(tclib):
    return tclib


def poisson_trace(rate, k, seed=0, lo=2 ** 12, hi=2 ** 22):
    r = np.random.default_rng(seed)
    t = np.cumsum(r.exponential(1.0 / rate, k))
    n = np.round(np.exp(r.uniform(np.log(lo), np.log(hi), k))).astype(np.uint32)
    return t, n


def test_requests_start_in_arrival_order(tc):
    t, n = poisson_trace(20_000, 500, seed=1)
    r = tc.simulate(t, n, n_streams=3, max_batch_clouds=16, launch_s=100e-6, per_point_s=1 / 20e9)
    assert np.all(r["t_start"] >= t)
    assert np.all(np.diff(r["t_start"]) >= 0)
    assert np.all(r["t_done"] > r["t_start"])


def test_batches_respect_limits(tc):
    t, n = poisson_trace(50_000, 800, seed=2)
    r = tc.simulate(t, n, max_batch_clouds=8, max_batch_points=1 << 23, launch_s=100e-6, per_point_s=1 / 20e9)
    b = r["batch_of"]
    for bid in np.unique(b):
        members = b == bid
        assert members.sum() <= 8
        assert n[members].sum() <= 1 << 23
        assert len(np.unique(r["t_start"][members])) == 1  # one launch time per batch


def test_budget_is_never_exceeded_and_blocks_admission(tc):
    t, n = poisson_trace(20_000, 300, seed=3)
    budget = 4 * tc.result_bytes(1 << 22)
    r = tc.simulate(t, n, result_budget_bytes=budget, launch_s=50e-6, per_point_s=1 / 20e9, hold_s=2e-3)
    assert 0 < r["peak_reserved"] <= budget
    free = tc.simulate(t, n, launch_s=50e-6, per_point_s=1 / 20e9, hold_s=2e-3)
    assert free["peak_reserved"] > budget  # without the cap more would have been in flight
    assert np.mean(r["t_start"] - t) > np.mean(free["t_start"] - t)


def test_head_of_line_waits_for_memory(tc):
    # Budget for one result at a time: request i+1 cannot start before request
    # i's result has been released (hold_s after it finished).
    n = np.full(5, 1 << 20, np.uint32)
    t = np.arange(5) * 1e-6
    hold = 1e-3
    r = tc.simulate(t, n, result_budget_bytes=tc.result_bytes(1 << 20), launch_s=1e-4, per_point_s=0.0,
                    hold_s=hold)
    for i in range(4):
        assert r["t_start"][i + 1] >= r["t_done"][i] + hold - 1e-12


def test_continuous_batching_overlaps_batches(tc):
    # A long batch is running on stream 0 when the second request arrives: it
    # goes out at once on stream 1 instead of queueing behind the first.
    t = np.array([0.0, 1e-4])
    n = np.array([1 << 22, 1 << 12], np.uint32)
    r = tc.simulate(t, n, n_streams=2, launch_s=1e-4, per_point_s=1e-9)
    assert r["t_start"][1] == pytest.approx(1e-4)
    assert r["t_start"][1] < r["t_done"][0]
    assert r["t_done"][1] < r["t_done"][0]
    one = tc.simulate(t, n, n_streams=1, launch_s=1e-4, per_point_s=1e-9)
    assert one["t_start"][1] == pytest.approx(one["t_done"][0])


def test_batch_size_follows_load(tc):
    lo_t, lo_n = poisson_trace(200, 200, seed=4)
    hi_t, hi_n = poisson_trace(100_000, 2000, seed=4)
    kw = dict(launch_s=100e-6, per_point_s=1 / 20e9)
    lo = tc.simulate(lo_t, lo_n, **kw)
    hi = tc.simulate(hi_t, hi_n, **kw)
    assert len(np.unique(lo["batch_of"])) == len(lo_t)  # light load: every request goes out alone
    assert len(hi_t) / len(np.unique(hi["batch_of"])) > 4  # heavy load: requests share launches


def test_batching_beats_one_at_a_time_at_equal_load(tc):
    # Same modelled device; the one-at-a-time loop pays the fixed launch cost
    # per request, the server once per batch.
    t, n = poisson_trace(20_000, 2000, seed=5)
    kw = dict(launch_s=150e-6, per_point_s=1 / 25e9)
    served = tc.simulate(t, n, n_streams=2, max_batch_clouds=256, **kw)
    plain = tc.simulate(t, n, n_streams=1, max_batch_clouds=1, **kw)
    lat_s, lat_p = served["t_done"] - t, plain["t_done"] - t
    assert np.percentile(lat_s, 99) < np.percentile(lat_p, 99)
    span_s = served["t_done"].max() - t[0]
    span_p = plain["t_done"].max() - t[0]
    assert span_s <= span_p


def test_requests_that_can_never_run_are_skipped(tc):
    t = np.array([0.0, 1e-6, 2e-6])
    n = np.array([1 << 12, 1 << 25, 1 << 12], np.uint32)
    r = tc.simulate(t, n, max_batch_points=1 << 24, launch_s=1e-5, per_point_s=0.0)
    assert r["t_start"][1] == -1 and r["t_done"][1] == -1
    assert r["t_done"][0] > 0 and r["t_done"][2] > 0


def test_result_bytes_cover_the_layout(tc):
    # Worst case: every array at n entries, nodes at the octree bound, 256 B alignment.
    n = 1000
    b = tc.result_bytes(n, 30)
    up = lambda x: (x + 255) // 256 * 256  # noqa: E731
    want = up(4 * n) + 3 * up(n) + 4 * up(4 * n) + up(64 * tc.octree_node_bound(n, 30))
    assert b == want
    want63 = up(8 * n) + 3 * up(n) + 4 * up(4 * n) + up(64 * tc.octree_node_bound(n, 63))
    assert tc.result_bytes(n, 63) == want63


def test_queue_saturation_rejects_with_busy(tc):
    """A bounded admission queue: arrivals beyond max_queued waiting requests are
    rejected (TC_ERR_BUSY) instead of queueing without bound; the others run in
    order, and with room to spare nothing is rejected."""
    t = np.arange(200) * 1e-7  # a burst far faster than the modelled device
    n = np.full(200, 1 << 18, np.uint32)
    r = tc.simulate(t, n, n_streams=1, max_batch_clouds=4, launch_s=1e-4, per_point_s=1e-9, max_queued=16)
    busy = r["status"] == tc._lib.TC_ERR_BUSY
    ok = r["status"] == 0
    assert busy.sum() > 0 and ok.sum() + busy.sum() == 200
    assert np.all(r["t_start"][busy] < 0) and np.all(r["t_done"][ok] > 0)
    assert np.all(np.diff(r["t_start"][ok]) >= 0)
    # Never more than 16 waiting: every accepted request was admitted while
    # fewer than 16 earlier accepted ones had not yet started.
    started = r["t_start"][ok]
    for i, ta in enumerate(t[ok]):
        waiting = np.sum((t[ok][:i] <= ta) & (started[:i] > ta))
        assert waiting < 16
    free = tc.simulate(t, n, n_streams=1, max_batch_clouds=4, launch_s=1e-4, per_point_s=1e-9, max_queued=1000)
    assert np.all(free["status"] == 0)


def test_cancellation_of_waiting_requests(tc):
    """Requests cancelled while they wait never start; the rest are served as
    without them, and a cancel after the start has no effect."""
    t = np.arange(40) * 1e-6
    n = np.full(40, 1 << 20, np.uint32)
    t_cancel = np.full(40, np.inf)
    t_cancel[10:20] = 1e-4  # still queued behind the first batches then
    t_cancel[0] = 1.0  # long after it started
    r = tc.simulate(t, n, n_streams=1, max_batch_clouds=2, launch_s=1e-4, per_point_s=1e-9, t_cancel=t_cancel)
    cancelled = r["status"] == tc._lib.TC_ERR_CANCELLED
    assert cancelled[10:20].all() and cancelled.sum() == 10
    assert np.all(r["t_start"][cancelled] < 0)
    kept = ~cancelled
    assert np.all(r["t_done"][kept] > 0) and np.all(np.diff(r["t_start"][kept]) >= 0)
    base = tc.simulate(t, n, n_streams=1, max_batch_clouds=2, launch_s=1e-4, per_point_s=1e-9)
    assert r["t_done"][kept][-1] < base["t_done"][-1]  # the cancelled work is not done


def test_allocator_failing_after_k_allocations(tc):
    """Injected result-allocator failures: after k placements the next ones
    fail; a failed placement ends the batch and the request waits (in order),
    so a finite run of failures only delays requests, and a permanent one
    leaves the remaining requests unserved (TC_ERR_NOMEM) instead of hanging."""
    t = np.arange(30) * 1e-6
    n = np.full(30, 1 << 16, np.uint32)
    kw = dict(n_streams=2, max_batch_clouds=4, launch_s=1e-4, per_point_s=1e-9)
    base = tc.simulate(t, n, **kw)
    delayed = tc.simulate(t, n, alloc_fail_after=5, alloc_fail_count=7, **kw)
    assert np.all(delayed["status"] == 0) and np.all(delayed["t_done"] > 0)
    assert np.all(np.diff(delayed["t_start"]) >= 0)
    assert np.array_equal(delayed["t_start"][:5], base["t_start"][:5])
    assert np.all(delayed["t_start"] >= base["t_start"] - 1e-15)
    stuck = tc.simulate(t, n, alloc_fail_after=5, alloc_fail_count=0xFFFFFFFF, check_status=False, **kw)
    assert stuck["result"] == tc._lib.TC_ERR_NOMEM
    assert np.sum(stuck["status"] == 0) == 5 and np.sum(stuck["status"] == tc._lib.TC_ERR_NOMEM) == 25
    assert np.all(stuck["t_start"][5:] < 0)
