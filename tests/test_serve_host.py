"""Serving-loop scheduler logic on the CPU through the NULL backend (no GPU,
nothing computed): admission by memory budget, FIFO continuous batching,
request timeouts, cancellation, staging-allocator failure injection and
shutdown."""
import threading
import time

import numpy as np
import pytest

import axhel_amd as ax

N = 4096
MIB = 1 << 20


def req(server, limbs=1, timeout_us=0, op=ax.OP_ADD):
    a = np.zeros(limbs * N, dtype=np.uint64)
    b = np.zeros(limbs * N, dtype=np.uint64)
    r = np.zeros(limbs * N, dtype=np.uint64)
    status, ticket = server.submit(op, N, limbs, r, a, b, moduli=[97] * limbs, timeout_us=timeout_us)
    return status, ticket, (a, b, r)


def test_requests_complete_and_stats():
    s = ax.Server(backend=ax.BACKEND_NULL, null_latency_us=200, memory_budget_bytes=64 * MIB)
    keep = []
    tickets = []
    for i in range(20):
        st, t, bufs = req(s, limbs=1 + i % 3)
        assert st == ax.OK
        tickets.append(t)
        keep.append(bufs)
    for t in tickets:
        assert s.wait(t) == ax.OK
    stt = s.stats()
    assert stt["submitted"] == 20 and stt["completed"] == 20
    assert stt["limbs_completed"] == sum(1 + i % 3 for i in range(20))
    assert stt["p99_us"] >= stt["p50_us"] > 0
    assert stt["bytes_in_use"] == 0
    s.close()


def test_admission_by_memory_budget():
    # each request stages 3 buffers of 1 MiB (op1, op2, result)
    s = ax.Server(backend=ax.BACKEND_NULL, null_latency_us=20000, memory_budget_bytes=7 * MIB, slots=4)
    st, t, _ = req(s, limbs=N * 8 // 8)  # 4096 limbs * 4096 words = 128 MiB per buffer > budget
    assert st == ax.REJECTED and t is None
    keep, tickets = [], []
    for _ in range(5):
        st, t, bufs = req(s, limbs=1)
        assert st == ax.OK
        tickets.append(t)
        keep.append(bufs)
    time.sleep(0.005)
    stt = s.stats()
    assert stt["bytes_in_use"] <= 7 * MIB  # at most two requests admitted at once
    for t in tickets:
        assert s.wait(t) == ax.OK
    stt = s.stats()
    assert stt["peak_bytes_in_use"] <= 7 * MIB
    assert stt["completed"] == 5 and stt["rejected"] == 1
    s.close()


def test_continuous_batching_admits_while_a_batch_runs():
    s = ax.Server(backend=ax.BACKEND_NULL, null_latency_us=150000, slots=2, memory_budget_bytes=64 * MIB)
    t0 = time.perf_counter()
    st, ta, ka = req(s)
    time.sleep(0.05)
    st, tb, kb = req(s)  # arrives while A's batch is on the device
    assert s.wait(ta) == ax.OK
    t_a = time.perf_counter() - t0
    assert s.wait(tb) == ax.OK
    t_b = time.perf_counter() - t0
    # B ran in the second slot, overlapping A: it did not wait for A's end
    assert t_b < t_a + 0.12, (t_a, t_b)
    s.close()


def test_queue_timeout_and_cancel():
    s = ax.Server(backend=ax.BACKEND_NULL, null_latency_us=100000, slots=1, memory_budget_bytes=64 * MIB)
    st, ta, ka = req(s)
    time.sleep(0.01)  # A is running in the only slot
    st, tb, kb = req(s, timeout_us=20000)  # cannot start within 20 ms
    st, tc, kc = req(s)
    assert s.cancel(ta) == ax.REJECTED  # already running
    assert s.cancel(tc) == ax.OK
    assert s.wait(tc) == ax.CANCELLED
    assert s.wait(tb) == ax.TIMEOUT
    assert s.wait(ta) == ax.OK
    stt = s.stats()
    assert stt["timed_out"] == 1 and stt["cancelled"] == 1 and stt["completed"] == 1
    s.close()


def test_wait_timeout_keeps_ticket():
    s = ax.Server(backend=ax.BACKEND_NULL, null_latency_us=80000, slots=1, memory_budget_bytes=64 * MIB)
    st, t, k = req(s)
    assert s.wait(t, timeout_us=1000) == ax.PENDING  # still running
    assert s.wait(t) == ax.OK
    assert s.wait(t) == ax.INVALID_ARGUMENT  # ticket released after completion
    s.close()


def test_two_waiters_on_one_ticket():
    """Exactly one of two threads waiting on the same ticket collects the
    status; the other learns that the ticket was collected."""
    s = ax.Server(backend=ax.BACKEND_NULL, null_latency_us=60000, slots=1, memory_budget_bytes=64 * MIB)
    for _ in range(5):
        st, t, k = req(s)
        assert st == ax.OK
        got = []
        ths = [threading.Thread(target=lambda: got.append(s.wait(t))) for _ in range(2)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        assert sorted(got) == sorted([ax.OK, ax.INVALID_ARGUMENT]), got
    s.close()


def test_staging_allocation_failures_are_reported():
    s = ax.Server(backend=ax.BACKEND_NULL, null_latency_us=100, memory_budget_bytes=256 * MIB, fail_every_alloc=4)
    keep, tickets = [], []
    for _ in range(12):
        st, t, bufs = req(s)
        assert st == ax.OK
        tickets.append(t)
        keep.append(bufs)
    res = [s.wait(t) for t in tickets]
    assert ax.NO_MEMORY in res and ax.OK in res
    assert set(res) <= {ax.OK, ax.NO_MEMORY}
    stt = s.stats()
    assert stt["failed"] == res.count(ax.NO_MEMORY)
    assert stt["bytes_in_use"] == 0
    s.close()


def test_batch_failures_are_reported_and_release_memory():
    """Every second batch fails after its staging was reserved: its requests
    end NO_MEMORY, the others OK, and all staging returns to the pool."""
    s = ax.Server(backend=ax.BACKEND_NULL, null_latency_us=200, slots=1, memory_budget_bytes=256 * MIB,
                  fail_every_batch=2)
    keep, res = [], []
    for _ in range(8):  # one request per batch: submit, then wait
        st, t, bufs = req(s)
        assert st == ax.OK
        keep.append(bufs)
        res.append(s.wait(t))
    assert res == [ax.OK, ax.NO_MEMORY] * 4, res
    stt = s.stats()
    assert stt["failed"] == 4 and stt["completed"] == 4
    assert stt["bytes_in_use"] == 0
    s.close()


def test_invalid_requests_are_refused():
    s = ax.Server(backend=ax.BACKEND_NULL)
    a = np.zeros(N, dtype=np.uint64)
    assert s.submit(ax.OP_ADD, 3000, 1, a, a, a, moduli=[97])[0] == ax.INVALID_ARGUMENT  # not a power of two
    assert s.submit(ax.OP_ADD, N, 1, a, a, None, moduli=[97])[0] == ax.INVALID_ARGUMENT  # missing operand2
    assert s.submit(ax.OP_MUL, N, 1, a, a, a, moduli=[97], input_mod_factor=3)[0] == ax.UNSUPPORTED_FACTOR
    assert s.submit(ax.OP_REDUCE, N, 1, a, a, moduli=[97], input_mod_factor=1, output_mod_factor=1)[0] \
        == ax.UNSUPPORTED_FACTOR
    assert s.submit(ax.OP_NTT_FORWARD, N, 1, a, a)[0] == ax.INVALID_ARGUMENT  # no plans
    s.close()


def test_shutdown_cancels_queued_and_finishes_running():
    s = ax.Server(backend=ax.BACKEND_NULL, null_latency_us=50000, slots=1, memory_budget_bytes=64 * MIB)
    st, ta, ka = req(s)
    time.sleep(0.01)
    st, tb, kb = req(s)
    out = {}

    def waiter():
        out["b"] = s.wait(tb)

    th = threading.Thread(target=waiter)
    th.start()
    time.sleep(0.005)
    ax.lib().axhel_hip_server_destroy(s._h)
    th.join(timeout=5)
    assert out.get("b") == ax.CANCELLED
    s._h = ax._p()


def test_latency_record_is_bounded():
    """Past `latency_samples` completions the percentiles come from a uniform
    sample of that many latencies (the server's memory stays bounded), while
    count, mean and max stay exact: 400 requests through a record of 16."""
    s = ax.Server(backend=ax.BACKEND_NULL, null_latency_us=50, memory_budget_bytes=256 * MIB, latency_samples=16)
    keep, tickets = [], []
    for i in range(400):
        st, t, bufs = req(s)
        assert st == ax.OK
        tickets.append(t)
        keep.append(bufs)
        if len(tickets) == 50:
            for t in tickets:
                assert s.wait(t) == ax.OK
            tickets, keep = [], []
    for t in tickets:
        assert s.wait(t) == ax.OK
    stt = s.stats()
    assert stt["completed"] == 400
    assert 0 < stt["p50_us"] <= stt["p99_us"] <= stt["max_us"]
    assert 0 < stt["mean_us"] <= stt["max_us"]
    s.reset_stats()
    assert s.stats()["completed"] == 0 and s.stats()["max_us"] == 0
    s.close()


def test_staging_pool_stays_within_budget():
    """Cached staging blocks are freed before a new block would take the pool
    past the budget: 1 MiB blocks from small requests, then 2 MiB blocks for
    larger ones, within 8 MiB (without the trimming the pool would hold both
    sets, 12 MiB)."""
    s = ax.Server(backend=ax.BACKEND_NULL, null_latency_us=2000, memory_budget_bytes=8 * MIB, slots=2)
    for limbs in (1, 64, 1, 64):  # 3 buffers of 1 MiB, then of 2 MiB, per request
        keep, tickets = [], []
        for _ in range(2):
            st, t, bufs = req(s, limbs=limbs)
            assert st == ax.OK
            tickets.append(t)
            keep.append(bufs)
        for t in tickets:
            assert s.wait(t) == ax.OK
        stt = s.stats()
        assert stt["bytes_in_use"] == 0
        assert stt["pool_bytes"] <= 8 * MIB and stt["peak_pool_bytes"] <= 8 * MIB, stt
    assert s.stats()["completed"] == 8
    s.close()
