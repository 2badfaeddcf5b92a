"""The server's admission policy (ssa_amd_admission_plan), host only.

The server's workers call the same function with each request's plan (its
first-tier cells, the device bytes its buffers hold, and the share of the
chip's resident workgroup slots its kernel shape fills); here it is called
directly with chosen queues and states.
"""
import pytest

ssa = pytest.importorskip("libssa_amd")

GB = 1 << 30


def state(**kw):
    s = dict(max_batch=16, batch_cells=2.5e11, aging_ms=750.0, assist_ms=50.0, device_budget=10 * GB,
             device_in_use=0, occupancy_in_use=0.0, running=0)
    s.update(kw)
    return s


def req(wait_ms=0.0, cells=1e10, device_bytes=GB, occupancy=1.0):
    return (wait_ms, cells, device_bytes, occupancy)


def test_idle_device_takes_cheapest_first_within_cell_budget():
    q = [req(5, 2e11), req(4, 1e10), req(3, 5e10), req(2, 1e11)]
    # cheapest first: 1e10, 5e10, 1e11 (sum 1.6e11); 2e11 would pass 2.5e11
    assert ssa.admission_plan(q, state()) == [1, 2, 3]


def test_aged_request_goes_first():
    # the aged 2e11 first, then the cheapest while the sum stays within 2.5e11
    q = [req(800, 2e11), req(4, 1e10), req(3, 5e10)]
    assert ssa.admission_plan(q, state()) == [0, 1]
    q = [req(800, 1e11), req(4, 1e10), req(3, 5e10)]
    assert ssa.admission_plan(q, state()) == [0, 1, 2]


def test_max_batch():
    q = [req(i, 1e9) for i in range(10)]
    assert len(ssa.admission_plan(q, state(max_batch=4))) == 4


def test_over_budget_request_is_deferred_while_others_run():
    # 6 GiB in use of 10: a request of 5 GiB waits, one of 3 GiB goes
    st = state(device_in_use=6 * GB, running=1, occupancy_in_use=0.0)
    assert ssa.admission_plan([req(100, device_bytes=5 * GB, occupancy=0.5)], st) == []
    assert ssa.admission_plan([req(100, device_bytes=3 * GB, occupancy=0.5)], st) == [0]
    # the batch stops at the first candidate that does not fit
    q = [req(10, 1e9, 2 * GB, 0.1), req(9, 2e9, 2 * GB, 0.1), req(8, 3e9, 2 * GB, 0.1)]
    assert ssa.admission_plan(q, st) == [0, 1]


def test_request_larger_than_the_budget_runs_alone_on_an_idle_device():
    big = req(1, device_bytes=40 * GB)
    assert ssa.admission_plan([big], state()) == [0]
    # a cheaper request goes first; the big one then runs alone, or once aged first
    assert ssa.admission_plan([big, req(0, 1e9)], state()) == [1]
    assert ssa.admission_plan([req(900, device_bytes=40 * GB), req(0, 1e9)], state()) == [0]
    assert ssa.admission_plan([big], state(running=1, occupancy_in_use=0.2)) == []


def test_occupancy_beside_running_batches():
    # a full-chip grid running: another full-chip batch waits until the oldest
    # request reaches the assist limit, then queues behind it
    st = state(running=1, occupancy_in_use=1.0, device_in_use=GB)
    assert ssa.admission_plan([req(10, occupancy=1.0)], st) == []
    assert ssa.admission_plan([req(60, occupancy=1.0)], st) == [0]
    # workgroups that fit the slots the running grid leaves start at once
    st = state(running=1, occupancy_in_use=0.3)
    assert ssa.admission_plan([req(1, occupancy=0.5)], st) == [0]
    assert ssa.admission_plan([req(1, occupancy=0.5), req(0, occupancy=0.4)], st) == []  # 0.3 + 0.9 > 1
    # with nothing running, occupancy never defers
    assert ssa.admission_plan([req(0, occupancy=1.0)], state()) == [0]


def test_empty_queue():
    assert ssa.admission_plan([], state()) == []
