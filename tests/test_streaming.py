"""The streaming scheduler's host logic (dvb-rx_amd/streaming.py) with a
stand-in decoder of known duration: latency accounting, batching and
device-slot admission."""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dvb-rx_amd"))

import streaming  # noqa: E402


def test_arrivals_rates():
    arr, kind = streaming.arrivals(2.0, 1.0, 3)
    assert (np.diff(arr) >= 0).all() and arr.min() >= 0 and arr.max() < 2.0
    for k, (_, count, _, _, _, rate, _) in enumerate(streaming.STREAMS):
        got = (kind == k).sum() / 2.0
        want = count * rate
        assert abs(got - want) < 4 * np.sqrt(want * 2.0) / 2.0 + 1, (k, got, want)


def test_schedule_constant_decode():
    arr = np.array([0.0, 0.1, 0.1, 0.25, 1.0])
    calls = []

    def dec(idx):
        calls.append(list(idx))
        return 0.2

    done, batches = streaming.schedule(arr, dec)
    # t=0: frame 0 alone -> done 0.2; then frames 1,2 (arrived by 0.2) ->
    # 0.4; frame 3 (0.25 <= 0.4) -> 0.6; frame 4 waits for its arrival
    assert calls == [[0], [1, 2], [3], [4]]
    assert np.allclose(done, [0.2, 0.4, 0.4, 0.6, 1.2])
    assert all((done - arr) >= 0.2 - 1e-12)


def test_schedule_max_batch_and_slots():
    arr = np.zeros(10)
    done, batches = streaming.schedule(arr, lambda idx: 1.0, max_batch=4)
    assert [b[1] - b[0] for b in batches] == [4, 4, 2]
    assert np.allclose(done, [1] * 4 + [2] * 4 + [3] * 2)
    # 3 device slots: frame j is admitted when frame j-3 completes
    done, batches = streaming.schedule(arr, lambda idx: 1.0, max_batch=8, slots=3)
    assert [b[1] - b[0] for b in batches] == [3, 3, 3, 1]
    assert np.allclose(done, [1] * 3 + [2] * 3 + [3] * 3 + [4])


class _FakeContext:
    """Stand-in decoder context: a batch enqueued at time t completes at
    t + dur on the wall clock."""

    def __init__(self, dur):
        self.dur = dur
        self.until = 0.0
        self.batches = []

    def start(self, i0, i1):
        self.until = time.perf_counter() + self.dur
        self.batches.append((i0, i1))

    def query(self):
        return time.perf_counter() >= self.until


def test_live_scheduler_overlaps_contexts():
    """Two contexts, 200 ms per batch, frames every 80 ms: a frame never
    waits for a busy context while the other is idle, so no latency passes
    one batch time plus one wait for a context (under 0.36 s; one context
    at a time would reach about two batch times), and both contexts
    work."""
    arr = np.arange(8) * 0.08
    ctx = [_FakeContext(0.2), _FakeContext(0.2)]
    done, batches = streaming.run_live(arr, ctx, lambda dk, g, i0, i1: dk.start(i0, i1), max_batch=64)
    lat = done - arr
    assert np.isfinite(done).all()
    assert sum(b[2] - b[1] for b in batches) == 8
    assert (lat >= 0.2 - 1e-3).all() and (lat < 0.36).all(), lat
    assert ctx[0].batches and ctx[1].batches
    seen = sorted(i for c in ctx for b in c.batches for i in range(*b))
    assert seen == list(range(8))


def test_live_scheduler_single_context_batches_and_slots():
    """One context: frames that arrive while it is busy form the next
    batch; with 2 device slots a frame waits until frame j-2 completes."""
    arr = np.zeros(5)
    ctx = [_FakeContext(0.02)]
    done, batches = streaming.run_live(arr, ctx, lambda dk, g, i0, i1: dk.start(i0, i1), max_batch=64, slots=2)
    assert [b for b in ctx[0].batches] == [(0, 2), (2, 4), (4, 5)]
    assert (np.diff(done[[0, 2, 4]]) >= 0.02 - 1e-3).all()


def test_live_scheduler_queues_by_kind():
    """Two queues (even and odd frames), one context, all frames at t=0:
    each batch holds one queue's frames, the queue with the older head
    first, and every frame completes once."""
    arr = np.linspace(0.0, 0.001, 6)
    groups = [np.array([0, 2, 4]), np.array([1, 3, 5])]
    ctx = [_FakeContext(0.01)]
    seen = []
    time.sleep(0.002)
    done, batches = streaming.run_live(arr, ctx, lambda dk, g, s, e: (seen.append((g, s, e)), dk.start(s, e)),
                                       groups, max_batch=64)
    assert np.isfinite(done).all()
    assert [b[0] for b in batches] == [g for g, _, _ in seen]
    assert sorted(i for g, s, e in seen for i in groups[g][s:e]) == list(range(6))
    assert seen[0][0] == 0


def test_in_order_delivery():
    """A frame leaves only after every earlier frame of its own stream;
    other streams do not hold it back."""
    done = np.array([5.0, 1.0, 2.0, 7.0, 3.0])
    sid = np.array([0, 0, 1, 1, 0])
    assert np.allclose(streaming.in_order(done, sid), [5.0, 5.0, 2.0, 7.0, 5.0])


def test_arrivals_streams():
    arr, kind = streaming.arrivals(0.5, 1.0, 9)
    arr2, kind2, sid = streaming.arrivals(0.5, 1.0, 9, streams=True)
    assert np.array_equal(arr, arr2) and np.array_equal(kind, kind2)
    assert set(np.unique(sid).tolist()) == set(range(sum(s[1] for s in streaming.STREAMS)))
    for s in np.unique(sid):
        assert len(set(kind[sid == s].tolist())) == 1
