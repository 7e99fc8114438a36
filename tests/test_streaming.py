"""The streaming scheduler's host logic (dvb-rx_amd/streaming.py) with a
stand-in decoder of known duration: latency accounting, batching and
device-slot admission."""
import os
import sys

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
