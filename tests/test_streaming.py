"""The streaming workload's traffic (dvb-rx_amd/streaming.py): the
receivers' arrival rates and their frames' receiver streams."""
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


def test_arrivals_streams():
    arr, kind = streaming.arrivals(0.5, 1.0, 9)
    arr2, kind2, sid = streaming.arrivals(0.5, 1.0, 9, streams=True)
    assert np.array_equal(arr, arr2) and np.array_equal(kind, kind2)
    assert set(np.unique(sid).tolist()) == set(range(sum(s[1] for s in streaming.STREAMS)))
    for s in np.unique(sid):
        assert len(set(kind[sid == s].tolist())) == 1
