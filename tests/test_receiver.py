"""The receiver loop's host logic (dvb-rx_amd/receiver.py rx_loop) with
simulated decoder contexts: batching per receiver with the flush, bounded
admission, in-order hand-off."""
import time

import numpy as np

from helpers import ldpcgpu  # noqa: F401  (puts the package on the path)
import receiver  # noqa: E402


class FakeCtx:
    """A context whose work takes `dt` seconds of wall time per launch,
    plus `per` seconds for each frame with an odd index (so later frames
    can finish first)."""

    def __init__(self, n, dt, per=0.0):
        self.end = [None] * n
        self.dt, self.per = dt, per
        self.seen = []

    def submit(self, c, frames):
        self.seen.append((c, list(frames)))
        self.end[c] = time.perf_counter() + self.dt + self.per * int((np.asarray(frames) % 2).sum())

    def poll(self, c):
        return self.end[c] is not None and time.perf_counter() >= self.end[c]


def test_every_frame_sent_at_once_and_delivered_in_order():
    arr = np.arange(40) * 5e-3
    sid = np.arange(40) % 4
    fc = FakeCtx(2, 3e-3)
    got = []
    done, dl, sent, launches = receiver.rx_loop(arr, sid, fc.submit, fc.poll, 2, 64, batch_max=1,
                                                on_deliver=lambda s, f: got.append((s, f)))
    assert np.isfinite(dl).all() and (dl >= done - 1e-9).all() and (done - arr >= 3e-3 - 1e-4).all()
    assert (sent - arr < 5e-3).all()  # sent on arrival (within the loop's step)
    for s in range(4):
        mine = [f for r, f in got if r == s]
        assert mine == sorted(mine) and len(mine) == 10
    assert sum(n for _, n, _ in launches) == 40


def test_batches_per_receiver_full_or_flushed():
    # receiver 0: 9 frames 1 ms apart; receiver 1: 3 frames, slow
    arr = np.concatenate([np.arange(9) * 5e-3, [2.5e-3, 12.5e-3, 22.5e-3]])
    sid = np.concatenate([np.zeros(9, int), np.ones(3, int)])
    o = np.argsort(arr, kind="stable")
    arr, sid = arr[o], sid[o]
    fc = FakeCtx(1, 1e-3)
    done, dl, sent, launches = receiver.rx_loop(arr, sid, fc.submit, fc.poll, 1, 64, batch_max=4, flush_ms=100.0)
    # receiver 0 sends batches of 4 when full: frames 1-4 at the 4th's
    # arrival, 5-8 at the 8th's; its 9th and receiver 1's 3 frames wait for
    # the end of input (sent then, before the 100 ms flush)
    r0 = np.nonzero(sid == 0)[0]
    assert np.allclose(sent[r0[:4]], sent[r0[3]]) and sent[r0[3]] - arr[r0[3]] < 5e-3
    assert np.allclose(sent[r0[4:8]], sent[r0[7]])
    assert (sent - arr <= 100e-3).all()
    for _, fr in fc.seen:
        for s in (0, 1):
            assert len([f for f in fr if sid[f] == s]) <= 8


def test_flush_sends_a_waiting_batch():
    # one receiver whose batch never fills: the flush sends it after 10 ms
    arr = np.array([0.0, 1e-3, 60e-3])
    sid = np.zeros(3, int)
    fc = FakeCtx(1, 1e-3)
    done, dl, sent, _ = receiver.rx_loop(arr, sid, fc.submit, fc.poll, 1, 64, batch_max=64, flush_ms=20.0)
    assert 20e-3 - 1e-4 <= sent[0] <= 30e-3 and sent[1] == sent[0]
    # the third frame arrives last: the input has ended, its batch goes
    assert sent[2] - arr[2] < 5e-3


def test_capacity_bounds_each_launch():
    arr = np.zeros(50)
    sid = np.arange(50) % 5
    fc = FakeCtx(2, 2e-3)
    receiver.rx_loop(arr, sid, fc.submit, fc.poll, 2, 12, batch_max=1)
    assert max(len(fr) for _, fr in fc.seen) <= 12
    assert sorted(f for _, fr in fc.seen for f in fr) == list(range(50))


def test_hand_off_waits_for_earlier_frames():
    # frame 1 (odd) takes longer than frame 2 of the same receiver on the
    # other context: frame 2 is done first but handed on after frame 1
    arr = np.array([0.0, 1e-3, 2e-3])
    sid = np.zeros(3, int)
    fc = FakeCtx(3, 1e-3, per=50e-3)
    done, dl, sent, _ = receiver.rx_loop(arr, sid, fc.submit, fc.poll, 3, 1, batch_max=1)
    assert done[2] < done[1]
    assert dl[2] >= done[1] - 1e-9 and dl[0] <= dl[1] <= dl[2]
