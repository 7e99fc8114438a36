#!/usr/bin/env python3
"""The traffic of the streaming workload: sixteen receivers and their
frame arrivals (receiver.py runs them on the GPU).

  8 x DVB-S2 8PSK 3/4 normal frames at 27.5 Msym/s  (22194 symbols a frame
      with pilots: 1239 frames/s each)
  4 x DVB-S2X QPSK 11/20 normal frames at 52 Msym/s (33282 symbols: 1562/s)
  2 x DVB-T2 PLPs, 64800-bit 3/5 on 256QAM (8100 cells a block, 7.3 Mcells/s
      of data cells: 900 blocks/s)
  2 x DVB-T2 PLPs, 16200-bit 2/3 on 64QAM (2700 cells: 2700 blocks/s)

Each receiver sends Poisson arrivals at its real rate.  The DVB-S2/S2X
frames are sent at their MODCOD's threshold + a margin (S2: QEF + 0.3 dB;
S2X: Shannon + 1 dB + 1.0 dB); the DVB-T2 blocks 0.3 dB above the lowest
Es/N0 at which 16 such blocks all decode.
"""
import numpy as np

# (name, count, pls or -1, short_frame, rate, frames/s per stream, Es/N0 dB)
STREAMS = [
    ("s2_8psk_3_4", 8, 0x38, 0, "3_4", 27.5e6 / 22194, None),
    ("s2x_qpsk_11_20", 4, 0x88, 0, "11_20", 52e6 / 33282, None),
    ("t2_256qam_3_5", 2, -1, 0, "3_5", 900.0, 18.7),
    ("t2_64qam_2_3_short", 2, -1, 1, "2_3", 2700.0, 14.1),
]
MARGIN_DB = {0x38: 0.3, 0x88: 1.0}  # above the MODCOD's threshold (S2: QEF, S2X: Shannon + 1 dB)
# the T2 streams' PLPs: (cod, mod, fec_type), rotated, type-0 time
# interleaver
T2_PLP = {"t2_256qam_3_5": (1, 3, 1), "t2_64qam_2_3_short": (2, 2, 0)}


def arrivals(duration, load_scale, seed, streams=False):
    """Sorted arrival times and stream kinds of all frames of the trace
    (and, with streams=True, each frame's receiver stream, 0..15)."""
    rng = np.random.default_rng(seed)
    t, kind, sid = [], [], []
    for k, (_, count, _, _, _, rate, _) in enumerate(STREAMS):
        for s in range(count):
            lam = rate * load_scale
            n = rng.poisson(lam * duration * 1.2) + 16
            ts = np.cumsum(rng.exponential(1.0 / lam, n))
            ts = ts[ts < duration]
            t.append(ts)
            kind.append(np.full(len(ts), k, np.int32))
            sid.append(np.full(len(ts), len(sid), np.int32))
    t = np.concatenate(t)
    kind = np.concatenate(kind)
    sid = np.concatenate(sid)
    order = np.argsort(t, kind="stable")
    if streams:
        return t[order], kind[order], sid[order]
    return t[order], kind[order]
