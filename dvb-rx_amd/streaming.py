#!/usr/bin/env python3
"""Streaming request loop on one GPU: frame latency under live traffic.

Sixteen receivers deliver LDPC frames at their real rates (Poisson
arrivals):

  8 x DVB-S2 8PSK 3/4 normal frames at 27.5 Msym/s  (22194 symbols a frame
      with pilots: 1239 frames/s each)
  4 x DVB-S2X QPSK 11/20 normal frames at 52 Msym/s (33282 symbols: 1562/s)
  2 x DVB-T2 PLPs, 64800-bit 3/5 on 256QAM (8100 cells a block, 7.3 Mcells/s
      of data cells: 900 blocks/s)
  2 x DVB-T2 PLPs, 16200-bit 2/3 on 64QAM (2700 cells: 2700 blocks/s)

Every frame's int8 LLRs are written by the GPU channel into a device-
resident slot when it is admitted (slots are the admission limit: a frame
arriving while all slots are in use waits for one).  The scheduler runs
on a simulated clock driven by measured GPU time: whenever the decoder is
idle it takes every frame that has arrived (up to max_batch, mixed codes
and frame sizes) as one batch, decodes it (ldpc_gpu_decode_batch, or the
learned decoder), and the batch completes after its measured wall time
(frame table upload, launches, synchronisation).  A frame's latency is
its batch's completion time minus its arrival time.

DVB-T2 QAM constellations are not among the DVB-S2/S2X MODCODs of the
GPU channel; the T2 streams' LLRs come from its +-1 per bit channel at an
Es/N0 placing them at the same distance above their code's threshold.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import ldpcgpu  # noqa: E402

# (name, count, pls or -1, short_frame, rate, frames/s per stream, Es/N0 dB)
STREAMS = [
    ("s2_8psk_3_4", 8, 0x38, 0, "3_4", 27.5e6 / 22194, None),
    ("s2x_qpsk_11_20", 4, 0x88, 0, "11_20", 52e6 / 33282, None),
    ("t2_256qam_3_5", 2, -1, 0, "3_5", 900.0, 3.0),
    ("t2_64qam_2_3_short", 2, -1, 1, "2_3", 2700.0, 3.6),
]
MARGIN_DB = {0x38: 0.3, 0x88: 1.0}  # above the MODCOD's threshold (S2: QEF, S2X: Shannon + 1 dB)


def arrivals(duration, load_scale, seed):
    """Sorted arrival times and stream kinds of all frames of the trace."""
    rng = np.random.default_rng(seed)
    t, kind = [], []
    for k, (_, count, _, _, _, rate, _) in enumerate(STREAMS):
        for s in range(count):
            lam = rate * load_scale
            n = rng.poisson(lam * duration * 1.2) + 16
            ts = np.cumsum(rng.exponential(1.0 / lam, n))
            ts = ts[ts < duration]
            t.append(ts)
            kind.append(np.full(len(ts), k, np.int32))
    t = np.concatenate(t)
    kind = np.concatenate(kind)
    order = np.argsort(t, kind="stable")
    return t[order], kind[order]


def schedule(arr, decode, max_batch=4096, slots=None):
    """Greedy batch scheduler on a simulated clock.  decode(idx) decodes
    the frames idx (positions in arr, in arrival order) and returns its
    duration in seconds.  A frame is admitted at its arrival, or when a
    device slot frees if `slots` frames are resident (frames complete in
    order, so frame j gets the slot of frame j - slots).  Whenever the
    decoder is idle it takes every admitted frame (up to max_batch).
    Returns per-frame completion times and the batches (first, end,
    seconds)."""
    n = len(arr)
    done = np.full(n, np.inf)
    batches = []

    def admit(j):
        return max(arr[j], done[j - slots]) if slots is not None and j >= slots else arr[j]

    t = 0.0
    i = 0
    while i < n:
        t = max(t, admit(i))
        j = i
        while j < n and j - i < max_batch and admit(j) <= t:
            j += 1
        d = decode(np.arange(i, j))
        t += d
        done[i:j] = t
        batches.append((i, j, d))
        i = j
    return done, batches


def run(dec, duration=1.0, load_scale=1.0, learned=False, tables=None, seed=5, max_batch=4096, slots=8192):
    """Generate the trace on the GPU and run the scheduler with real
    decodes.  Returns latency percentiles and throughput."""
    from nms_train import DevBuf, codeword_pool

    arr, kind = arrivals(duration, load_scale, seed)
    n = len(arr)
    info = []
    for name, count, pls, short, rate, r, esn0 in STREAMS:
        fec = ldpcgpu.fec_lookup(short, rate)
        code = ldpcgpu.code_lookup(short, rate)
        e = ldpcgpu.qef_db(pls) + MARGIN_DB[pls] if pls >= 0 else esn0
        info.append((pls, short, rate, fec, code, e))
    N = np.array([info[k][3].nldpc for k in kind], np.int64)
    KL = np.array([info[k][3].kldpc for k in kind], np.int64)
    llr_off = np.concatenate([[0], np.cumsum(N)[:-1]])
    out_off = np.concatenate([[0], np.cumsum((KL + 7) // 8)[:-1]])
    dv = DevBuf(dec)
    try:
        d_llr = dv.empty(int(N.sum()))
        d_out = dv.empty(int(((KL + 7) // 8).sum()))
        d_it = dv.empty(4 * n)
        d_ok = dv.empty(n)
        rng = np.random.default_rng(seed + 1)
        buf = np.zeros(int(N.sum()), np.int8)
        refs = [None] * len(STREAMS)
        rows_all = np.zeros(n, np.int64)
        for k, (pls, short, rate, fec, code, e) in enumerate(info):
            sel = np.nonzero(kind == k)[0]
            if len(sel) == 0:
                continue
            pool = codeword_pool(short, rate, 64, seed + 10 + k)
            refs[k] = np.packbits(pool[:, :fec.kldpc], axis=1)
            rows = rng.integers(0, 64, len(sel)).astype(np.int32)
            rows_all[sel] = rows
            d_pool = dv.put(pool)
            # the frames of one stream kind are generated together, then
            # placed in their slots in arrival order
            d_rows = dv.put(rows)
            d_e = dv.put(np.full(len(sel), e, np.float32))
            tmp = dv.empty(len(sel) * fec.nldpc)
            ldpcgpu.channel(pls, fec.nldpc, d_pool, d_rows, len(sel), d_e, seed * 977 + k, tmp, stream=dec.stream)
            dec.synchronize()
            host = dv.get(tmp, (len(sel), fec.nldpc), np.int8)
            for q, f in enumerate(sel):
                buf[llr_off[f]:llr_off[f] + fec.nldpc] = host[q]
        dec.upload(d_llr, buf)
        frames = np.zeros(n, ldpcgpu.FRAME_DTYPE)
        frames["code"] = [info[k][4] for k in kind]
        frames["nbits"] = KL
        frames["group"] = -1
        frames["llr_off"] = llr_off
        frames["out_off"] = out_off
        frames["max_iter"] = 50
        if learned:
            for k, (pls, short, rate, fec, code, e) in enumerate(info):
                t = (tables or {}).get(code)
                if t is None:  # no trained table: the plain arithmetic
                    dec.set_nms(code, np.full((1, ldpcgpu.NMS_DEGS), 16, np.uint8),
                                np.full((1, ldpcgpu.NMS_DEGS), 16, np.int16))
                else:
                    dec.set_nms(code, t[0], t[1])
        learned_frames = frames.copy()
        learned_frames["max_iter"] = 30

        def decode(idx):
            # idx: consecutive trace positions; results at those positions
            i0 = int(idx[0])
            t0 = time.perf_counter()
            if learned:
                dec.decode_batch_nms(learned_frames[idx], d_llr, d_out, d_it + 4 * i0, d_ok + i0, 50)
            else:
                dec.decode_batch(frames[idx], d_llr, d_out, d_it + 4 * i0, d_ok + i0)
            dec.synchronize()
            return time.perf_counter() - t0

        # warm the kernels
        decode(np.arange(min(64, n)))
        done, batches = schedule(arr, decode, max_batch, slots)
        out = np.zeros(int(((KL + 7) // 8).sum()), np.uint8)
        it = np.zeros(n, np.int32)
        dec.download(out, d_out)
        dec.download(it, d_it)
    finally:
        dv.free()
        if learned:
            for k, inf in enumerate(info):
                dec.set_nms(inf[4], None, None)
    lat = (done - arr) * 1e3
    errors = 0
    for f in range(n):
        k = kind[f]
        ref = refs[k][rows_all[f]]
        o = out[out_off[f]:out_off[f] + len(ref)]
        errors += int(not np.array_equal(o, ref))
    busy = sum(b[2] for b in batches)
    span = max(done[-1], duration)
    return {
        "decoder": "learned" if learned else "plain", "frames": int(n), "duration_s": duration,
        "load_scale": load_scale, "device_slots": slots, "offered_gbit_s": round(float(KL.sum()) / duration / 1e9, 4),
        "decoded_gbit_s": round(float(KL.sum()) / span / 1e9, 4),
        "capacity_gbit_s": round(float(KL.sum()) / busy / 1e9, 3),
        "gpu_busy": round(busy / span, 4), "batches": len(batches),
        "avg_batch": round(n / len(batches), 1), "avg_iterations": round(float(it.mean()), 3),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                       "max": round(float(lat.max()), 3), "mean": round(float(lat.mean()), 3)},
        "frame_errors_ldpc": errors,
    }


def learned_tables():
    """Trained tables shipped with the package, by code id."""
    out = {}
    d = os.path.join(HERE, "nms")
    for fn in sorted(os.listdir(d)):
        if fn.startswith("eval_") or not fn.endswith(".json"):
            continue
        t = json.load(open(os.path.join(d, fn)))
        out[ldpcgpu.code_lookup(t["short_frame"], t["rate"])] = (np.array(t["S"], np.uint8),
                                                                 np.array(t["B"], np.int16))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--duration", type=float, default=1.0)
    ap.add_argument("--load", type=float, default=1.0)
    ap.add_argument("--learned", action="store_true")
    a = ap.parse_args()
    dec = ldpcgpu.Decoder(0)
    res = run(dec, a.duration, a.load, a.learned, learned_tables() if a.learned else None)
    dec.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
