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
arriving while all slots are in use waits for one).  Two schedulers take
every frame that has arrived (up to max_batch, mixed codes and frame
sizes) as one batch whenever a decoder is idle, and decode it with
ldpc_gpu_decode_batch (or the learned decoder):

  schedule()  one batch at a time on a simulated clock driven by measured
              GPU time: a batch completes after its measured wall time
              (frame table upload, launches, synchronisation);
  run_live()  in real time with several decoder contexts, each with its
              own stream, so that batches overlap on the GPU's idle CUs;
              completion is seen by polling the contexts.

A frame's latency is its batch's completion time minus its arrival time.

The DVB-S2/S2X streams' LLRs come from the GPU channel (MODCOD
constellation, AWGN, max-log demapper).  The DVB-T2 streams' blocks are
rotated-QAM cells of encoded frames (workloads.t2_blocks, AWGN) put
through the GPU T2 front end (ldpc_gpu_t2_demap: cell and bit
deinterleaving, soft values), 0.3 dB above the lowest Es/N0 at which 16
such blocks all decode.
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


def in_order(done, sid):
    """Delivery times when each receiver stream hands its frames on in
    arrival order, as the reference's output thread does
    (libdvbrx/src/rxmt.c:571-600): a frame leaves once it and every earlier
    frame of its stream are decoded.  Frames are in arrival order."""
    out = np.empty_like(done)
    last = {}
    for f in range(len(done)):
        s = int(sid[f])
        out[f] = max(done[f], last.get(s, -np.inf))
        last[s] = out[f]
    return out


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


def t2_frames(dec, dv, name, esn0, sel, seed, fec, buf, llr_off, refs, rows_all, k):
    """The soft values of a T2 stream's blocks (positions sel of the trace)
    through the GPU T2 front end, placed in their slots; the reference bits
    of each into refs / rows_all."""
    import workloads

    cod, mod, ft = T2_PLP[name]
    nb = len(sel)
    # at most 1000 blocks per TI-block, as in a T2 frame (the cell
    # interleaver's shifts bound it)
    plp = ldpcgpu.T2Plp(cod, mod, ft, 1, min(255, max(8, -(-nb // 1000))), 0)
    cells, csi, _, bchcw = workloads.t2_blocks(plp, nb, esn0, seed=seed, distinct=64)
    refs[k] = bchcw
    rows_all[sel] = np.arange(nb) % len(bchcw)
    d_cells, d_csi = dv.put(cells), dv.put(csi)
    tmp = dv.empty(nb * fec.nldpc)
    d_err, d_sig = dv.empty(4 * nb), dv.empty(4 * nb)
    ldpcgpu.t2_demap(dec, plp, nb, d_cells, d_csi, tmp, d_err, d_sig)
    dec.synchronize()
    host = dv.get(tmp, (nb, fec.nldpc), np.int8)
    for q, f in enumerate(sel):
        buf[llr_off[f]:llr_off[f] + fec.nldpc] = host[q]


def run(dec, duration=1.0, load_scale=1.0, learned=False, tables=None, seed=5, max_batch=4096, slots=8192,
        lanes=None):
    """Generate the trace on the GPU and decode it.  lanes=None: the
    scheduler on the simulated clock with one decode at a time (each
    batch's measured wall time advances the clock).  lanes=[decoders]:
    live, in real time, with one batch in flight per decoder context
    (run_live).  Returns latency percentiles and throughput."""
    from nms_train import DevBuf, codeword_pool

    arr, kind, sid = arrivals(duration, load_scale, seed, streams=True)
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
        d_it_mix = dv.empty(4 * n)
        d_ok_mix = dv.empty(n)
        mixed, mixed_at = [], [0]
        rng = np.random.default_rng(seed + 1)
        buf = np.zeros(int(N.sum()), np.int8)
        refs = [None] * len(STREAMS)
        rows_all = np.zeros(n, np.int64)
        for k, (pls, short, rate, fec, code, e) in enumerate(info):
            sel = np.nonzero(kind == k)[0]
            if len(sel) == 0:
                continue
            if STREAMS[k][0] in T2_PLP:
                t2_frames(dec, dv, STREAMS[k][0], e, sel, seed + 10 + k, fec, buf, llr_off, refs, rows_all, k)
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
        # the T2 streams take the reference's T2 FEC thread's cap (dvbt2.c:127)
        frames["max_iter"] = [30 if STREAMS[k][0] in T2_PLP else 50 for k in kind]
        decs = lanes if lanes else [dec]
        if learned:
            for dk in decs:
                for k, (pls, short, rate, fec, code, e) in enumerate(info):
                    t = (tables or {}).get(code)
                    if t is None:  # no trained table: the plain arithmetic
                        dk.set_nms(code, np.full((1, ldpcgpu.NMS_DEGS), 16, np.uint8),
                                   np.full((1, ldpcgpu.NMS_DEGS), 16, np.int16))
                    else:
                        dk.set_nms(code, t[0], t[1])
        learned_frames = frames.copy()
        learned_frames["max_iter"] = 30
        # in learned runs, streams whose code has no trained table keep the
        # plain decoder, and so do the T2 streams (the reference's T2 FEC
        # thread has none, and the tables were trained on DVB-S2 frames)
        plain_kind = [STREAMS[k][0] in T2_PLP or (tables or {}).get(info[k][4]) is None
                      for k in range(len(STREAMS))]
        plain_pos = np.array([plain_kind[k] for k in kind])

        def enqueue(dk, pos):
            # trace positions pos (ascending); iteration counts and flags
            # land at those positions.  In learned runs the plain-decoder
            # frames of the range go in a call of their own (the
            # live scheduler's queues never mix them)
            i0, i1 = int(pos[0]), int(pos[-1]) + 1
            assert i1 - i0 == len(pos)
            if not learned:
                dk.decode_batch(frames[i0:i1], d_llr, d_out, d_it + 4 * i0, d_ok + i0)
                return
            for plain in (True, False):
                idx = np.nonzero(plain_pos[i0:i1] == plain)[0] + i0
                if len(idx) == 0:
                    continue
                if len(idx) == i1 - i0:
                    tab, it_p, ok_p = (frames if plain else learned_frames)[i0:i1], d_it + 4 * i0, d_ok + i0
                else:
                    # a mixed range: the subset's results go to scratch and
                    # are moved to their positions after the run
                    tab = (frames if plain else learned_frames)[idx]
                    j = mixed_at[0]
                    if j + len(idx) > n:
                        raise RuntimeError("streaming: mixed-range scratch exhausted")
                    mixed_at[0] += len(idx)
                    mixed.append((idx, j))
                    it_p, ok_p = d_it_mix + 4 * j, d_ok_mix + j
                if plain:
                    dk.decode_batch(tab, d_llr, d_out, it_p, ok_p)
                else:
                    dk.decode_batch_nms(tab, d_llr, d_out, it_p, ok_p, 50)

        def decode(idx):
            t0 = time.perf_counter()
            enqueue(dec, idx)
            dec.synchronize()
            return time.perf_counter() - t0

        if lanes:
            # one queue per launch kind (degree class x frame size, T2 or
            # not), each context launching one kind at a time on its own
            # stream
            kkey = []
            for k, inf in enumerate(info):
                p = ldpcgpu.code_params(inf[4])
                # T2 frames queue apart: their non-converging blocks run
                # to the T2 cap and would hold back S2 frames of the kind
                kkey.append((p["degree_class"], p["n"], STREAMS[k][0] in T2_PLP, learned and plain_kind[k]))
            key = np.array([kkey.index(kkey[k]) for k in kind])
            # frames are renumbered so that each queue's frames (in arrival
            # order) sit at consecutive positions
            perm = np.concatenate([np.nonzero(key == u)[0] for u in np.unique(key)])
            frames[:] = frames[perm]
            learned_frames[:] = learned_frames[perm]
            plain_pos[:] = plain_pos[perm]
            qstart = np.concatenate([[0], np.cumsum([int((key == u).sum()) for u in np.unique(key)])])
            groups = [perm[qstart[g]:qstart[g + 1]] for g in range(len(qstart) - 1)]
            for dk in decs:
                dk.set_concurrent(False)
            # grow every context's scratch to the largest batch first (a
            # device allocation mid-run would wait for all streams): up to
            # max_batch frames of each queue
            for dk in decs:
                for g in range(len(groups)):
                    m = min(max_batch, len(groups[g]))
                    enqueue(dk, np.arange(qstart[g], qstart[g] + m))
                dk.synchronize()
            done, batches = run_live(arr, decs, lambda dk, g, s0, e: enqueue(dk, np.arange(qstart[g] + s0, qstart[g] + e)),
                                     groups, max_batch, slots)
            # results sit at the renumbered positions
            back = np.empty(n, np.int64)
            back[perm] = np.arange(n)
        else:
            # warm the kernels (the warm-up's mixed-range results are
            # dropped: the run decodes those frames again)
            decode(np.arange(min(64, n)))
            mixed.clear()
            mixed_at[0] = 0
            done, batches = schedule(arr, decode, max_batch, slots)
            back = np.arange(n)
        out = np.zeros(int(((KL + 7) // 8).sum()), np.uint8)
        it = np.zeros(n, np.int32)
        dec.download(out, d_out)
        dec.download(it, d_it)
        if mixed:
            it_mix = np.zeros(mixed_at[0], np.int32)
            dec.download(it_mix, d_it_mix)
            for idx, j in mixed:
                it[idx] = it_mix[j:j + len(idx)]
        it = it[back]
    finally:
        dv.free()
        if learned:
            for dk in (lanes or [dec]):
                for k, inf in enumerate(info):
                    dk.set_nms(inf[4], None, None)
        for dk in (lanes or []):
            dk.set_concurrent(True)
    lat = (done - arr) * 1e3
    lat_io = (in_order(done, sid) - arr) * 1e3
    errors = 0
    nl = len(lanes) if lanes else 1
    for f in range(n):
        k = kind[f]
        ref = refs[k][rows_all[f]]
        o = out[out_off[f]:out_off[f] + len(ref)]
        errors += int(not np.array_equal(o, ref))
    errors_bch = bch_errors(dec, info, kind, out, out_off, refs, rows_all)
    busy = sum(b[-1] for b in batches)
    span = max(float(done.max()), duration)
    return {
        "decoder": "learned" if learned else "plain", "frames": int(n), "duration_s": duration,
        "load_scale": load_scale, "device_slots": slots,
        "scheduler": f"live, {nl} contexts in flight" if lanes else "simulated clock, one batch at a time",
        "offered_gbit_s": round(float(KL.sum()) / duration / 1e9, 4),
        "decoded_gbit_s": round(float(KL.sum()) / span / 1e9, 4),
        "capacity_gbit_s": round(float(KL.sum()) / busy * nl / 1e9, 3),
        "gpu_busy": round(busy / span / nl, 4), "batches": len(batches),
        "avg_batch": round(n / len(batches), 1), "avg_iterations": round(float(it.mean()), 3),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                       "max": round(float(lat.max()), 3), "mean": round(float(lat.mean()), 3)},
        "latency_ms_in_order": {"p50": round(float(np.percentile(lat_io, 50)), 3),
                                "p99": round(float(np.percentile(lat_io, 99)), 3),
                                "max": round(float(lat_io.max()), 3)},
        "frames_over_100ms_cap": int((lat_io > 100.0).sum()),
        "latency_ms_by_stream": {
            STREAMS[k][0]: {"p50": round(float(np.percentile(lat[kind == k], 50)), 3),
                            "p99": round(float(np.percentile(lat[kind == k], 99)), 3)}
            for k in range(len(STREAMS)) if (kind == k).any()},
        "frame_errors_ldpc": errors,
        "frame_errors_after_bch": errors_bch,
    }


def bch_errors(dec, info, kind, out, out_off, refs, rows_all):
    """Frames not equal to their BCH codeword after BCH decoding (on the
    GPU, one call per stream kind, after the run: not in the latencies).
    The T2 streams need it: the 256-QAM blocks leave some frames with
    residual LDPC errors (in the reference's chain as here), which BCH
    corrects, as the reference's FEC thread does (dvbt2.c:2390-2395)."""
    errors = 0
    for k, (pls, short, rate, fec, code, e) in enumerate(info):
        sel = np.nonzero(kind == k)[0]
        if len(sel) == 0:
            continue
        kb = fec.kldpc // 8
        cw = np.stack([out[out_off[f]:out_off[f] + kb] for f in sel])
        d_cw, d_res = dec.malloc(cw.nbytes), dec.malloc(4 * len(sel))
        try:
            dec.upload(d_cw, cw)
            ldpcgpu.bch_decode(dec, short, rate, len(sel), d_cw, kb, d_res)
            dec.synchronize()
            res = np.zeros(len(sel), np.int32)
            dec.download(cw, d_cw)
            dec.download(res, d_res)
        finally:
            dec.free(d_cw)
            dec.free(d_res)
        for q, f in enumerate(sel):
            errors += int(res[q] < 0 or not np.array_equal(cw[q], refs[k][rows_all[f]]))
    return errors


def run_live(arr, decs, enqueue, groups=None, max_batch=4096, slots=None, poll_s=20e-6):
    """Real-time scheduler with several decoder contexts.  Frames arrive at
    arr (seconds from the start, sorted) on the wall clock; a frame is
    admitted at its arrival, or when a device slot frees (slots frames
    resident, as schedule()).  groups: lists of trace positions in arrival
    order (frames of one launch kind each; default: one list of all).
    Whenever a context is idle it takes the group whose oldest admitted
    frame arrived first and batches all of that group's admitted frames
    (up to max_batch): enqueue(context, group, s, e) enqueues positions
    s..e-1 of the group without waiting.  Completion is seen by polling
    each context (Decoder.query).  Returns per-frame completion times and
    the batches (group, first, end, seconds from enqueue to completion)."""
    n = len(arr)
    groups = [np.arange(n)] if groups is None else [np.asarray(g) for g in groups]
    done = np.full(n, np.inf)
    batches = []
    busy = [None] * len(decs)  # (group, s, e, enqueue time) per context
    heads = [0] * len(groups)  # first position of each group not yet enqueued
    finished = 0
    t0 = time.perf_counter()

    def admitted(j, now):
        return arr[j] <= now and (slots is None or j < slots or done[j - slots] <= now)

    while finished < n:
        now = time.perf_counter() - t0
        for k, dk in enumerate(decs):
            b = busy[k]
            if b is not None and dk.query():
                now = time.perf_counter() - t0
                done[groups[b[0]][b[1]:b[2]]] = now
                batches.append((b[0], b[1], b[2], now - b[3]))
                finished += b[2] - b[1]
                busy[k] = None
        for k, dk in enumerate(decs):
            if busy[k] is not None:
                continue
            best = None
            for g, idx in enumerate(groups):
                h = heads[g]
                if h < len(idx) and admitted(idx[h], now) and (best is None or arr[idx[h]] < arr[groups[best][heads[best]]]):
                    best = g
            if best is None:
                break
            idx = groups[best]
            s0 = e = heads[best]
            while e < len(idx) and e - s0 < max_batch and admitted(idx[e], now):
                e += 1
            enqueue(dk, best, s0, e)
            busy[k] = (best, s0, e, now)
            heads[best] = e
        if all(b is None for b in busy):
            pend = [arr[idx[heads[g]]] for g, idx in enumerate(groups) if heads[g] < len(idx)]
            wait = (min(pend) if pend else 0.0) - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        else:
            time.sleep(poll_s)
    return done, batches


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
    ap.add_argument("--contexts", type=int, default=0, help="live scheduler with this many contexts (0: simulated clock)")
    a = ap.parse_args()
    dec = ldpcgpu.Decoder(0)
    lanes = [dec] + [ldpcgpu.Decoder(0) for _ in range(a.contexts - 1)] if a.contexts > 0 else None
    res = run(dec, a.duration, a.load, a.learned, learned_tables() if a.learned else None, lanes=lanes)
    for d in lanes or [dec]:
        d.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
