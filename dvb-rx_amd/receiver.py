#!/usr/bin/env python3
"""DVB receivers sharing one GPU: the FEC stage of libdvbrx's multi-threaded
receiver (libdvbrx/src/rxmt.c), with the GPU as its LDPC workers.

Sixteen receivers (streams.STREAMS: 8 DVB-S2 8PSK 3/4, 4 DVB-S2X QPSK
11/20, 4 DVB-T2 PLPs) deliver frames at their real rates (Poisson
arrivals).  A frame arrives as what the receiver's PL / OFDM stage hands
to its FEC stage: the frame's received symbols and noise variance
(DVB-S2/S2X), or the FEC block's time-deinterleaved cells and channel
state (DVB-T2).  Those are made before the clock starts (they are the
radio input); everything after arrival runs inside the timed loop:

  batching   each receiver collects its frames in an open batch, sent when
             it holds batch_max frames or when its first frame has waited
             flush_ms (rxmt.c:264-272, :474-480: a batch per MODCOD, sent
             when full or after max_latency_ms, 100 ms by default);
             batch_max = 1 sends every frame at once (dynamic batching);
  admission  each decoder context owns a fixed device region for up to
             `capacity` frames (symbols, LLRs, decoded bits, results);
             sent batches wait, in the order sent, until a context is idle
             and takes as many whole batches as its region holds, so the
             device memory in use is bounded (as rxmt.c's pool of batch
             buffers bounds its own, :222-228);
  GPU work   per context, on its own stream, one launch of mixed codes and
             frame sizes: the frames' symbols / cells copied in, the
             max-log demapper (ldpc_gpu_demap) or the T2 front end
             (ldpc_gpu_t2_demap), one ldpc_gpu_decode_batch of all of them
             (S2 up to 50 iterations, T2 up to 30, the reference's T2 FEC
             cap), BCH (ldpc_gpu_bch_decode), results copied out;
  delivery   a receiver hands its frames on in arrival order: a frame
             leaves once it and every earlier frame of its receiver are
             decoded (rxmt.c:571-600), to the receiver's sink.

Latency is delivery time minus arrival time, front end and decode
included.  The loop runs on the wall clock; completion is seen by polling
the contexts.
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
import streaming  # noqa: E402
from streaming import STREAMS, T2_PLP  # noqa: E402

FLUSH_MS = 100.0  # rxmt.c:127, max_latency_ms


def rx_loop(arr, sid, submit, poll, ncontexts, capacity, batch_max=1, flush_ms=FLUSH_MS, on_deliver=None,
            poll_s=20e-6, t0=None):
    """The receivers' batching, admission and delivery on the wall clock.

    arr: arrival times (s, sorted), sid: each frame's receiver.
    submit(c, frames): enqueue the GPU work of the frames (trace indices)
    on context c, without waiting; poll(c): True once context c's work is
    done (its results then read).  Returns (done, delivered, sent,
    launches): per frame the time its decode was seen done, the time its
    receiver handed it on, the time its batch was sent; and per launch
    (context, frames, seconds)."""
    n = len(arr)
    done = np.full(n, np.inf)
    delivered = np.full(n, np.inf)
    sent = np.full(n, np.inf)
    launches = []
    open_b = {}          # receiver -> [first frame's arrival, frames]
    ready = []           # sent batches, oldest first: lists of frames
    busy = [None] * ncontexts
    order = {}           # receiver -> its frames in arrival order
    for f in range(n):
        order.setdefault(int(sid[f]), []).append(f)
    nxt = {s: 0 for s in order}
    t0 = time.perf_counter() if t0 is None else t0
    i = 0
    finished = 0

    def send(s, now):
        fr = open_b.pop(s)[1]
        sent[fr] = now
        ready.append(fr)

    while finished < n:
        now = time.perf_counter() - t0
        # arrivals into their receiver's open batch; a full batch is sent
        while i < n and arr[i] <= now:
            s = int(sid[i])
            b = open_b.setdefault(s, [arr[i], []])
            b[1].append(i)
            if len(b[1]) >= batch_max:
                send(s, now)
            i += 1
        # batches whose first frame has waited flush_ms; and all of them
        # once the input has ended
        for s in [s for s, b in open_b.items() if i >= n or now - b[0] >= flush_ms * 1e-3]:
            send(s, now)
        # completions, then in-order delivery per receiver
        for c in range(ncontexts):
            if busy[c] is not None and poll(c):
                now = time.perf_counter() - t0
                fr, ts = busy[c]
                done[fr] = now
                launches.append((c, len(fr), now - ts))
                finished += len(fr)
                busy[c] = None
                for s in {int(sid[f]) for f in fr}:
                    lst = order[s]
                    while nxt[s] < len(lst) and done[lst[nxt[s]]] <= now:
                        f = lst[nxt[s]]
                        delivered[f] = now
                        if on_deliver:
                            on_deliver(s, f)
                        nxt[s] += 1
        # idle contexts take whole sent batches, oldest first, up to their
        # capacity
        for c in range(ncontexts):
            if busy[c] is not None or not ready:
                continue
            take = []
            while ready and len(take) + len(ready[0]) <= capacity:
                take += ready.pop(0)
            if not take:  # a batch larger than a region: split it
                take, ready[0] = ready[0][:capacity], ready[0][capacity:]
            submit(c, np.array(take, np.int64))
            busy[c] = (take, time.perf_counter() - t0)
        if all(b is None for b in busy) and not ready:
            nxt_t = arr[i] if i < n else None
            if open_b:
                fl = min(b[0] for b in open_b.values()) + flush_ms * 1e-3
                nxt_t = fl if nxt_t is None else min(nxt_t, fl)
            if nxt_t is not None:
                wait = nxt_t - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
        else:
            time.sleep(poll_s)
    return done, delivered, sent, launches


class Trace:
    """The receivers' input: arrival times, receiver and stream kind of each
    frame, and per stream kind the frames' received symbols (DVB-S2/S2X:
    the MODCOD's symbols through AWGN, made by the GPU channel from pool
    codewords) or cells (DVB-T2: workloads.t2_blocks, each block its own
    TI-block), with the reference BCH codewords."""

    def __init__(self, dec, duration, load_scale, seed):
        import workloads
        from nms_train import DevBuf, codeword_pool

        self.arr, self.kind, self.sid = streaming.arrivals(duration, load_scale, seed, streams=True)
        n = len(self.arr)
        self.n = n
        self.pos = np.zeros(n, np.int64)  # frame -> index within its kind
        self.kinds = []
        for k, (name, count, pls, short, rate, r, esn0) in enumerate(STREAMS):
            sel = np.nonzero(self.kind == k)[0]
            self.pos[sel] = np.arange(len(sel))
            fec = ldpcgpu.fec_lookup(short, rate)
            code = ldpcgpu.code_lookup(short, rate)
            kd = dict(name=name, pls=pls, short=short, rate=rate, fec=fec, code=code, count=len(sel),
                      t2=name in T2_PLP, max_iter=30 if name in T2_PLP else 50)
            if kd["t2"]:
                cod, mod, ft = T2_PLP[name]
                nb = max(1, len(sel))
                # every block its own TI-block (type 0): its cell
                # interleaver shift is 0, so any run of blocks demaps as
                # one call with as many TI-blocks (at most 255, the
                # L1 field's range)
                cells, csi, refs, rows = [], [], [], []
                for c0 in range(0, nb, 255):
                    m = min(255, nb - c0)
                    c, h, _, cw = workloads.t2_blocks(ldpcgpu.T2Plp(cod, mod, ft, 1, m, 0), m, esn0,
                                                      seed=seed + 10 + k + 1000 * c0, distinct=min(64, m))
                    rows.append(np.arange(m) % len(cw) + sum(len(x) for x in refs))
                    cells.append(c)
                    csi.append(h)
                    refs.append(cw)
                kd.update(cod=cod, mod=mod, ft=ft, cells=np.concatenate(cells), csi=np.concatenate(csi),
                          refs=np.concatenate(refs), rows=np.concatenate(rows),
                          ncells=ldpcgpu.t2_geometry(ldpcgpu.T2Plp(cod, mod, ft, 1, 1, 0))[3])
            else:
                e = ldpcgpu.qef_db(pls) + streaming.MARGIN_DB[pls]
                pool = codeword_pool(short, rate, 64, seed + 10 + k)
                rows = np.random.default_rng(seed + 20 + k).integers(0, 64, max(1, len(sel))).astype(np.int32)
                ns = ldpcgpu.demap_symbols(pls)
                dv = DevBuf(dec)
                try:
                    d_pool, d_rows = dv.put(pool), dv.put(rows)
                    d_e = dv.put(np.full(len(rows), e, np.float32))
                    d_sym, d_n0 = dv.empty(len(rows) * ns * 8), dv.empty(len(rows) * 4)
                    rc = ldpcgpu.gpu_lib().ldpc_gpu_channel_symbols(pls, fec.nldpc, d_pool, d_rows, len(rows), d_e,
                                                                    seed * 977 + k, d_sym, d_n0, dec.stream)
                    if rc:
                        raise RuntimeError(f"ldpc_gpu_channel_symbols: error {rc}")
                    dec.synchronize()
                    sym = dv.get(d_sym, (len(rows), ns, 2), np.float32)
                    n0 = dv.get(d_n0, len(rows), np.float32)
                finally:
                    dv.free()
                kd.update(sym=sym, n0=n0, refs=np.packbits(pool[:, :fec.kldpc], axis=1), rows=rows, ns=ns,
                          esn0_db=e)
            self.kinds.append(kd)

    def ref(self, f):
        kd = self.kinds[self.kind[f]]
        return kd["refs"][kd["rows"][self.pos[f]]]


class GpuFec:
    """The GPU work of the loop: per decoder context a fixed device region
    for `capacity` frames, and submit / poll / results on the context's
    stream."""

    def __init__(self, decs, trace, capacity, learned_tables=None):
        self.decs = decs
        self.tr = trace
        self.cap = capacity
        self.mx_sym = max([kd["ns"] * 8 for kd in trace.kinds if not kd["t2"]] +
                          [kd["ncells"] * 12 for kd in trace.kinds if kd["t2"]])
        self.regions = []
        for dk in decs:
            r = dict(sym=dk.malloc(capacity * self.mx_sym), n0=dk.malloc(4 * capacity),
                     llr=dk.malloc(capacity * 64800), out=dk.malloc(capacity * 8100),
                     it=dk.malloc(4 * capacity), ok=dk.malloc(capacity), bch=dk.malloc(4 * capacity),
                     err=dk.malloc(4 * capacity), sig=dk.malloc(4 * capacity))
            self.regions.append(r)
        self.batch = [None] * len(decs)
        self.results = {}

    def close(self):
        for dk, r in zip(self.decs, self.regions):
            for p in r.values():
                dk.free(p)
        self.regions = []

    def layout(self, frames):
        """The frames of a launch grouped by stream kind (each kind's frames
        consecutive, so BCH runs per kind over one stride)."""
        kinds = self.tr.kind[frames]
        order = np.argsort(kinds, kind="stable")
        return frames[order]

    def submit(self, c, frames):
        tr, dk, r = self.tr, self.decs[c], self.regions[c]
        frames = self.layout(frames)
        n = len(frames)
        ft = np.zeros(n, ldpcgpu.FRAME_DTYPE)
        segs = []
        q = 0
        sym_off = 0
        out_off = 0
        for k, kd in enumerate(tr.kinds):
            sel = frames[tr.kind[frames] == k]
            if len(sel) == 0:
                continue
            N, kb = kd["fec"].nldpc, kd["fec"].kldpc // 8
            p = tr.pos[sel]
            if kd["t2"]:
                for c0 in range(0, len(sel), 255):
                    m = min(255, len(sel) - c0)
                    cells = np.ascontiguousarray(kd["cells"][p[c0:c0 + m]])
                    csi = np.ascontiguousarray(kd["csi"][p[c0:c0 + m]])
                    dk.upload(r["sym"] + sym_off, cells)
                    dk.upload(r["sym"] + sym_off + cells.nbytes, csi)
                    ldpcgpu.t2_demap(dk, ldpcgpu.T2Plp(kd["cod"], kd["mod"], kd["ft"], 1, m, 0), m, r["sym"] + sym_off,
                                     r["sym"] + sym_off + cells.nbytes, r["llr"] + q * 64800 + c0 * N,
                                     r["err"] + 4 * (q + c0), r["sig"] + 4 * (q + c0))
                    sym_off += cells.nbytes + csi.nbytes
            else:
                sym = np.ascontiguousarray(kd["sym"][p])
                dk.upload(r["sym"] + sym_off, sym)
                dk.upload(r["n0"] + 4 * q, np.ascontiguousarray(kd["n0"][p]))
                ldpcgpu.demap(dk, kd["pls"], len(sel), r["sym"] + sym_off, r["n0"] + 4 * q, r["llr"] + q * 64800)
                sym_off += sym.nbytes
            for m in range(len(sel)):
                ft[q + m] = (kd["code"], kd["max_iter"], kd["fec"].kldpc, -1, q * 64800 + m * N, out_off + m * kb)
            segs.append((k, q, len(sel), out_off))
            q += len(sel)
            out_off += len(sel) * kb
        dk.decode_batch(ft, r["llr"], r["out"], r["it"], r["ok"])
        for k, q0, cnt, o0 in segs:
            kd = tr.kinds[k]
            ldpcgpu.bch_decode(dk, kd["short"], kd["rate"], cnt, r["out"] + o0, kd["fec"].kldpc // 8, r["bch"] + 4 * q0)
        self.batch[c] = (frames, segs, out_off)

    def poll(self, c):
        if not self.decs[c].query():
            return False
        frames, segs, out_bytes = self.batch[c]
        dk, r = self.decs[c], self.regions[c]
        n = len(frames)
        out = np.zeros(out_bytes, np.uint8)
        it = np.zeros(n, np.int32)
        ok = np.zeros(n, np.uint8)
        bch = np.zeros(n, np.int32)
        dk.download(out, r["out"])
        dk.download(it, r["it"])
        dk.download(ok, r["ok"])
        dk.download(bch, r["bch"])
        for k, q0, cnt, o0 in segs:
            kb = self.tr.kinds[k]["fec"].kldpc // 8
            for m in range(cnt):
                f = int(frames[q0 + m])
                self.results[f] = (out[o0 + m * kb:o0 + (m + 1) * kb].copy(), int(it[q0 + m]), int(ok[q0 + m]),
                                   int(bch[q0 + m]))
        self.batch[c] = None
        return True


def run(decs, duration=0.5, load_scale=1.0, seed=5, capacity=4096, batch_max=1, flush_ms=FLUSH_MS, trace=None):
    """The receivers on the GPU for `duration` seconds of traffic; decs:
    decoder contexts (one batch in flight each).  Returns the latency and
    throughput record; with trace given (a Trace), its frames."""
    tr = trace or Trace(decs[0], duration, load_scale, seed)
    g = GpuFec(decs, tr, capacity)
    sink = {}

    def on_deliver(s, f):
        sink.setdefault(s, []).append(f)

    try:
        for dk in decs:
            dk.set_concurrent(False)
        # warm every context with a full region of each kind (kernel code
        # loaded, the contexts' decoder scratch grown: no device allocation
        # inside the loop)
        for c in range(len(decs)):
            for k in range(len(tr.kinds)):
                sel = np.nonzero(tr.kind == k)[0][:capacity]
                if len(sel):
                    g.submit(c, sel)
                    while not g.poll(c):
                        time.sleep(1e-4)
        g.results.clear()
        t0 = time.perf_counter()
        done, delivered, sent, launches = rx_loop(tr.arr, tr.sid, g.submit, g.poll, len(decs), capacity, batch_max,
                                                  flush_ms, on_deliver, t0=t0)
    finally:
        for dk in decs:
            dk.set_concurrent(True)
        g.close()
    n = tr.n
    # delivery in arrival order per receiver, every frame once
    for s, lst in sink.items():
        assert lst == sorted(lst) and len(set(lst)) == len(lst)
    assert sum(len(v) for v in sink.values()) == n
    errors = sum(int(not np.array_equal(g.results[f][0][:len(tr.ref(f))], tr.ref(f))) for f in range(n))
    bch_fail = sum(int(g.results[f][3] < 0) for f in range(n))
    lat = (delivered - tr.arr) * 1e3
    wait = (sent - tr.arr) * 1e3
    kl = np.array([tr.kinds[k]["fec"].kldpc for k in tr.kind], np.int64)
    busy = sum(x[2] for x in launches)
    span = max(float(delivered.max()), duration)
    it = np.array([g.results[f][1] for f in range(n)])

    def pct(a):
        return {"p50": round(float(np.percentile(a, 50)), 3), "p99": round(float(np.percentile(a, 99)), 3),
                "max": round(float(a.max()), 3)}

    rec = {
        "decoder": "plain", "frames": int(n), "duration_s": duration, "load_scale": load_scale,
        "contexts": len(decs), "capacity_frames_per_context": capacity,
        "batching": f"per receiver, sent at {batch_max} frame(s) or after {flush_ms:g} ms",
        "offered_gbit_s": round(float(kl.sum()) / duration / 1e9, 4),
        "decoded_gbit_s": round(float(kl.sum()) / span / 1e9, 4),
        "gpu_busy": round(busy / span / len(decs), 4), "launches": len(launches),
        "avg_launch_frames": round(n / max(1, len(launches)), 1), "avg_iterations": round(float(it.mean()), 3),
        "latency_ms": pct(lat), "batch_wait_ms": pct(wait),
        "frames_over_flush": int((wait > flush_ms + 1.0).sum()),
        "latency_ms_by_stream": {STREAMS[k][0]: pct(lat[tr.kind == k]) for k in range(len(STREAMS))
                                 if (tr.kind == k).any()},
        "frame_errors_after_bch": errors, "bch_failed": bch_fail,
        "includes": "symbol / cell upload, demapper or T2 front end, LDPC, BCH, results download, in-order hand-off",
    }
    return rec, tr, g.results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--duration", type=float, default=0.5)
    ap.add_argument("--load", type=float, default=1.0)
    ap.add_argument("--contexts", type=int, default=2)
    ap.add_argument("--capacity", type=int, default=4096)
    ap.add_argument("--batch-max", type=int, default=1)
    ap.add_argument("--flush-ms", type=float, default=FLUSH_MS)
    a = ap.parse_args()
    decs = [ldpcgpu.Decoder(0) for _ in range(a.contexts)]
    try:
        rec, _, _ = run(decs, a.duration, a.load, capacity=a.capacity, batch_max=a.batch_max, flush_ms=a.flush_ms)
    finally:
        for d in decs:
            d.close()
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
