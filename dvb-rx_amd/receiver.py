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
             frame sizes, enqueued without waiting: the frames' symbols /
             cells copied in from the receivers' page-locked input buffers
             (runs of consecutive frames, no gather on the host), the
             max-log demapper (ldpc_gpu_demap) or the T2 front end
             (ldpc_gpu_t2_demap), one ldpc_gpu_decode_batch of all of them
             (S2 up to 50 iterations, T2 up to 30, the reference's T2 FEC
             cap), BCH (ldpc_gpu_bch_decode), results copied out to the
             context's page-locked result buffers;
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
LANES = 4  # forked streams per context: one per launch of a batch of the four stream kinds


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

    def pin(self, dec):
        """Move the received symbols / cells (the receivers' input buffers)
        into page-locked host memory, so that their copies to the device
        are enqueued without waiting and run at the link's rate."""
        self.pinned = getattr(self, "pinned", [])
        if self.pinned:
            return
        for kd in self.kinds:
            for key in ("cells", "csi") if kd["t2"] else ("sym", "n0"):
                pa = dec.host_array(kd[key].shape, kd[key].dtype)
                pa[...] = kd[key]
                kd[key] = pa
                self.pinned.append((kd, key, pa))

    def release(self, dec):
        """The input back in ordinary memory, the page-locked buffers freed
        (no copy from them may be in flight)."""
        for kd, key, pa in getattr(self, "pinned", []):
            kd[key] = np.array(pa)
            dec.host_free(pa)
        self.pinned = []


class GpuFec:
    """The GPU work of the loop: per decoder context a fixed device region
    for `capacity` frames and page-locked host buffers for its results;
    submit enqueues everything on the context's stream and returns, poll
    collects a finished launch's results."""

    def __init__(self, decs, trace, capacity):
        self.decs = decs
        self.tr = trace
        self.cap = capacity
        self.mx_sym = max([kd["ns"] * 8 for kd in trace.kinds if not kd["t2"]] +
                          [kd["ncells"] * 12 for kd in trace.kinds if kd["t2"]])
        self.regions, self.host = [], []
        for dk in decs:
            r = dict(sym=dk.malloc(capacity * self.mx_sym), n0=dk.malloc(4 * capacity),
                     llr=dk.malloc(capacity * 64800), out=dk.malloc(capacity * 8100),
                     it=dk.malloc(4 * capacity), ok=dk.malloc(capacity), bch=dk.malloc(4 * capacity),
                     err=dk.malloc(4 * capacity), sig=dk.malloc(4 * capacity))
            self.regions.append(r)
            self.host.append(dict(out=dk.host_array(capacity * 8100, np.uint8), it=dk.host_array(capacity, np.int32),
                                  ok=dk.host_array(capacity, np.uint8), bch=dk.host_array(capacity, np.int32)))
        self.batch = [None] * len(decs)
        self.finished = []  # per completed launch: frames, segments, bits, iterations, flags, BCH results

    def close(self):
        # nothing in flight may still read or write the buffers (an error
        # can end the loop with launches pending)
        for dk in self.decs:
            try:
                dk.synchronize()
            except RuntimeError:
                pass
        for dk, r, h in zip(self.decs, self.regions, self.host):
            for p in r.values():
                dk.free(p)
            for a in h.values():
                dk.host_free(a)
        self.regions, self.host = [], []

    def submit(self, c, frames):
        """Enqueue the launch of `frames` on context c: per stream kind the
        frames' symbols (or cells and channel state) copied in, by runs of
        consecutive frames of the trace's pinned buffers, the demapper (or
        T2 front end), then one decode of all, BCH per kind and the
        results copied out to the context's pinned buffers."""
        tr, dk, r, h = self.tr, self.decs[c], self.regions[c], self.host[c]
        kinds, pos = tr.kind[frames], tr.pos[frames]
        o = np.lexsort((pos, kinds))
        frames, kinds, pos = frames[o], kinds[o], pos[o]
        n = len(frames)
        ft = np.zeros(n, ldpcgpu.FRAME_DTYPE)
        segs = []
        q = sym_off = out_off = 0
        for k in np.unique(kinds):
            kd = tr.kinds[k]
            a, b = np.searchsorted(kinds, [k, k + 1])
            p, m = pos[a:b], int(b - a)
            cut = np.flatnonzero(np.diff(p) != 1) + 1
            runs = list(zip(np.r_[0, cut], np.r_[cut, m]))
            N, kb = kd["fec"].nldpc, kd["fec"].kldpc // 8
            d_llr = r["llr"] + q * 64800
            if kd["t2"]:
                cells, csi = kd["cells"], kd["csi"]
                cb, hb = cells[0].nbytes, csi[0].nbytes
                d_cells = r["sym"] + sym_off
                d_csi = d_cells + m * cb
                for s0, e0 in runs:
                    dk.upload_async(d_cells + int(s0) * cb, cells.ctypes.data + int(p[s0]) * cb, int(e0 - s0) * cb)
                    dk.upload_async(d_csi + int(s0) * hb, csi.ctypes.data + int(p[s0]) * hb, int(e0 - s0) * hb)
                for c0 in range(0, m, 255):
                    mm = min(255, m - c0)
                    ldpcgpu.t2_demap(dk, ldpcgpu.T2Plp(kd["cod"], kd["mod"], kd["ft"], 1, mm, 0), mm,
                                     d_cells + c0 * cb, d_csi + c0 * hb, d_llr + c0 * N,
                                     r["err"] + 4 * (q + c0), r["sig"] + 4 * (q + c0))
                sym_off += m * (cb + hb)
            else:
                sym, n0 = kd["sym"], kd["n0"]
                sb = sym[0].nbytes
                for s0, e0 in runs:
                    dk.upload_async(r["sym"] + sym_off + int(s0) * sb, sym.ctypes.data + int(p[s0]) * sb,
                                    int(e0 - s0) * sb)
                    dk.upload_async(r["n0"] + 4 * (q + int(s0)), n0.ctypes.data + 4 * int(p[s0]), 4 * int(e0 - s0))
                ldpcgpu.demap(dk, kd["pls"], m, r["sym"] + sym_off, r["n0"] + 4 * q, d_llr)
                sym_off += m * sb
            blk = ft[q:q + m]
            blk["code"], blk["max_iter"], blk["nbits"], blk["group"] = kd["code"], kd["max_iter"], kd["fec"].kldpc, -1
            blk["llr_off"] = q * 64800 + np.arange(m, dtype=np.uint64) * N
            blk["out_off"] = out_off + np.arange(m, dtype=np.uint64) * kb
            segs.append((int(k), q, m, out_off))
            q += m
            out_off += m * kb
        dk.decode_batch(ft, r["llr"], r["out"], r["it"], r["ok"])
        for k, q0, cnt, o0 in segs:
            kd = tr.kinds[k]
            ldpcgpu.bch_decode(dk, kd["short"], kd["rate"], cnt, r["out"] + o0, kd["fec"].kldpc // 8, r["bch"] + 4 * q0)
        for name, nb in (("out", out_off), ("it", 4 * n), ("ok", n), ("bch", 4 * n)):
            dk.download_async(h[name].ctypes.data, r[name], nb)
        self.batch[c] = (frames, segs, out_off)

    def poll(self, c):
        if not self.decs[c].query():
            return False
        frames, segs, out_bytes = self.batch[c]
        h, n = self.host[c], len(frames)
        self.finished.append((frames, segs, h["out"][:out_bytes].copy(), h["it"][:n].copy(), h["ok"][:n].copy(),
                              h["bch"][:n].copy()))
        self.batch[c] = None
        return True

    def results(self):
        """{frame: (decoded kldpc bits packed, iterations, syndrome flag,
        BCH result)} of the completed launches."""
        res = {}
        for frames, segs, out, it, ok, bch in self.finished:
            for k, q0, cnt, o0 in segs:
                kb = self.tr.kinds[k]["fec"].kldpc // 8
                for m in range(cnt):
                    res[int(frames[q0 + m])] = (out[o0 + m * kb:o0 + (m + 1) * kb], int(it[q0 + m]), int(ok[q0 + m]),
                                                int(bch[q0 + m]))
        return res


def run(decs, duration=0.5, load_scale=1.0, seed=5, capacity=4096, batch_max=1, flush_ms=FLUSH_MS, trace=None,
        concurrent=LANES):
    """The receivers on the GPU for `duration` seconds of traffic; decs:
    decoder contexts (one batch in flight each), each with its decode's
    launches forked over `concurrent` streams (a batch mixes four codes,
    whose launches then overlap) or, with 0, one after another on its
    stream.  Returns the latency and throughput record, the trace and the
    frames' results."""
    tr = trace or Trace(decs[0], duration, load_scale, seed)
    sink = {}

    def on_deliver(s, f):
        sink.setdefault(s, []).append(f)

    g = None
    try:
        tr.pin(decs[0])
        g = GpuFec(decs, tr, capacity)
        for dk in decs:
            dk.set_concurrent(concurrent)
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
        g.finished.clear()
        t0 = time.perf_counter()
        done, delivered, sent, launches = rx_loop(tr.arr, tr.sid, g.submit, g.poll, len(decs), capacity, batch_max,
                                                  flush_ms, on_deliver, t0=t0)
    finally:
        for dk in decs:
            dk.set_concurrent(True)
        if g:
            g.close()
        tr.release(decs[0])
    results = g.results()
    n = tr.n
    # delivery in arrival order per receiver, every frame once
    for s, lst in sink.items():
        assert lst == sorted(lst) and len(set(lst)) == len(lst)
    assert sum(len(v) for v in sink.values()) == n
    errors = sum(int(not np.array_equal(results[f][0][:len(tr.ref(f))], tr.ref(f))) for f in range(n))
    bch_fail = sum(int(results[f][3] < 0) for f in range(n))
    lat = (delivered - tr.arr) * 1e3
    wait = (sent - tr.arr) * 1e3
    kl = np.array([tr.kinds[k]["fec"].kldpc for k in tr.kind], np.int64)
    busy = sum(x[2] for x in launches)
    span = max(float(delivered.max()), duration)
    it = np.array([results[f][1] for f in range(n)])

    def pct(a):
        return {"p50": round(float(np.percentile(a, 50)), 3), "p99": round(float(np.percentile(a, 99)), 3),
                "max": round(float(a.max()), 3)}

    rec = {
        "decoder": "plain", "frames": int(n), "duration_s": duration, "load_scale": load_scale,
        "contexts": len(decs), "capacity_frames_per_context": capacity, "launch_streams": int(concurrent),
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
        "frame_errors_by_stream": {STREAMS[k][0]: int(sum(not np.array_equal(results[f][0][:len(tr.ref(f))], tr.ref(f))
                                                        for f in np.nonzero(tr.kind == k)[0]))
                                   for k in range(len(STREAMS)) if (tr.kind == k).any()},
        "includes": "symbol / cell upload, demapper or T2 front end, LDPC, BCH, results download, in-order hand-off",
    }
    return rec, tr, results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--duration", type=float, default=0.5)
    ap.add_argument("--load", type=float, default=1.0)
    ap.add_argument("--contexts", type=int, default=2)
    ap.add_argument("--capacity", type=int, default=4096)
    ap.add_argument("--batch-max", type=int, default=1)
    ap.add_argument("--flush-ms", type=float, default=FLUSH_MS)
    ap.add_argument("--lanes", type=int, default=LANES, help="forked streams per context; 0: launches one after another")
    a = ap.parse_args()
    decs = [ldpcgpu.Decoder(0) for _ in range(a.contexts)]
    try:
        rec, _, _ = run(decs, a.duration, a.load, capacity=a.capacity, batch_max=a.batch_max, flush_ms=a.flush_ms,
                        concurrent=a.lanes)
    finally:
        for d in decs:
            d.close()
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
