#!/usr/bin/env python3
"""Frame / bit error rates of the learned decoder against the plain one.

For each Es/N0 of a sweep (0.1 dB steps by default), frames of the
MODCOD are generated on the GPU (pool codewords through ldpc_gpu_channel)
and decoded twice from the same LLRs:

  plain    layered offset min-sum, up to --plain-iters iterations, each
           frame with its own early stop (ldpc_gpu_decode_batch);
  learned  the learned table (iterations past its T rows reuse the last
           row) for up to --learned-iters iterations, then the plain
           decoder from the channel LLRs for frames still failing their
           checks (ldpc_gpu_decode_batch_nms, fallback --plain-iters).

Frame errors are counted twice: "ldpc" if any of the kldpc decoded bits
differs from the codeword, "bch" if more than the BCH code's t of them
do (the outer BCH decoder, a bounded-distance decoder, corrects up to t
errors in its nbch = kldpc bits, so this is the receiver's frame error).
Reported per point: FERs, BER, average iterations (learned: learned +
fallback iterations), fallback share, decode time and decoded info
Gbit/s; and the lowest Es/N0 of the sweep from which on each decoder
keeps the post-BCH FER <= the target.
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
from nms_train import DevBuf, codeword_pool  # noqa: E402


def evaluate(table, esn0_list, frames_per_point, plain_iters=50, learned_iters=30, batch=4096, pool_n=1024, seed=99,
             device=0, target=1e-4, progress=True):
    pls = table["pls"]
    short, rate = table["short_frame"], table["rate"]
    code = ldpcgpu.code_lookup(short, rate)
    fec = ldpcgpu.fec_lookup(short, rate)
    N, kl = fec.nldpc, fec.kldpc
    kb = (kl + 7) // 8
    S = np.array(table["S"], np.uint8)
    B = np.array(table["B"], np.int16)
    T = S.shape[0]
    pool = codeword_pool(short, rate, pool_n, seed)
    # reference bits of each pool row: the first kldpc codeword bits, packed
    ref = np.packbits(pool[:, :kl], axis=1)
    dec = ldpcgpu.Decoder(device)
    dec.set_nms(code, S, B)
    dv = DevBuf(dec)
    out = {"pls": pls, "rate": rate, "constellation": table["constellation"], "T": T, "plain_iters": plain_iters,
           "learned_iters": learned_iters, "bch_t": fec.t, "frames_per_point": frames_per_point, "points": []}
    try:
        d_pool = dv.put(pool)
        d_llr = dv.empty(batch * N)
        d_out = dv.empty(batch * kb)
        d_it = dv.empty(4 * batch)
        d_ok = dv.empty(batch)
        d_fb = dv.empty(batch)
        d_rows = dv.empty(4 * batch)
        d_esn0 = dv.empty(4 * batch)
        frames = np.zeros(batch, ldpcgpu.FRAME_DTYPE)
        frames["code"] = code
        frames["nbits"] = kl
        frames["group"] = -1
        frames["llr_off"] = np.arange(batch) * N
        frames["out_off"] = np.arange(batch) * kb
        rng = np.random.default_rng(seed)
        for pi, e in enumerate(esn0_list):
            acc = {m: dict(fe=0, fb_fe=0, be=0, it=0, fb=0, frames=0, sec=0.0) for m in ("plain", "learned")}
            done = 0
            while done < frames_per_point:
                n = min(batch, frames_per_point - done)
                rows = rng.integers(0, pool_n, batch).astype(np.int32)
                dec.upload(d_rows, rows)
                dec.upload(d_esn0, np.full(batch, e, np.float32))
                ldpcgpu.channel(pls, N, d_pool, d_rows, n, d_esn0, seed * 7919 + pi * 100003 + done, d_llr,
                                stream=dec.stream)
                for mode in ("plain", "learned"):
                    fr = frames[:n].copy()
                    dec.synchronize()
                    t = time.perf_counter()
                    if mode == "plain":
                        fr["max_iter"] = plain_iters
                        dec.decode_batch(fr, d_llr, d_out, d_it, d_ok)
                    else:
                        fr["max_iter"] = learned_iters
                        dec.decode_batch_nms(fr, d_llr, d_out, d_it, d_ok, plain_iters, d_fb)
                    dec.synchronize()
                    dt = time.perf_counter() - t
                    bits = dv.get(d_out, (batch, kb), np.uint8)[:n]
                    it = dv.get(d_it, batch, np.int32)[:n]
                    diff = np.unpackbits(bits ^ ref[rows[:n]], axis=1)[:, :kl]
                    errs = diff.sum(axis=1)
                    a = acc[mode]
                    a["fe"] += int((errs > 0).sum())
                    a["fb_fe"] += int((errs > fec.t).sum())
                    a["be"] += int(errs.sum())
                    a["it"] += int(it.sum())
                    a["frames"] += n
                    a["sec"] += dt
                    if mode == "learned":
                        a["fb"] += int(dv.get(d_fb, batch, np.uint8)[:n].sum())
                done += n
            pt = {"esn0_db": round(float(e), 3)}
            for mode, a in acc.items():
                pt[mode] = {"fer": a["fb_fe"] / a["frames"], "frame_errors": a["fb_fe"],
                            "fer_ldpc": a["fe"] / a["frames"], "ber": a["be"] / (a["frames"] * kl),
                            "avg_iterations": a["it"] / a["frames"], "seconds": round(a["sec"], 4),
                            "gbit_s": round(a["frames"] * kl / a["sec"] / 1e9, 3)}
                if mode == "learned":
                    pt[mode]["fallback_share"] = a["fb"] / a["frames"]
            out["points"].append(pt)
            if progress:
                print(f"# {e:.2f} dB plain FER {pt['plain']['fer']:.2e} (ldpc {pt['plain']['fer_ldpc']:.1e}) it {pt['plain']['avg_iterations']:.2f} "
                      f"{pt['plain']['gbit_s']:.2f} Gbit/s | learned FER {pt['learned']['fer']:.2e} "
                      f"(ldpc {pt['learned']['fer_ldpc']:.1e}) "
                      f"it {pt['learned']['avg_iterations']:.2f} fb {pt['learned']['fallback_share']:.3f} "
                      f"{pt['learned']['gbit_s']:.2f} Gbit/s", file=sys.stderr, flush=True)
    finally:
        dv.free()
        dec.close()
    # lowest Es/N0 from which on FER stays <= target
    for mode in ("plain", "learned"):
        thr = None
        for pt in sorted(out["points"], key=lambda p: -p["esn0_db"]):
            if pt[mode]["fer"] <= target:
                thr = pt["esn0_db"]
            else:
                break
        out[f"{mode}_esn0_at_target"] = thr
    out["target_fer"] = target
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--table", required=True)
    ap.add_argument("--from-db", type=float, required=True, help="first Es/N0, relative to the MODCOD's QEF")
    ap.add_argument("--to-db", type=float, required=True)
    ap.add_argument("--step", type=float, default=0.1)
    ap.add_argument("--frames", type=int, default=20000)
    ap.add_argument("--plain-iters", type=int, default=50)
    ap.add_argument("--learned-iters", type=int, default=30)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    table = json.load(open(a.table))
    qef = ldpcgpu.qef_db(table["pls"])
    grid = [qef + a.from_db + a.step * k for k in range(int(round((a.to_db - a.from_db) / a.step)) + 1)]
    res = evaluate(table, grid, a.frames, a.plain_iters, a.learned_iters)
    res["qef_db"] = qef
    with open(a.out, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps({k: res[k] for k in ("pls", "rate", "plain_esn0_at_target", "learned_esn0_at_target")}))


if __name__ == "__main__":
    main()
