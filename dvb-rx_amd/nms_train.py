#!/usr/bin/env python3
"""Train the learned offset min-sum table of one MODCOD's code on the GPU.

Data: a pool of codewords of the code (random information bits, BCH +
LDPC encoded by lib/libdvbs2sim.so); every training frame is a pool
codeword sent through the GPU channel (ldpc_gpu_channel: the MODCOD's
constellation and bit interleaver, AWGN at an Es/N0 drawn uniformly from
[QEF + lo, QEF + hi] per frame, max-log int8 LLRs) with fresh noise.

Model: T unrolled layered iterations, one (alpha, beta) pair per iteration
and check degree, starting from the plain decoder (alpha = beta = 1);
forward with the quantised integer table, backward through the float
model (include/ldpc_nms.h), Adam on the GPU, loss = sum over iterations
of the mean binary cross-entropy of the posteriors.

Writes a JSON table: float parameters, the quantised (S, B) integer table
that ldpc_gpu_set_nms takes, and the training record.
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

DEGS = ldpcgpu.NMS_DEGS


def codeword_pool(short, rate, n, seed):
    """n codewords (one byte per bit) of the code of (short, rate): the
    signs of noiseless frames of the frame source."""
    specs = [dict(short_frame=short, rate=rate, pls=-1, esn0_db=1000.0, stream=f) for f in range(n)]
    llr, _, layout = ldpcgpu.generate(specs, seed, with_ref=False)
    N = layout[0][1]
    return (llr.reshape(n, N) < 0).astype(np.uint8)


class DevBuf:
    """Device buffers allocated through a decoder context."""

    def __init__(self, dec):
        self.dec = dec
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.dec.malloc(max(a.nbytes, 1))
        self.dec.upload(p, a)
        self.ptrs.append(p)
        return p

    def empty(self, nbytes):
        p = self.dec.malloc(nbytes)
        self.ptrs.append(p)
        return p

    def get(self, p, shape, dtype):
        out = np.zeros(shape, dtype)
        self.dec.download(out, p)
        return out

    def free(self):
        for p in self.ptrs:
            self.dec.free(p)
        self.ptrs = []


def train(pls, T=10, steps=2000, batch=128, lr=1e-3, temp=0.5, pool_n=1024, seed=1, lo=-0.5, hi=1.0,
          log_every=100, device=0, progress=True):
    m = ldpcgpu.pls_lookup(pls)
    short, rate = m["short_frame"], m["rate"]
    code = ldpcgpu.code_lookup(short, rate)
    fec = ldpcgpu.fec_lookup(short, rate)
    N = fec.nldpc
    qef = ldpcgpu.qef_db(pls)
    t0 = time.time()
    pool = codeword_pool(short, rate, pool_n, seed)
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, pool_n, (steps, batch)).astype(np.int32)
    esn0 = (qef + rng.uniform(lo, hi, (steps, batch))).astype(np.float32)
    params = np.ones((2 * T, DEGS), np.float32)  # alpha rows, then beta rows: the plain decoder
    dec = ldpcgpu.Decoder(device)
    tr = ldpcgpu.NmsTrainer(code, T, batch, device)
    dv = DevBuf(dec)
    hist = []
    try:
        d_pool = dv.put(pool)
        d_rows = dv.put(rows)
        d_esn0 = dv.put(esn0)
        d_llr = dv.empty(batch * N)
        d_par = dv.put(params)
        d_m = dv.put(np.zeros_like(params))
        d_v = dv.put(np.zeros_like(params))
        d_grad = dv.empty(params.nbytes)
        d_loss = dv.empty(4)
        d_w = dv.put(np.ones(T, np.float32))
        np_ = params.size
        t1 = time.time()
        for k in range(steps):
            r = d_rows + 4 * batch * k
            ldpcgpu.channel(pls, N, d_pool, r, batch, d_esn0 + 4 * batch * k, seed * 1000003 + k, d_llr,
                            stream=tr.stream)
            tr.step(d_llr, d_pool, r, batch, d_par, d_w, temp, d_loss, d_grad)
            tr.adam(d_par, d_grad, d_m, d_v, np_, lr, k + 1)
            if (k + 1) % log_every == 0 or k + 1 == steps:
                tr.synchronize()
                loss = float(dv.get(d_loss, 1, np.float32)[0])
                hist.append((k + 1, loss))
                if progress:
                    print(f"# step {k + 1}/{steps} loss {loss:.5f} ({time.time() - t1:.1f} s)", file=sys.stderr,
                          flush=True)
        tr.synchronize()
        train_s = time.time() - t1
        params = dv.get(d_par, params.shape, np.float32)
    finally:
        dv.free()
        tr.close()
        dec.close()
    alpha, beta = params[:T], params[T:]
    S, B = ldpcgpu.nms_quantize(alpha, beta)
    used = sorted({int(d) for d in np.nonzero((np.abs(alpha - 1) > 1e-7) | (np.abs(beta - 1) > 1e-7))[1]})
    return {
        "pls": pls, "short_frame": short, "rate": rate, "constellation": m["constellation"], "code": code,
        "T": T, "degrees": used,
        "alpha": alpha.tolist(), "beta": beta.tolist(), "S": S.tolist(), "B": B.tolist(),
        "training": {"steps": steps, "batch": batch, "frames": steps * batch, "lr": lr, "optimizer": "adam",
                     "temp": temp, "loss_weights": "1 per iteration", "codeword_pool": pool_n,
                     "esn0_db": [qef + lo, qef + hi], "qef_db": qef, "seed": seed,
                     "seconds": round(train_s, 2), "setup_seconds": round(t1 - t0, 2), "loss": hist},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pls", type=lambda x: int(x, 0), required=True)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--temp", type=float, default=0.5)
    ap.add_argument("--pool", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--lo", type=float, default=-0.5, help="lowest training Es/N0 relative to the MODCOD's QEF")
    ap.add_argument("--hi", type=float, default=1.0, help="highest training Es/N0 relative to the QEF")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    res = train(a.pls, a.iters, a.steps, a.batch, a.lr, a.temp, a.pool, a.seed, a.lo, a.hi)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps({k: res[k] for k in ("pls", "rate", "constellation", "T", "degrees", "S", "B")}))


if __name__ == "__main__":
    main()
