#!/usr/bin/env python3
"""Frame error rate of a DVB-T2 PLP through the GPU FEC path, by Es/N0.

For each Es/N0, nblocks independently encoded and noised blocks
(workloads.t2_blocks with every block distinct) go through
ldpc_gpu_t2_decode (lanes 0, max_iter 30 as the reference's T2 FEC
thread).  Reported per point: blocks whose LDPC decoder did not reach a
codeword (syndrome flag), blocks whose LDPC output differs from the BCH
codeword, blocks BCH could not decode or left wrong (the receiver's
frame errors), average iterations, and the decode time.

    python3 dvb-rx_amd/t2_eval.py --cod 1 --mod 3 --fec-type 1 --from-db 17.8 --to-db 20.2 --step 0.4 \\
        --blocks 1024 --out t2_fer.json
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import ldpcgpu  # noqa: E402
import workloads  # noqa: E402


def point(dec, plp, nblocks, esn0, seed, max_iter=30):
    N, K, kbch, nc = ldpcgpu.t2_geometry(plp)
    cells, csi, info, bchcw = workloads.t2_blocks(plp, nblocks, esn0, seed=seed, distinct=nblocks)
    kb = K // 8
    sizes = dict(cells=cells.nbytes, csi=csi.nbytes, llr=nblocks * N, cw=nblocks * kb, it=4 * nblocks,
                 syn=nblocks, bch=4 * nblocks, err=4 * nblocks, sig=4 * nblocks)
    d = {k: dec.malloc(v) for k, v in sizes.items()}
    try:
        dec.upload(d["cells"], cells)
        dec.upload(d["csi"], csi)
        # the LDPC output before BCH: decode once without the BCH pass
        ldpcgpu.t2_demap(dec, plp, nblocks, d["cells"], d["csi"], d["llr"], d["err"], d["sig"])
        fr = np.zeros(nblocks, ldpcgpu.FRAME_DTYPE)
        fr["code"] = ldpcgpu.code_lookup(plp.fec_type == 0, code_name(plp))
        fr["max_iter"] = max_iter
        fr["nbits"] = K
        fr["group"] = -1
        fr["llr_off"] = np.arange(nblocks, dtype=np.uint64) * N
        fr["out_off"] = np.arange(nblocks, dtype=np.uint64) * kb
        t0 = time.perf_counter()
        dec.decode_batch(fr, d["llr"], d["cw"], d["it"], d["syn"])
        dec.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        cw = np.zeros((nblocks, kb), np.uint8)
        it = np.zeros(nblocks, np.int32)
        syn = np.zeros(nblocks, np.uint8)
        dec.download(cw, d["cw"])
        dec.download(it, d["it"])
        dec.download(syn, d["syn"])
        ldpc_wrong = int((cw != bchcw).any(1).sum())
        # then the whole job (front end, LDPC, BCH) as the receiver runs it
        ldpcgpu.t2_decode(dec, plp, nblocks, d["cells"], d["csi"], max_iter, 0, d["llr"], d["cw"], d["it"],
                          d["syn"], d["bch"], d["err"], d["sig"])
        dec.synchronize()
        res = np.zeros(nblocks, np.int32)
        dec.download(cw, d["cw"])
        dec.download(res, d["bch"])
        bad = int(((res < 0) | (cw[:, :kbch // 8] != info).any(1)).sum())
    finally:
        for v in d.values():
            dec.free(v)
    return {"esn0_db": round(esn0, 3), "blocks": nblocks, "ldpc_unconverged": int((syn == 0).sum()),
            "ldpc_wrong": ldpc_wrong, "fer_ldpc": ldpc_wrong / nblocks, "frame_errors_after_bch": bad,
            "fer_after_bch": bad / nblocks, "avg_iterations": round(float(it.mean()), 3),
            "ldpc_ms": round(ms, 3)}


def code_name(plp):
    sf = plp.fec_type == 0
    rate = workloads.T2_RATES[plp.cod]
    return "2_3_t2" if (not sf and plp.cod == 2) else "3_5_t2" if (sf and plp.cod == 1) else rate


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cod", type=int, default=1)
    ap.add_argument("--mod", type=int, default=3)
    ap.add_argument("--fec-type", type=int, default=1)
    ap.add_argument("--rotation", type=int, default=1)
    ap.add_argument("--from-db", type=float, required=True)
    ap.add_argument("--to-db", type=float, required=True)
    ap.add_argument("--step", type=float, default=0.2)
    ap.add_argument("--blocks", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    plp = ldpcgpu.T2Plp(a.cod, a.mod, a.fec_type, a.rotation, 8, 0)
    dec = ldpcgpu.Decoder(0)
    pts = []
    try:
        for k, db in enumerate(np.arange(a.from_db, a.to_db + 1e-9, a.step)):
            p = point(dec, plp, a.blocks, float(db), a.seed + k)
            print(json.dumps(p), flush=True)
            pts.append(p)
    finally:
        dec.close()
    res = {"plp": {"cod": a.cod, "mod": a.mod, "fec_type": a.fec_type, "rotation": a.rotation,
                   "code": code_name(plp)}, "max_iter": 30, "points": pts}
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
