#!/usr/bin/env python3
"""Write the golden fixtures of the LDPC decode path.

Inputs come from this repository's frame source (lib/libdvbs2sim.so,
seeded); expected outputs come from the reference decoder itself
(libdvbrx/src/ldpc.c), compiled from its own sources by oracle/Makefile into
oracle/_ref/ (AVX-512, AVX2 and scalar builds, which must agree).  The
reference's own tests hold no vectors for this path (tests/test_ldpc.c
asserts nothing and its inputs are generated by GNU Radio), so these
fixtures are the pins:

  golden.json  per set: generator spec, SHA-256 of each frame's LLRs,
               per-frame decode (nframes = 1: iterations, syndrome, SHA-256
               of the bits) and the batch decode of the whole set
               (drx_ldpc_decode semantics: one count, bits per frame)
  vectors.npz  full LLRs and decoded bits of a few frames

Run from the repository root after `python3 -c "import __graft_entry__ as
g; g.build()"` (needs oracle/_ref, i.e. the reference tree mounted).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import helpers  # noqa: E402
from helpers import gen, ldpcgpu, sha256  # noqa: E402

# (name, short, rate, pls, Es/N0 dB, frames, seed, max_iter, group size)
SETS = [
    ("s2_n1_2_qpsk_1.5dB", 0, "1_2", 0x10, 1.5, 64, 2024, 50, 64),
    ("s2_n3_4_unit_4.2dB", 0, "3_4", -1, 4.2, 16, 11, 50, 16),
    ("s2_n9_10_unit_6.6dB", 0, "9_10", -1, 6.6, 16, 12, 30, 16),
    ("s2_s3_5_unit_2.4dB", 1, "3_5", -1, 2.4, 32, 13, 30, 32),
    ("s2_s8_9_unit_6.5dB", 1, "8_9", -1, 6.5, 32, 14, 20, 32),
    ("s2x_n135_180_unit_4.0dB", 0, "135_180", -1, 4.0, 16, 15, 25, 16),
    ("s2x_s11_45_unit_-1.6dB", 1, "11_45", -1, -1.6, 32, 16, 40, 32),
    ("t2_n2_3_unit_3.3dB", 0, "2_3_t2", -1, 3.3, 16, 17, 30, 16),
    ("s2_n1_4_unit_-2.3dB", 0, "1_4", -1, -2.3, 8, 18, 50, 8),
    ("s2_n1_4_unit_-1.9dB", 0, "1_4", -1, -1.9, 8, 19, 50, 8),
]


def main():
    refs = helpers.ref_builds()
    if not refs:
        sys.exit("oracle/_ref is not built (needs the reference tree)")
    out = {"generator": "lib/libdvbs2sim.so sim_generate", "reference_builds": sorted(refs), "sets": []}
    vec = {}
    for name, short, rate, pls, esn0, nf, seed, mi, gsz in SETS:
        fec, llr, refbits = gen(short, rate, esn0, nf, seed, pls=pls)
        per_frame = []
        for f in range(nf):
            res = {k: helpers.ref_decode(R, fec, rate, llr[f:f + 1], mi) for k, R in refs.items()}
            its = {k: v[0] for k, v in res.items()}
            bits = {k: sha256(v[1][0]) for k, v in res.items()}
            assert len(set(its.values())) == 1 and len(set(bits.values())) == 1, (name, f, its)
            it, b = res["v4" if "v4" in res else sorted(res)[0]]
            per_frame.append({
                "llr_sha256": sha256(llr[f]),
                "iterations": int(mi if it < 0 else it),
                "syndrome_ok": int(it >= 0),
                "bits_sha256": sha256(b[0]),
                "bits_equal_sent": bool((b[0] == refbits[f]).all()),
            })
        # batch (group) decode in groups of gsz frames, with every build
        # whose lane count covers the group
        groups = []
        for g0 in range(0, nf, gsz):
            gl = llr[g0:g0 + gsz]
            res = {k: helpers.ref_decode(R, fec, rate, gl, mi) for k, R in refs.items() if R.drx_ldpc_lanes() >= len(gl)}
            assert res, "no build with enough lanes"
            its = {k: v[0] for k, v in res.items()}
            assert len(set(its.values())) == 1, (name, its)
            it, b = next(iter(res.values()))
            for k, v in res.items():
                assert (v[1] == b).all(), (name, k)
            groups.append({"first": g0, "count": len(gl), "result": int(it),
                           "bits_sha256": [sha256(b[f]) for f in range(len(gl))]})
        out["sets"].append({
            "name": name, "short_frame": short, "rate": rate, "pls": pls, "esn0_db": esn0, "frames": nf,
            "seed": seed, "max_iter": mi, "kldpc": fec.kldpc, "per_frame": per_frame, "groups": groups,
        })
        if name.startswith("s2_n1_2") or name.startswith("s2_s3_5"):
            res = helpers.ref_decode(refs[sorted(refs)[0]], fec, rate, llr[:2], mi)
            vec[name + "_llr"] = llr[:2]
            vec[name + "_group2_bits"] = res[1]
            vec[name + "_group2_result"] = np.array([res[0]], np.int32)
        print(name, "iterations", [p["iterations"] for p in per_frame][:12], "groups",
              [g["result"] for g in groups], flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    np.savez_compressed(os.path.join(HERE, "vectors.npz"), **vec)


if __name__ == "__main__":
    main()
