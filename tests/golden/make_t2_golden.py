#!/usr/bin/env python3
"""Write the golden fixture of the DVB-T2 FEC block path (t2_golden.json).

Inputs are regenerated from seeds: random cells and channel state
(tests/helpers.t2_random_cells), and encoded blocks through the transmit
model (tests/helpers.t2_codewords + t2_modulate).  Expected outputs come
from the reference's own decode_job (libdvbrx/src/dvbt2.c), compiled from
its sources by oracle/Makefile into oracle/_ref/libdrxt2ref_v3.so (AVX2,
32 lanes).  The fixture holds, per case, the generator parameters and the
SHA-256 of the reference's soft values, MER sums (float32 bits),
BBFRAMEs and BCH results, so that the oracle stays pinned where the
reference tree is absent.

Run from the repository root after `python3 -c "import __graft_entry__ as
g; g.build()"` (needs oracle/_ref, i.e. the reference tree mounted).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import helpers as H  # noqa: E402

# front-end cases: (cod, mod, fec_type, rotation, til_length, til_type, nblocks, snr_db, seed)
FRONT = [(c, m, f, r, 3, 0, 3, 12.0 + 3 * m, 100 + 17 * f + 5 * m + c)
         for f in (1, 0) for m in range(4) for c in (1, 4) for r in (0, 1)]
# whole-job cases: (cod, mod, fec_type, rotation, til_length, til_type, nblocks, snr_db, seed)
JOBS = [(0, 1, 0, 1, 2, 0, 6, 8.6, 11), (3, 2, 0, 1, 0, 0, 6, 15.0, 5), (2, 3, 0, 0, 4, 0, 6, 19.0, 7),
        (1, 3, 1, 1, 3, 0, 4, 18.4, 29)]


def sha(a):
    import hashlib

    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def front_inputs(case):
    cod, mod, ft, rot, tl, tt, nb, snr, seed = case
    N, nc, _ = H.t2_geometry(cod, mod, ft)
    return H.t2_random_cells(mod, nc, nb, np.random.default_rng(seed), snr_db=snr)


def job_inputs(case):
    cod, mod, ft, rot, tl, tt, nb, snr, seed = case
    rng = np.random.default_rng(seed)
    info, bits = H.t2_codewords(cod, ft, nb, rng)
    cells, csi = H.t2_modulate(cod, mod, ft, rot, (tl, tt), bits, snr, rng)
    return info, cells, csi


def main():
    refs = H.t2_ref_builds()
    if "v3" not in refs:
        sys.exit("needs oracle/_ref/libdrxt2ref_v3.so and an AVX2 host")
    L, lanes = refs["v3"]
    out = {"reference": "libdvbrx dvbt2.c decode_job, x86-64-v3 build (clang, -ffp-contract=fast), 32 lanes",
           "front_end": [], "jobs": []}
    for case in FRONT:
        cod, mod, ft, rot, tl, tt, nb, snr, seed = case
        cells, csi = front_inputs(case)
        _, _, err, sig, llr = H.ref_t2_job(L, cod, mod, ft, rot, (tl, tt), cells, csi, 1, with_llr=True)
        out["front_end"].append({"case": list(case), "llr_sha256": sha(llr), "err_sha256": sha(err),
                                 "sig_sha256": sha(sig)})
    for case in JOBS:
        cod, mod, ft, rot, tl, tt, nb, snr, seed = case
        info, cells, csi = job_inputs(case)
        bb, ok, err, sig, _ = H.ref_t2_job(L, cod, mod, ft, rot, (tl, tt), cells, csi, 30)
        out["jobs"].append({"case": list(case), "lanes": lanes, "max_iter": 30, "bb_sha256": sha(bb),
                            "ok": ok.tolist(), "err_sha256": sha(err), "sig_sha256": sha(sig),
                            "bb_equals_info": [bool((bb[b] == info[b]).all()) for b in range(nb)]})
    with open(os.path.join(HERE, "t2_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(f"wrote {len(out['front_end'])} front-end cases and {len(out['jobs'])} jobs")


if __name__ == "__main__":
    main()
