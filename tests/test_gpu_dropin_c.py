"""The drop-in from plain C (tests/c/dropin_ldpc.c, built by
__graft_entry__.build()): drx_ldpc_* called exactly as libdvbrx's callers
call them, batches of drx_ldpc_lanes() frames.  Each batch's return value
and every decoded byte must equal the oracle's group decode (the
reference's batch semantics) of the same frames."""
import os
import subprocess

import numpy as np
import pytest

from helpers import ROOT, Oracle, gen, oracle_bch

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "tests", "c", "dropin_ldpc")


@pytest.mark.parametrize("short,rate,pls,esn0,nf,mi", [(0, "1_2", 0x10, 1.5, 128, 50), (1, "3_5", -1, 2.4, 100, 30),
                                                     (0, "9_10", 0x2c, 6.6, 64, 50)])
def test_c_dropin_matches_oracle(tmp_path, short, rate, pls, esn0, nf, mi):
    assert os.path.exists(EXE), "tests/c/dropin_ldpc not built: run __graft_entry__.build()"
    print(f"# C drop-in {short} {rate} pls {pls} {esn0} dB x{nf}", flush=True)
    outf = tmp_path / "out.bin"
    r = subprocess.run([EXE, str(short), rate, hex(pls) if pls >= 0 else "-1", str(esn0), str(nf), str(mi), str(outf)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = dict(line.split(" ", 1) for line in r.stdout.strip().splitlines())
    lanes = int(lines["lanes"])
    assert lanes == 64
    got = [int(x) for x in lines["batches"].split()]
    fec, llr, ref = gen(short, rate, esn0, nf, 2024, pls=pls)
    kb = fec.kldpc // 8
    out = np.fromfile(outf, np.uint8).reshape(nf, kb)
    o = Oracle(short, rate, fec.kldpc)
    want = []
    for f0 in range(0, nf, lanes):
        it, bits = o.group(llr[f0:f0 + lanes], mi)
        want.append(int(it))
        assert (out[f0:f0 + lanes] == bits).all(), f0
    assert got == want
    errors = int(sum((out[f] != ref[f]).any() for f in range(nf)))
    assert lines["frames"] == f"{nf} in_error {errors}"


@pytest.mark.parametrize("short,rate,esn0", [(0, "1_2", 1.5), (1, "3_5", 2.4)])
def test_c_dropin_bch_matches_oracle(tmp_path, short, rate, esn0):
    """drx_fec_lookup and drx_bch_* of the drop-in header, from C as rx.c
    calls them after the LDPC decoder: decoded frames with 0 .. t + 2 bits
    flipped, each result and codeword equal to the oracle's BCH decoder
    (the reference's drx_bch_decode restated)."""
    assert os.path.exists(EXE), "tests/c/dropin_ldpc not built: run __graft_entry__.build()"
    print(f"# C drop-in BCH {short} {rate} {esn0} dB", flush=True)
    nf = 40
    outf, bchf = tmp_path / "out.bin", tmp_path / "bch.bin"
    r = subprocess.run([EXE, str(short), rate, "-1", str(esn0), str(nf), "50", str(outf), str(bchf)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = dict(line.split(" ", 1) for line in r.stdout.strip().splitlines())
    got = [int(x) for x in lines["bch"].split()]
    fec, _, _ = gen(short, rate, esn0, 1, 2024)
    kb = fec.kldpc // 8
    dec = np.fromfile(outf, np.uint8).reshape(nf, kb)
    cws = np.fromfile(bchf, np.uint8).reshape(nf, kb)
    seen = set()
    for f in range(nf):
        cw = dec[f].copy()
        for k in range(f % (fec.t + 3)):
            bit = (131 * k + 7 * f) % fec.kldpc
            cw[bit // 8] ^= 0x80 >> (bit % 8)
        want, wcw = oracle_bch(short, rate, cw)
        assert got[f] == want, (f, got[f], want)
        assert (cws[f] == wcw).all(), f
        seen.add(want)
    assert -1 in seen and 0 in seen and fec.t in seen
