"""Pins of the oracle's BCH decoder (oracle_bch_decode, restating
libdvbrx/src/bch.c:106-290): against the reference decoder built from its
own sources, on codewords with 0 .. t errors (always corrected back to the
codeword) and with more (the reference's result, -1 or another codeword,
reproduced byte for byte)."""
import numpy as np
import pytest

from helpers import flip, gen, oracle, oracle_bch, ref_bch, ref_builds

REFS = ref_builds()
needs_ref = pytest.mark.skipif(not REFS, reason="oracle/_ref not built (needs the reference tree)")
FECS = [(0, "1_4"), (0, "1_2"), (0, "2_3"), (0, "9_10"), (1, "1_4"), (1, "1_2"), (1, "8_9"), (0, "2_3_t2"),
        (0, "11_20"), (1, "11_45")]


def _codeword(short, rate, seed):
    fec, _, ref = gen(short, rate, 1000.0, 1, seed)
    return fec, ref[0]


@pytest.mark.parametrize("short,rate", FECS)
def test_corrects_up_to_t(short, rate):
    fec, cw = _codeword(short, rate, 11)
    assert oracle().oracle_bch_nbch(short, rate.encode()) == fec.kldpc
    rng = np.random.default_rng(len(rate) + 100 * short)
    for w in sorted({0, 1, 2, fec.t // 2, fec.t - 1, fec.t}):
        for trial in range(3):
            bits = rng.choice(fec.kldpc, w, replace=False)
            r, out = oracle_bch(short, rate, flip(cw, bits))
            assert r == w and (out == cw).all(), (w, trial)


@needs_ref
@pytest.mark.parametrize("short,rate", FECS)
def test_matches_reference(short, rate):
    fec, cw = _codeword(short, rate, 12)
    rng = np.random.default_rng(7 + len(rate) + 100 * short)
    R = REFS.get("v4") or next(iter(REFS.values()))
    fails = 0
    for w in [0, 1, fec.t, fec.t + 1, fec.t + 1, fec.t + 2, fec.t + 3, 2 * fec.t + 3, 40]:
        bits = rng.choice(fec.kldpc, w, replace=False)
        bad = flip(cw, bits)
        r, out = oracle_bch(short, rate, bad)
        rr, rout = ref_bch(R, fec, rate, bad)
        assert r == rr and (out == rout).all(), (w, r, rr)
        fails += r < 0
    assert fails >= 3  # beyond t the decoder mostly gives up


def test_unknown_rate():
    assert oracle().oracle_bch_nbch(0, b"7_7") == -1
    assert oracle_bch(0, "7_7", np.zeros(10, np.uint8))[0] == -2
