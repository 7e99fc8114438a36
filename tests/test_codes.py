"""Code registry, layer schedule and frame source (CPU).

* every rate of both frame sizes resolves to a code whose tables give the
  reference's link count (checked when the registry is built) and whose
  dependency levels are a valid schedule;
* the encoder's codewords satisfy every parity check of the oracle's
  parity-check matrix, and the reference decoder (oracle/_ref) accepts them
  in zero iterations, which ties the generated tables to the reference's;
* BCH codewords of the frame source decode with zero corrections in the
  reference's BCH decoder;
* the channel model's LLR statistics follow from its definition.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import ALL_RATES, Oracle, gen, kmodel, ldpcgpu, ref_builds, ref_decode

CODES = [(s, r) for s in (0, 1) for r in ALL_RATES[s]]


def test_registry_counts():
    L = ldpcgpu.gpu_lib()
    ids = {ldpcgpu.code_lookup(s, r) for s, r in CODES}
    assert -1 not in ids
    assert len(ids) == L.ldpc_gpu_code_count()
    assert L.ldpc_gpu_code_lookup(0, b"7_8") == -1 and L.ldpc_gpu_code_lookup(1, b"9_10") == -1


@pytest.mark.parametrize("short,rate", CODES)
def test_code_params(short, rate):
    code = ldpcgpu.code_lookup(short, rate)
    p = ldpcgpu.code_params(code)
    fec = ldpcgpu.fec_lookup(short, rate)
    assert p["n"] == (16200 if short else 64800) == fec.nldpc
    assert p["q"] * 360 == p["n"] - p["k"]
    assert fec.kldpc <= p["k"] and fec.kbch < fec.kldpc
    assert 2 <= p["max_degree"] <= 28
    assert kmodel().kmodel_check_levels(code) == 0


def test_t2_rates_share_s2_bch():
    """T2 rate names resolve to the FEC row of the same numeric rate."""
    a = ldpcgpu.fec_lookup(0, "2_3_t2")
    b = ldpcgpu.fec_lookup(0, "2_3")
    assert (a.nldpc, a.kbch, a.t) == (b.nldpc, b.kbch, b.t)


def _encode(short, rate, seed):
    L = ldpcgpu.sim_lib()
    fec = ldpcgpu.fec_lookup(short, rate)
    code = ldpcgpu.code_lookup(short, rate)
    k = ldpcgpu.code_params(code)["k"]
    info = np.random.default_rng(seed).integers(0, 2, k).astype(np.uint8)
    cw = np.zeros(fec.nldpc, np.uint8)
    assert L.sim_ldpc_encode(short, rate.encode(), info.ctypes.data, cw.ctypes.data) == 0
    return fec, info, cw


@pytest.mark.parametrize("short,rate", CODES)
def test_encoder_codewords(short, rate):
    fec, info, cw = _encode(short, rate, 3)
    assert (cw[:len(info)] == info).all()  # systematic
    llr = np.where(cw == 1, -90, 90).astype(np.int8)
    o = Oracle(short, rate, fec.kldpc)
    assert o.check(llr) == 1
    # any single parity bit flipped breaks the checks
    for v in (len(info), fec.nldpc - 1, (len(info) + fec.nldpc) // 2):
        bad = llr.copy()
        bad[v] = -bad[v]
        assert o.check(bad) == 0


REFS = ref_builds()


@pytest.mark.skipif(not REFS, reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("short,rate", CODES)
def test_encoder_codewords_pass_reference_checks(short, rate):
    fec, info, cw = _encode(short, rate, 4)
    llr = np.where(cw == 1, -90, 90).astype(np.int8)[None]
    R = next(iter(REFS.values()))
    it, out = ref_decode(R, fec, rate, llr, 5)
    assert it == 0
    bits = np.unpackbits(out[0])[:fec.kldpc]
    assert (bits == info[:fec.kldpc]).all()


@pytest.mark.skipif(not REFS, reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("short,rate", [(0, "1_2"), (0, "9_10"), (1, "3_5"), (1, "8_9"), (0, "2_9_VLSNR"), (0, "2_3_t2")])
def test_bch_codewords_pass_reference_bch(short, rate):
    """Noiseless frames: the sent BCH codeword decodes with zero
    corrections; three flipped bits are corrected (t >= 8)."""
    fec, _, ref = gen(short, rate, 1000.0, 1, 9)
    R = next(iter(REFS.values()))
    b = R.drx_bch_new(C.byref(fec.as_struct()))
    assert b
    try:
        cw = ref[0].copy()
        assert R.drx_bch_decode(b, cw.ctypes.data) == 0
        bad = ref[0].copy()
        for bit in (0, 1000, fec.kldpc - 1):
            bad[bit // 8] ^= 0x80 >> (bit % 8)
        assert R.drx_bch_decode(b, bad.ctypes.data) == 3
        assert (bad == ref[0]).all()
    finally:
        R.drx_bch_free(b)


def test_unit_channel_statistics():
    """llr = quantise(4 y / sigma^2), y = +-1 + N(0, sigma^2): the mean of
    the LLR magnitude given the sent bit is 4/sigma^2 (before clipping)."""
    esn0 = 2.0
    fec, llr, ref = gen(0, "1_2", esn0, 2, 123)
    sigma2 = 1.0 / (10 ** (esn0 / 10))
    bits = np.unpackbits(ref[0])[:fec.kldpc]
    signed = llr[0][:fec.kldpc].astype(np.float64) * np.where(bits == 1, -1, 1)
    assert abs(signed.mean() - 4 / sigma2) / (4 / sigma2) < 0.05
    assert llr.min() >= -127 and llr.max() <= 127


def test_modcod_frames_decode():
    """A QPSK 1/2 frame through the MODCOD path (pls 0x10) at 3 dB decodes."""
    fec, llr, ref = gen(0, "1_2", 3.0, 2, 5, pls=0x10)
    o = Oracle(0, "1_2", fec.kldpc)
    bits, it, ok = o.frames(llr, 30)
    assert ok.all() and (bits == ref).all()
