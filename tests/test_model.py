"""CPU models of the two kernel schedules against the oracle.

tests/model/libkmodel.so compiles the device code's node routines
(dvb-rx_amd/csrc/ldpc_node.h, ldpc_pk.h) for the host and drives them in
the kernels' order:

* kmodel_decode: the three-phase layer schedule of the one-frame kernel
  (dependency levels inside a layer, int8 messages);
* kmodel_pk_pair: the packed two-frame kernel (two frames in the high and
  low bytes of 16-bit lanes), the second frame entering its slot a few
  iterations after the first, as when a slot is refilled.

Both must reproduce the oracle's per-frame decode bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import ALL_RATES, Oracle, esn0_for, gen, kmodel, kmodel_decode, ldpcgpu

CODES = [(s, r) for s in (0, 1) for r in ALL_RATES[s]]


@pytest.mark.parametrize("short,rate", CODES)
def test_layer_schedule_model(short, rate):
    fec, llr, _ = gen(short, rate, esn0_for(short, rate), 1, 900 + ALL_RATES[short].index(rate), stream0=short)
    code = ldpcgpu.code_lookup(short, rate)
    mi = 8
    bits, it, ok = Oracle(short, rate, fec.kldpc).frames(llr, mi)
    out, mit, mok = kmodel_decode(code, llr[0], mi, fec.kldpc)
    assert (mit, mok) == (int(it[0]), int(ok[0]))
    assert (out == bits[0]).all()


def pk_pair(code, a, b, delay, max_iter, nbits):
    L = kmodel()
    L.kmodel_pk_pair.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p]
    a = np.ascontiguousarray(a, np.int8)
    b = np.ascontiguousarray(b, np.int8)
    kb = (nbits + 7) // 8
    oa, ob = np.zeros(kb, np.uint8), np.zeros(kb, np.uint8)
    it = np.zeros(2, np.int32)
    ok = np.zeros(2, np.int32)
    rc = L.kmodel_pk_pair(code, a.ctypes.data, b.ctypes.data, delay, max_iter, nbits, oa.ctypes.data,
                          ob.ctypes.data, it.ctypes.data, ok.ctypes.data)
    assert rc == 0
    return oa, ob, it, ok


# codes with nodes sharing bits (in few or many layers) and without, of
# every slot count class
PK_CASES = [(0, "1_2", 0), (0, "1_2", 3), (0, "3_5", 5), (0, "9_10", 2), (1, "1_4", 4), (1, "8_9", 1),
            (0, "2_9_VLSNR", 2), (1, "11_45", 3), (0, "135_180", 1), (0, "2_3_t2", 2), (1, "3_5_t2", 0),
            (0, "1_4", 1), (0, "1_3", 0), (0, "2_5", 2), (0, "13_45", 1)]


@pytest.mark.parametrize("short,rate,delay", PK_CASES)
def test_packed_pair_model(short, rate, delay):
    fec, llr, _ = gen(short, rate, esn0_for(short, rate), 2, 40 + delay, stream0=7)
    code = ldpcgpu.code_lookup(short, rate)
    mi = 10
    bits, it, ok = Oracle(short, rate, fec.kldpc).frames(llr, mi)
    oa, ob, pit, pok = pk_pair(code, llr[0], llr[1], delay, mi, fec.kldpc)
    assert list(pit) == list(it) and list(pok) == list(ok)
    assert (oa == bits[0]).all() and (ob == bits[1]).all()


def test_packed_pair_extremes():
    """Saturated and zero LLRs in the two bytes of one lane: +127 next to
    -128 and 0 (the unpacking must not carry between the slots)."""
    fec = ldpcgpu.fec_lookup(1, "1_2")
    code = ldpcgpu.code_lookup(1, "1_2")
    rng = np.random.default_rng(2)
    a = rng.choice([-128, -127, 127, 0, 1, -1], fec.nldpc).astype(np.int8)
    b = rng.choice([-128, 127, 0], fec.nldpc).astype(np.int8)
    bits, it, ok = Oracle(1, "1_2", fec.kldpc).frames(np.stack([a, b]), 4)
    oa, ob, pit, pok = pk_pair(code, a, b, 0, 4, fec.kldpc)
    assert list(pit) == list(it) and list(pok) == list(ok)
    assert (oa == bits[0]).all() and (ob == bits[1]).all()


def pk_pair_nms(code, a, b, delay, S, B, max_iter, fallback, nbits):
    L = kmodel()
    vp = C.c_void_p
    L.kmodel_pk_pair_nms.argtypes = [C.c_int, vp, vp, C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp,
                                     vp, vp, vp]
    a = np.ascontiguousarray(a, np.int8)
    b = np.ascontiguousarray(b, np.int8)
    S = np.ascontiguousarray(S, np.uint8)
    B = np.ascontiguousarray(B, np.int16)
    kb = (nbits + 7) // 8
    oa, ob = np.zeros(kb, np.uint8), np.zeros(kb, np.uint8)
    it, ok, fb = np.zeros(2, np.int32), np.zeros(2, np.int32), np.zeros(2, np.int32)
    rc = L.kmodel_pk_pair_nms(code, a.ctypes.data, b.ctypes.data, delay, S.ctypes.data, B.ctypes.data, S.shape[0],
                              max_iter, fallback, nbits, oa.ctypes.data, ob.ctypes.data, it.ctypes.data,
                              ok.ctypes.data, fb.ctypes.data)
    assert rc == 0
    return oa, ob, it, ok, fb


@pytest.mark.parametrize("short,rate,delay,esn0", [(0, "3_5", 2, 2.3), (1, "1_2", 3, 1.2), (0, "2_3", 0, 3.1),
                                                   (1, "8_9", 1, 6.0), (0, "2_9_VLSNR", 2, -2.0)])
def test_packed_pair_learned_model(short, rate, delay, esn0):
    """The packed learned arithmetic (per-slot scale / offset by
    iteration and check degree, plain fallback after the learned cap)
    against the oracle's learned decoder."""
    fec, llr, _ = gen(short, rate, esn0, 2, 60 + delay, stream0=3)
    code = ldpcgpu.code_lookup(short, rate)
    T = 6
    rng = np.random.default_rng(delay)
    S = rng.integers(8, 20, (T, 33)).astype(np.uint8)
    B = rng.integers(-16, 48, (T, 33)).astype(np.int16)
    o = Oracle(short, rate, fec.kldpc)
    for learned_max, fallback in ((T, 30), (3, 0), (8, 30)):
        bits, it, ok, fb = o.nms_frames(S, B, llr, learned_max, fallback)
        oa, ob, pit, pok, pfb = pk_pair_nms(code, llr[0], llr[1], delay, S, B, learned_max, fallback, fec.kldpc)
        assert list(pit) == list(it) and list(pok) == list(ok) and list(pfb) == list(fb), (learned_max, fallback)
        assert (oa == bits[0]).all() and (ob == bits[1]).all()
