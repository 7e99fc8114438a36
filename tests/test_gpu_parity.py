"""Parity of the HIP decoder (through the C-ABI) with the CPU oracle:
decoded bits, iteration counts and syndrome flags must be identical."""
import numpy as np
import pytest

import helpers
from helpers import ALL_RATES, Oracle, esn0_for, gen, ldpcgpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["auto", "single"])
def dec(request):
    """Both kernels: two frames per workgroup (packed 16-bit arithmetic) and
    one frame per workgroup."""
    d = ldpcgpu.Decoder(0, kernel=request.param)
    yield d
    d.close()


def run_frames(dec, fec, code, llr, max_iter, group=-1):
    n, N = llr.shape
    kb = (fec.kldpc + 7) // 8
    fr = np.zeros(n, ldpcgpu.FRAME_DTYPE)
    for f in range(n):
        fr[f] = (code, max_iter, fec.kldpc, group, f * N, f * kb)
    out, it, ok = dec.decode_host(fr, llr.reshape(-1), n * kb)
    return out.reshape(n, kb), it, ok


def test_single_code_64_per_frame(dec):
    """64 frames, DVB-S2 normal 1/2, QPSK, Es/N0 1.5 dB, max 50 iterations,
    each frame with its own early stop (the reference called per frame)."""
    fec, llr, ref = gen(0, "1_2", 1.5, 64, seed=2024, pls=0x10)
    code = ldpcgpu.code_lookup(0, "1_2")
    o_bits, o_it, o_ok = Oracle(0, "1_2", fec.kldpc).frames(llr, 50)
    g_bits, g_it, g_ok = run_frames(dec, fec, code, llr, 50)
    assert (g_it == o_it).all(), (g_it, o_it)
    assert (g_ok == o_ok).all()
    assert (g_bits == o_bits).all()
    assert o_ok.all() and (o_bits == ref).all()


def test_single_code_64_group(dec):
    """Same 64 frames as one group: the reference's batch call semantics."""
    fec, llr, ref = gen(0, "1_2", 1.5, 64, seed=2024, pls=0x10)
    code = ldpcgpu.code_lookup(0, "1_2")
    o_it, o_bits = Oracle(0, "1_2", fec.kldpc).group(llr, 50)
    g_bits, g_it, g_ok = run_frames(dec, fec, code, llr, 50, group=0)
    assert (g_it == o_it).all()
    assert g_ok.all() == (o_it >= 0)
    assert (g_bits == o_bits).all()


# Es/N0 near each code's threshold, so that frames take many iterations,
# some fail, and the dependent tails of the classic codes are exercised.
@pytest.mark.parametrize("short", [0, 1])
def test_all_codes_per_frame(dec, short):
    """Every rate of both frame sizes, 3 noisy frames each, max 25."""
    for rate in ALL_RATES[short]:
        fec, llr, ref = gen(short, rate, esn0_for(short, rate), 3, seed=77, stream0=1000)
        code = ldpcgpu.code_lookup(short, rate)
        o_bits, o_it, o_ok = Oracle(short, rate, fec.kldpc).frames(llr, 25)
        g_bits, g_it, g_ok = run_frames(dec, fec, code, llr, 25)
        assert (g_it == o_it).all(), (rate, g_it, o_it)
        assert (g_ok == o_ok).all(), rate
        assert (g_bits == o_bits).all(), rate


def test_mixed_batch_one_launch(dec):
    """Normal and short frames of many codes in one call, per-frame stop."""
    rng = np.random.default_rng(5)
    specs = []
    for f in range(48):
        short = int(rng.random() < 0.25)
        rate = ["1_4", "1_3", "2_5", "1_2", "3_5", "2_3", "3_4", "4_5", "5_6", "8_9"][rng.integers(10)]
        specs.append(dict(short_frame=short, rate=rate, esn0_db=esn0_for(short, rate) + 0.3, stream=f))
    llr, ref, layout = ldpcgpu.generate(specs, 99)
    fr = ldpcgpu.frames_for(layout, 40)
    out, it, ok = dec.decode_host(fr, llr, len(ref))
    for f, (code, n, kldpc, lo, ro) in enumerate(layout):
        short = int(n == 16200)
        rate = specs[f]["rate"]
        kb = (kldpc + 7) // 8
        o_bits, o_it, o_ok = Oracle(short, rate, kldpc).frames(llr[lo:lo + n][None], 40)
        assert it[f] == o_it[0] and ok[f] == o_ok[0], (f, rate)
        assert (out[ro:ro + kb] == o_bits[0]).all(), (f, rate)


def test_many_frames_refill(dec):
    """More frames than workgroups of a few codes, iteration counts spread
    wide: slots are refilled while their partner runs, workgroups move
    between code segments; every frame must still equal its own decode."""
    rng = np.random.default_rng(17)
    codes = [(0, "1_2"), (0, "3_4"), (1, "3_5"), (1, "1_4"), (0, "9_10"), (1, "8_9")]
    specs = []
    for f in range(700):
        short, rate = codes[rng.integers(len(codes))]
        specs.append(dict(short_frame=short, rate=rate, esn0_db=esn0_for(short, rate) + rng.uniform(-0.2, 0.8),
                          stream=10_000 + f))
    llr, ref, layout = ldpcgpu.generate(specs, 123)
    fr = ldpcgpu.frames_for(layout, 30)
    out, it, ok = dec.decode_host(fr, llr, len(ref))
    by = {}
    for f, (code, n, kldpc, lo, ro) in enumerate(layout):
        by.setdefault((int(n == 16200), specs[f]["rate"], kldpc), []).append(f)
    for (short, rate, kldpc), fl in by.items():
        n = 16200 if short else 64800
        kb = (kldpc + 7) // 8
        lls = np.stack([llr[layout[f][3]:layout[f][3] + n] for f in fl])
        o_bits, o_it, o_ok = Oracle(short, rate, kldpc).frames(lls, 30)
        for k, f in enumerate(fl):
            ro = layout[f][4]
            assert it[f] == o_it[k] and ok[f] == o_ok[k], (rate, f, it[f], o_it[k])
            assert (out[ro:ro + kb] == o_bits[k]).all(), (rate, f)
    assert len(set(it.tolist())) > 5  # the counts really spread


def test_groups_and_singles_in_one_call(dec):
    """Two groups (different codes) and per-frame frames mixed."""
    fa, la, _ = gen(0, "3_4", 3.95, 8, seed=3)
    fb, lb, _ = gen(1, "3_5", 2.5, 12, seed=4)
    ca, cb = ldpcgpu.code_lookup(0, "3_4"), ldpcgpu.code_lookup(1, "3_5")
    llr = np.concatenate([la.reshape(-1), lb.reshape(-1)])
    kba, kbb = fa.kldpc // 8, fb.kldpc // 8
    fr = np.zeros(20, ldpcgpu.FRAME_DTYPE)
    for f in range(8):
        fr[f] = (ca, 30, fa.kldpc, 7 if f < 6 else -1, f * 64800, f * kba)
    for f in range(12):
        fr[8 + f] = (cb, 20, fb.kldpc, 3 if f % 2 == 0 else -1, 8 * 64800 + f * 16200, 8 * kba + f * kbb)
    out, it, ok = dec.decode_host(fr, llr, 8 * kba + 12 * kbb)
    oa, ob = Oracle(0, "3_4", fa.kldpc), Oracle(1, "3_5", fb.kldpc)
    ga_it, ga_bits = oa.group(la[:6], 30)
    for f in range(6):
        assert it[f] == (ga_it if ga_it >= 0 else 30) and (out[f * kba:(f + 1) * kba] == ga_bits[f]).all()
    sa_bits, sa_it, sa_ok = oa.frames(la[6:], 30)
    for k, f in enumerate((6, 7)):
        assert it[f] == sa_it[k] and ok[f] == sa_ok[k] and (out[f * kba:(f + 1) * kba] == sa_bits[k]).all()
    ev = [f for f in range(12) if f % 2 == 0]
    gb_it, gb_bits = ob.group(lb[ev], 20)
    for k, f in enumerate(ev):
        o0 = 8 * kba + f * kbb
        assert it[8 + f] == (gb_it if gb_it >= 0 else 20) and (out[o0:o0 + kbb] == gb_bits[k]).all()
    od = [f for f in range(12) if f % 2]
    sb_bits, sb_it, sb_ok = ob.frames(lb[od], 20)
    for k, f in enumerate(od):
        o0 = 8 * kba + f * kbb
        assert it[8 + f] == sb_it[k] and ok[8 + f] == sb_ok[k] and (out[o0:o0 + kbb] == sb_bits[k]).all()


def test_special_inputs(dec):
    """All +127 (zero codeword): 0 iterations; all zero: never converges;
    a clean codeword: 0 iterations and its own bits; -128 everywhere."""
    fec, llr_clean, ref = gen(0, "2_3", 400.0, 1, seed=8)
    code = ldpcgpu.code_lookup(0, "2_3")
    N = 64800
    cases = np.stack([np.full(N, 127, np.int8), np.zeros(N, np.int8), llr_clean[0], np.full(N, -128, np.int8),
                      np.random.default_rng(1).integers(-128, 128, N).astype(np.int8)])
    o_bits, o_it, o_ok = Oracle(0, "2_3", fec.kldpc).frames(cases, 10)
    g_bits, g_it, g_ok = run_frames(dec, fec, code, cases, 10)
    assert (g_it == o_it).all() and (g_ok == o_ok).all() and (g_bits == o_bits).all()
    assert g_it[0] == 0 and g_ok[0] == 1 and not g_bits[0].any()
    assert g_it[1] == 10 and g_ok[1] == 0
    assert g_it[2] == 0 and g_ok[2] == 1 and (g_bits[2] == ref[0]).all()


def test_max_iter_zero_and_nbits(dec):
    fec, llr, ref = gen(1, "1_2", 2.0, 4, seed=9)
    code = ldpcgpu.code_lookup(1, "1_2")
    o_bits, o_it, o_ok = Oracle(1, "1_2", fec.kldpc).frames(llr, 0)
    g_bits, g_it, g_ok = run_frames(dec, fec, code, llr, 0)
    assert (g_it == 0).all() and (g_ok == o_ok).all() and (g_bits == o_bits).all()
    # fewer output bits than K (shortened codes): first nbits only
    fr = np.zeros(1, ldpcgpu.FRAME_DTYPE)
    fr[0] = (code, 20, 1001, -1, 0, 0)
    out, it, ok = dec.decode_host(fr, llr[0], 126)
    full, _, _ = Oracle(1, "1_2", fec.kldpc).frames(llr[:1], 20)
    bits = np.unpackbits(full[0])[:1001]
    assert (np.unpackbits(out)[:1001] == bits).all() and not np.unpackbits(out)[1001:].any()


def test_bad_arguments(dec):
    fr = np.zeros(1, ldpcgpu.FRAME_DTYPE)
    fr[0] = (999, 10, 0, -1, 0, 0)
    with pytest.raises(RuntimeError):
        dec.decode_host(fr, np.zeros(64800, np.int8), 8100)
    fr[0] = (ldpcgpu.code_lookup(0, "1_2"), 10, 0, -1, 4, 0)  # misaligned LLRs
    with pytest.raises(RuntimeError):
        dec.decode_host(fr, np.zeros(64808, np.int8), 8100)
    out, it, ok = dec.decode_host(np.zeros(0, ldpcgpu.FRAME_DTYPE), np.zeros(8, np.int8), 0)
    assert len(it) == 0


def test_drx_dropin_matches_reference_batch():
    """drx_ldpc_* of libldpcgpu.so behave as the reference's batch call."""
    L = ldpcgpu.gpu_lib()
    assert L.drx_ldpc_lanes() == 64
    fec, llr, ref = gen(1, "2_3", 3.0, 64, seed=12)
    d = ldpcgpu.DrxLdpc(fec, "2_3")
    try:
        it, bits = d.decode(llr, 30)
        o_it, o_bits = Oracle(1, "2_3", fec.kldpc).group(llr, 30)
        assert it == o_it and (bits == o_bits).all()
        it2, bits2 = d.decode(llr[:5], 30)
        o_it2, o_bits2 = Oracle(1, "2_3", fec.kldpc).group(llr[:5], 30)
        assert it2 == o_it2 and (bits2 == o_bits2).all()
        assert d.decode(np.zeros((65, 16200), np.int8), 10)[0] == -1
        assert d.decode(np.zeros((0, 16200), np.int8), 10)[0] == 0
    finally:
        d.close()


def test_reference_batch_directly_when_built():
    """Where the reference was built (oracle/_ref), the drop-in equals it."""
    refs = helpers.ref_builds()
    if not refs:
        pytest.skip("reference decoder not built here")
    fec, llr, ref = gen(0, "3_5", 2.3, 16, seed=13)
    d = ldpcgpu.DrxLdpc(fec, "3_5")
    try:
        it, bits = d.decode(llr, 50)
        for name, R in refs.items():
            lanes = R.drx_ldpc_lanes()
            if lanes < 16:
                continue
            r_it, r_bits = helpers.ref_decode(R, fec, "3_5", llr, 50)
            assert it == r_it and (bits == r_bits).all(), name
    finally:
        d.close()


@pytest.mark.parametrize("short,rate,esn0", [(0, "1_2", 1.05), (1, "3_4", 3.95), (0, "3_5", 2.35), (1, "1_4", -1.7)])
def test_many_groups_two_pass_against_oracle(short, rate, esn0):
    """40 groups of 6 frames near threshold (stops spread, some groups fail):
    the two-pass group path equals the oracle's lock-step decode of every
    group, and the exact one-frame path; how many calls fell back to the
    exact path is reported."""
    fec, llr, _ = gen(short, rate, esn0, 240, seed=77)
    code = ldpcgpu.code_lookup(short, rate)
    n, N = llr.shape
    kb = (fec.kldpc + 7) // 8
    fr = np.zeros(n, ldpcgpu.FRAME_DTYPE)
    for f in range(n):
        fr[f] = (code, 25, fec.kldpc, f // 6, f * N, f * kb)
    res = {}
    for kernel in ("auto", "single"):
        d = ldpcgpu.Decoder(0, kernel=kernel)
        try:
            res[kernel] = d.decode_host(fr, llr.reshape(-1), n * kb)
            print(f"# {rate}: {kernel} fallbacks {d.group_fallbacks()}", flush=True)
        finally:
            d.close()
    for a, b in zip(res["auto"], res["single"]):
        assert (a == b).all()
    out, it, ok = res["auto"]
    o = Oracle(short, rate, fec.kldpc)
    for g in range(0, n // 6, 4):  # every fourth group against the oracle
        o_it, o_bits = o.group(llr[6 * g:6 * g + 6], 25)
        for k in range(6):
            f = 6 * g + k
            assert it[f] == (o_it if o_it >= 0 else 25), (g, k)
            assert (out[f * kb:(f + 1) * kb] == o_bits[k]).all(), (g, k)
    assert len(set(int(x) for x in it[::6])) > 3, "group counts should spread"
