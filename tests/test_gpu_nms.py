"""GPU parity of the learned offset min-sum training kernels and the GPU
channel (include/ldpc_nms.h), against the CPU oracle.

* the training step's forward (integer, quantised parameters) gives the
  oracle's posteriors exactly; its loss and gradient agree with the
  oracle's float64 accumulation to float32 rounding;
* Adam on the GPU equals the textbook update;
* the GPU channel: deterministic per seed, unit-channel LLR statistics,
  noiseless hard decisions equal the codeword, and MODCOD frames at the
  quasi-error-free threshold + 1 dB decode.
"""
import os

import numpy as np
import pytest

from helpers import NMS_DEGS, Oracle, codewords, gen, ldpcgpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dec():
    d = ldpcgpu.Decoder(0)
    yield d
    d.close()


class Dev:
    """Device copies of host arrays through a decoder context."""

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


@pytest.mark.parametrize("short,rate,esn0,nf,T", [(1, "1_2", 1.0, 5, 4), (0, "3_5", 2.4, 2, 3), (1, "8_9", 6.0, 3, 3)])
def test_train_step_matches_oracle(dec, short, rate, esn0, nf, T):
    fec, llr, ref = gen(short, rate, esn0, nf, 500 + nf)
    cw = codewords(short, rate, ref)
    code = ldpcgpu.code_lookup(short, rate)
    rng = np.random.default_rng(nf)
    alpha = rng.uniform(0.6, 1.1, (T, NMS_DEGS)).astype(np.float32)
    beta = rng.uniform(0.0, 2.0, (T, NMS_DEGS)).astype(np.float32)
    w = np.linspace(0.2, 1.0, T).astype(np.float32)
    o = Oracle(short, rate, fec.kldpc)
    loss, ga, gb, post = o.nms_train(1, llr.astype(np.float32), cw, alpha, beta, w, temp=0.5)

    dv = Dev(dec)
    tr = ldpcgpu.NmsTrainer(code, T, nf)
    try:
        d_llr = dv.put(llr)
        d_cw = dv.put(cw[::-1].copy())  # rows reversed: frame f uses row nf-1-f
        d_row = dv.put(np.arange(nf - 1, -1, -1, dtype=np.int32))
        d_par = dv.put(np.concatenate([alpha, beta]))
        d_w = dv.put(w)
        d_loss = dv.empty(4)
        d_grad = dv.empty(4 * 2 * T * NMS_DEGS)
        d_post = dv.empty(llr.nbytes)
        tr.step(d_llr, d_cw, d_row, nf, d_par, d_w, 0.5, d_loss, d_grad, d_post)
        tr.synchronize()
        gpost = dv.get(d_post, llr.shape, np.int8)
        gloss = float(dv.get(d_loss, 1, np.float32)[0])
        grad = dv.get(d_grad, (2 * T, NMS_DEGS), np.float32)
    finally:
        tr.close()
        dv.free()
    assert (gpost.astype(np.float32) == post).all()
    assert abs(gloss - loss) <= 1e-4 * abs(loss)
    g_ref = np.concatenate([ga, gb])
    scale = np.abs(g_ref).max()
    assert scale > 0
    assert np.abs(grad - g_ref).max() <= 2e-3 * scale, np.abs(grad - g_ref).max() / scale
    assert ((grad != 0) == (g_ref != 0)).all()


def test_adam_matches_textbook(dec):
    rng = np.random.default_rng(9)
    n = 2 * 3 * NMS_DEGS
    p = (1.0 + 0.3 * rng.normal(size=n)).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    dv = Dev(dec)
    tr = ldpcgpu.NmsTrainer(ldpcgpu.code_lookup(1, "1_2"), 3, 1)
    try:
        d_p, d_m, d_v = dv.put(p), dv.put(m), dv.put(v)
        for step in range(1, 4):
            g = rng.normal(size=n).astype(np.float32)
            d_g = dv.put(g)
            tr.adam(d_p, d_g, d_m, d_v, n, 1e-2, step)
            tr.synchronize()
            m = 0.9 * m + 0.1 * g
            v = 0.999 * v + 0.001 * g * g
            p = p - 1e-2 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
            # projection: alpha (first half) into [0.25, 1.5], beta into [-8, 16]
            p[:n // 2] = np.clip(p[:n // 2], 0.25, 1.5)
            p[n // 2:] = np.clip(p[n // 2:], -8.0, 16.0)
        got = dv.get(d_p, n, np.float32)
    finally:
        tr.close()
        dv.free()
    assert np.allclose(got, p, rtol=1e-5, atol=1e-6)


def _channel(dec, pls, n, cw, esn0, seed):
    dv = Dev(dec)
    try:
        nf = len(esn0)
        d_cw = dv.put(cw)
        d_row = dv.put(np.arange(nf, dtype=np.int32) % cw.shape[0])
        d_e = dv.put(np.asarray(esn0, np.float32))
        d_llr = dv.empty(nf * n)
        ldpcgpu.channel(pls, n, d_cw, d_row, nf, d_e, seed, d_llr, stream=dec.stream)
        dec.synchronize()
        return dv.get(d_llr, (nf, n), np.int8)
    finally:
        dv.free()


def test_channel_unit_statistics_and_seed(dec):
    fec, _, ref = gen(0, "1_2", 1000.0, 2, 3)
    cw = codewords(0, "1_2", ref)
    esn0 = 2.0
    a = _channel(dec, -1, fec.nldpc, cw, [esn0, esn0], 11)
    b = _channel(dec, -1, fec.nldpc, cw, [esn0, esn0], 11)
    c = _channel(dec, -1, fec.nldpc, cw, [esn0, esn0], 12)
    assert (a == b).all() and (a != c).any()
    s2 = 1.0 / 10 ** (esn0 / 10)
    signed = a.astype(np.float64) * np.where(cw == 1, -1, 1)
    assert abs(signed.mean() - 4 / s2) < 0.03 * (4 / s2)
    # the spread of 4y/s2 is 4/s
    assert abs(signed.std() - 4 / np.sqrt(s2)) < 0.05 * (4 / np.sqrt(s2))


@pytest.mark.parametrize("pls", [0x10, 0x30, 0x48, 0x58, 0x5c])
def test_channel_noiseless_and_threshold(dec, pls):
    m = ldpcgpu.pls_lookup(pls)
    short, rate = m["short_frame"], m["rate"]
    fec, _, ref = gen(short, rate, 1000.0, 4, 21)
    cw = codewords(short, rate, ref)
    hi = _channel(dec, pls, fec.nldpc, cw, [40.0] * 4, 5)
    assert ((hi < 0) == (cw == 1)).all() and (hi != 0).all()
    llr = _channel(dec, pls, fec.nldpc, cw, [ldpcgpu.qef_db(pls) + 1.0] * 8, 6)
    o = Oracle(short, rate, fec.kldpc)
    bits, it, ok = o.frames(llr, 50)
    assert ok.all()
    assert (bits == np.concatenate([ref, ref])).all()


def _decode_nms(dec, llr, codes, S, B, learned_max, fallback, nbits):
    """Learned decode of frames (rows of llr, one code each) through the
    C-ABI: (bits rows, iters, ok, fell)."""
    n = llr.shape[1]
    kb = (nbits + 7) // 8
    frames = np.zeros(len(llr), ldpcgpu.FRAME_DTYPE)
    frames["code"] = codes
    frames["max_iter"] = learned_max
    frames["nbits"] = nbits
    frames["group"] = -1
    frames["llr_off"] = np.arange(len(llr)) * n
    frames["out_off"] = np.arange(len(llr)) * kb
    dv = Dev(dec)
    try:
        d_llr = dv.put(llr)
        d_out = dv.empty(len(llr) * kb)
        d_it = dv.empty(4 * len(llr))
        d_ok = dv.empty(len(llr))
        d_fb = dv.empty(len(llr))
        dec.decode_batch_nms(frames, d_llr, d_out, d_it, d_ok, fallback, d_fb)
        dec.synchronize()
        return (dv.get(d_out, (len(llr), kb), np.uint8), dv.get(d_it, len(llr), np.int32),
                dv.get(d_ok, len(llr), np.uint8), dv.get(d_fb, len(llr), np.uint8))
    finally:
        dv.free()


@pytest.mark.parametrize("short,rate,esn0", [(0, "3_5", 2.3), (0, "2_3", 3.1), (1, "1_2", 1.2), (0, "9_10", 6.4),
                                             (1, "8_9", 6.1), (0, "2_9_VLSNR", -2.0), (0, "13_45", -1.0)])
def test_learned_decode_matches_oracle(dec, short, rate, esn0):
    fec, llr, _ = gen(short, rate, esn0, 12, 700, stream0=short)
    code = ldpcgpu.code_lookup(short, rate)
    T = 6
    rng = np.random.default_rng(len(rate))
    S = rng.integers(9, 20, (T, NMS_DEGS)).astype(np.uint8)
    B = rng.integers(-16, 48, (T, NMS_DEGS)).astype(np.int16)
    o = Oracle(short, rate, fec.kldpc)
    dec.set_nms(code, S, B)
    try:
        for learned_max, fallback in ((T, 30), (3, 0), (9, 40)):
            want = o.nms_frames(S, B, llr, learned_max, fallback)
            got = _decode_nms(dec, llr, code, S, B, learned_max, fallback, fec.kldpc)
            for w, g, name in zip(want, got, ("bits", "iters", "ok", "fell")):
                assert (w == g).all(), (name, learned_max, fallback, w, g)
    finally:
        dec.set_nms(code, None, None)


def test_learned_decode_mixed_batch_refill(dec):
    """Several codes in one call, many frames per workgroup slot, fallback
    in the middle of refills: every frame as the oracle decodes it."""
    sets = [(0, "3_5", 2.7), (0, "2_3", 3.5), (1, "1_2", 1.6)]
    tabs = {}
    rows = []
    rng = np.random.default_rng(77)
    for k, (short, rate, esn0) in enumerate(sets):
        fec, llr, _ = gen(short, rate, esn0, 40, 800 + k)
        code = ldpcgpu.code_lookup(short, rate)
        T = 5
        S = rng.integers(13, 17, (T, NMS_DEGS)).astype(np.uint8)
        B = rng.integers(0, 32, (T, NMS_DEGS)).astype(np.int16)
        tabs[code] = (S, B, fec, llr, Oracle(short, rate, fec.kldpc))
        dec.set_nms(code, S, B)
    try:
        # one buffer: frames of all codes interleaved, N-byte slots of 64800
        order = [(code, f) for f in range(40) for code in tabs]
        n = 64800
        buf = np.zeros((len(order), n), np.int8)
        for r, (code, f) in enumerate(order):
            llr = tabs[code][3][f]
            buf[r, :len(llr)] = llr
        frames = np.zeros(len(order), ldpcgpu.FRAME_DTYPE)
        kb = max((t[2].kldpc + 7) // 8 for t in tabs.values())
        for r, (code, f) in enumerate(order):
            frames[r] = (code, 12, tabs[code][2].kldpc, -1, r * n, r * kb)
        dv = Dev(dec)
        try:
            d_llr = dv.put(buf)
            d_out = dv.empty(len(order) * kb)
            d_it = dv.empty(4 * len(order))
            d_ok = dv.empty(len(order))
            d_fb = dv.empty(len(order))
            dec.decode_batch_nms(frames, d_llr, d_out, d_it, d_ok, 30, d_fb)
            dec.synchronize()
            out = dv.get(d_out, (len(order), kb), np.uint8)
            it = dv.get(d_it, len(order), np.int32)
            ok = dv.get(d_ok, len(order), np.uint8)
            fb = dv.get(d_fb, len(order), np.uint8)
        finally:
            dv.free()
        assert fb.any() and not fb.all()
        for code, (S, B, fec, llr, o) in tabs.items():
            wb, wit, wok, wfb = o.nms_frames(S, B, llr, 12, 30)
            rows_c = [r for r, (c, f) in enumerate(order) if c == code]
            nb = (fec.kldpc + 7) // 8
            assert (out[rows_c, :nb] == wb).all()
            assert (it[rows_c] == wit).all() and (ok[rows_c] == wok).all() and (fb[rows_c] == wfb).all()
    finally:
        for code in tabs:
            dec.set_nms(code, None, None)


def test_learned_decode_rejects_bad_calls(dec):
    code = ldpcgpu.code_lookup(1, "1_2")
    fr = np.zeros(1, ldpcgpu.FRAME_DTYPE)
    fr["code"] = code
    fr["max_iter"] = 5
    fr["group"] = -1
    with pytest.raises(RuntimeError):  # no table for the code
        dec.decode_batch_nms(fr, 1, 1, 1, 1, 10)
    S = np.full((3, NMS_DEGS), 16, np.uint8)
    B = np.full((3, NMS_DEGS), 16, np.int16)
    dec.set_nms(code, S, B)
    try:
        fr["group"] = 0
        with pytest.raises(RuntimeError):  # groups are not learned-decoded
            dec.decode_batch_nms(fr, 1, 1, 1, 1, 10)
    finally:
        dec.set_nms(code, None, None)


TABLES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dvb-rx_amd", "nms")


@pytest.mark.parametrize("name,dB", [("s2_n3_5_8psk", 0.2), ("s2_n2_3_16apsk", 0.2)])
def test_trained_table_gain(dec, name, dB):
    """What the trained table buys over the plain decoder, on the same
    frames (GPU channel, Es/N0 = QEF + 0.2 dB, 4096 frames):
    * at an equal budget of 15 iterations (no fallback) it leaves fewer
      frames unconverged;
    * as deployed (up to 30 learned iterations, plain fallback) it needs
      fewer iterations on average, with no more frame errors after the
      BCH code's correction (> t bit errors)."""
    import json

    from nms_train import codeword_pool

    table = json.load(open(os.path.join(TABLES, name + ".json")))
    pls, short, rate = table["pls"], table["short_frame"], table["rate"]
    code = ldpcgpu.code_lookup(short, rate)
    fec = ldpcgpu.fec_lookup(short, rate)
    N, kl = fec.nldpc, fec.kldpc
    kb = (kl + 7) // 8
    nf = 4096
    pool = codeword_pool(short, rate, 256, 5)
    ref = np.packbits(pool[:, :kl], axis=1)
    rows = np.random.default_rng(6).integers(0, 256, nf).astype(np.int32)
    dec.set_nms(code, np.array(table["S"], np.uint8), np.array(table["B"], np.int16))
    dv = Dev(dec)
    try:
        d_pool = dv.put(pool)
        d_rows = dv.put(rows)
        d_e = dv.put(np.full(nf, ldpcgpu.qef_db(pls) + dB, np.float32))
        d_llr = dv.empty(nf * N)
        ldpcgpu.channel(pls, N, d_pool, d_rows, nf, d_e, 31, d_llr, stream=dec.stream)
        dec.synchronize()
        frames = np.zeros(nf, ldpcgpu.FRAME_DTYPE)
        frames["code"] = code
        frames["nbits"] = kl
        frames["group"] = -1
        frames["llr_off"] = np.arange(nf) * N
        frames["out_off"] = np.arange(nf) * kb
        d_out, d_it, d_ok, d_fb = dv.empty(nf * kb), dv.empty(4 * nf), dv.empty(nf), dv.empty(nf)

        def run(learned, max_iter, fallback):
            fr = frames.copy()
            fr["max_iter"] = max_iter
            if learned:
                dec.decode_batch_nms(fr, d_llr, d_out, d_it, d_ok, fallback, d_fb)
            else:
                dec.decode_batch(fr, d_llr, d_out, d_it, d_ok)
            dec.synchronize()
            bits = dv.get(d_out, (nf, kb), np.uint8)
            errs = np.unpackbits(bits ^ ref[rows], axis=1)[:, :kl].sum(axis=1)
            return dv.get(d_it, nf, np.int32), dv.get(d_ok, nf, np.uint8), errs

        p_it, p_ok, _ = run(False, 15, 0)
        l_it, l_ok, _ = run(True, 15, 0)
        assert (l_ok == 0).sum() < (p_ok == 0).sum(), ((l_ok == 0).sum(), (p_ok == 0).sum())
        p_it, p_ok, p_err = run(False, 50, 0)
        l_it, l_ok, l_err = run(True, 30, 50)
        assert l_it.mean() < 0.97 * p_it.mean(), (l_it.mean(), p_it.mean())
        assert (l_err > fec.t).sum() <= (p_err > fec.t).sum()
    finally:
        dv.free()
        dec.set_nms(code, None, None)
