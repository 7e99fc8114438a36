"""Test-side wrappers: the CPU oracle (oracle/liboracle.so), the reference
decoder compiled from its own sources (oracle/_ref/, present where it was
built), the CPU model of the kernel schedule (tests/model/libkmodel.so), and
the product's bindings (dvb-rx_amd/ldpcgpu.py)."""
import ctypes as C
import glob
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "dvb-rx_amd"))

import ldpcgpu  # noqa: E402

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
KMODEL_SO = os.path.join(ROOT, "tests", "model", "libkmodel.so")

ALL_RATES = {
    0: ["1_4", "1_3", "2_5", "1_2", "3_5", "2_3", "3_4", "4_5", "5_6", "8_9", "9_10", "2_9_VLSNR", "13_45",
        "9_20", "90_180", "96_180", "11_20", "100_180", "104_180", "26_45", "18_30", "28_45", "23_36",
        "116_180", "20_30", "124_180", "25_36", "128_180", "13_18", "132_180", "22_30", "135_180",
        "140_180", "7_9", "154_180", "2_3_t2"],
    1: ["1_4", "1_3", "2_5", "1_2", "3_5", "2_3", "3_4", "4_5", "5_6", "8_9", "11_45", "4_15", "14_45",
        "7_15", "8_15", "26_45", "32_45", "1_5_VLSNR_SF2", "11_45_VLSNR_SF2", "1_5_VLSNR", "4_15_VLSNR",
        "1_3_VLSNR", "3_5_t2"],
}

ESN0 = {"1_4": -2.0, "1_3": -1.1, "2_5": -0.2, "1_2": 1.0, "3_5": 2.2, "2_3": 3.0, "3_4": 4.0, "4_5": 4.6,
        "5_6": 5.2, "8_9": 6.2, "9_10": 6.4}


def esn0_for(short, rate):
    """Es/N0 (dB) of the unit channel near a code's waterfall."""
    if rate in ESN0:
        return ESN0[rate] + (0.3 if short else 0.0)
    num, den = [int(x) for x in rate.replace("_VLSNR", "").replace("_SF2", "").replace("_t2", "").split("_")]
    r = num / den
    # rough BPSK-per-component threshold of the unit channel
    return float(10 * np.log10(2 ** (2 * r) - 1) + 0.4)


_oracle = None
NMS_DEGS = 33


def oracle():
    global _oracle
    if _oracle is None:
        L = C.CDLL(ORACLE_SO)
        vp = C.c_void_p
        L.oracle_ldpc_new.restype = vp
        L.oracle_ldpc_new.argtypes = [C.c_int, C.c_char_p, C.c_int]
        L.oracle_ldpc_free.argtypes = [vp]
        L.oracle_ldpc_decode.argtypes = [vp, vp, vp, C.c_int, C.c_int]
        L.oracle_ldpc_decode_frames.argtypes = [vp, vp, vp, C.c_int, C.c_int, vp, vp]
        L.oracle_ldpc_check.argtypes = [vp, vp]
        L.oracle_check_node.argtypes = [C.c_int, vp, vp, C.c_int, vp, vp]
        L.oracle_ldpc_run.argtypes = [vp, vp, C.c_int, C.c_int, vp]
        L.oracle_nms_scale_q.argtypes = [C.c_float]
        L.oracle_nms_offset_q.argtypes = [C.c_float]
        L.oracle_nms_check_node.argtypes = [C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, vp, vp]
        L.oracle_nms_decode_frames.argtypes = [vp, vp, vp, C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp]
        L.oracle_nms_train.argtypes = [vp, C.c_int, vp, vp, C.c_int, C.c_int, vp, vp, vp, C.c_float, vp, vp, vp]
        L.oracle_nms_train.restype = C.c_double
        L.oracle_demap_symbols.argtypes = [C.c_int]
        L.oracle_demap_ring.argtypes = [C.c_int]
        L.oracle_demap_points.argtypes = [C.c_int, vp, vp]
        L.oracle_demap_frame.argtypes = [C.c_int, vp, C.c_float, vp]
        L.oracle_bch_nbch.argtypes = [C.c_int, C.c_char_p]
        L.oracle_bch_decode.argtypes = [C.c_int, C.c_char_p, vp]
        ip = C.POINTER(C.c_int)
        L.oracle_t2_geometry.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip, C.c_char_p, C.c_int]
        L.oracle_t2_cellperm.argtypes = [C.c_int, vp]
        L.oracle_t2_shift.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp]
        L.oracle_t2_bitsrc.argtypes = [C.c_int, C.c_int, C.c_int, vp]
        L.oracle_t2_demap.argtypes = [C.c_int] * 7 + [vp] * 5
        _oracle = L
    return _oracle


def demap_pls_codes():
    """Every PLS code (pilot bit clear) with a constellation."""
    out = []
    for pls in range(0, 0x100, 2):
        if oracle().oracle_demap_symbols(pls) > 0:
            out.append(pls)
    return out


def demap_points(pls):
    """Unit-energy constellation of a MODCOD by label: complex array."""
    re = np.zeros(256, np.float32)
    im = np.zeros(256, np.float32)
    n = oracle().oracle_demap_points(pls, re.ctypes.data, im.ctypes.data)
    assert n > 0, hex(pls)
    return re[:n] + 1j * im[:n]


def demap_order(pls):
    """The bit interleaver's column per label bit (MSB first), read back
    from the oracle demapper: moving symbol 0 from the point labelled 0 to
    the one whose label has bit b alone set flips the sign of one LLR, at
    order[b] * nsym; None for QPSK (no interleaver: bits 2j, 2j+1)."""
    if ldpcgpu.pls_lookup(pls)["constellation"] == "QPSK":
        return None
    pts = demap_points(pls)
    nb = int(np.log2(len(pts)))
    ns = oracle().oracle_demap_symbols(pls)
    base = np.zeros((ns, 2), np.float32)
    base[:] = [pts[0].real, pts[0].imag]
    ref = oracle_demap(pls, base, 1e-3)
    order = []
    for b in range(nb):
        s = base.copy()
        p = pts[1 << (nb - 1 - b)]
        s[0] = [p.real, p.imag]
        diff = np.nonzero((oracle_demap(pls, s, 1e-3) < 0) != (ref < 0))[0]
        flips = [int(d) for d in diff if d % ns == 0]
        assert len(flips) == 1, (hex(pls), b)
        order.append(flips[0] // ns)
    return order


def modulate(pls, cw, esn0_db, rng):
    """Codeword bits (N,) -> noisy unit-energy symbols (nsym, 2) float32 and
    the noise variance: label bit b of symbol j is codeword bit
    order[b] * nsym + j (zero past the codeword), AWGN of variance n0."""
    pts = demap_points(pls)
    nb = int(np.log2(len(pts)))
    ns = oracle().oracle_demap_symbols(pls)
    order = demap_order(pls)
    N = len(cw)
    lab = np.zeros(ns, np.int64)
    for b in range(nb):
        pos = (2 * np.arange(ns) + b) if order is None else order[b] * ns + np.arange(ns)
        bit = np.where(pos < N, cw[np.minimum(pos, N - 1)], 0)
        lab = lab * 2 + bit
    n0 = 10 ** (-esn0_db / 10)
    y = pts[lab] + np.sqrt(n0 / 2) * (rng.normal(size=ns) + 1j * rng.normal(size=ns))
    return np.stack([y.real, y.imag], 1).astype(np.float32), np.float32(n0)


def oracle_demap(pls, sym, n0):
    """Oracle LLRs of one frame: sym (nsym, 2) float32 (re, im)."""
    sym = np.ascontiguousarray(sym, np.float32)
    nldpc = 16200 if ldpcgpu.pls_lookup(pls)["short_frame"] else 64800
    llr = np.zeros(nldpc, np.int8)
    assert oracle().oracle_demap_frame(pls, sym.ctypes.data, np.float32(n0), llr.ctypes.data) == 0
    return llr


def ref_demap(R, pls, sym, n0):
    """drx_demap_frame of a reference build on one frame."""
    m = ldpcgpu.pls_lookup(pls)
    R.drx_demap_new.restype = C.c_void_p
    R.drx_demap_new.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    R.drx_demap_frame.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_void_p]
    R.drx_demap_symbols.argtypes = [C.c_void_p]
    R.drx_demap_free.argtypes = [C.c_void_p]
    d = R.drx_demap_new(m["constellation"].encode(), m["rate"].encode(), m["short_frame"])
    assert d, hex(pls)
    try:
        sym = np.ascontiguousarray(sym, np.float32)
        assert R.drx_demap_symbols(d) == len(sym)
        llr = np.zeros(16200 if m["short_frame"] else 64800, np.int8)
        R.drx_demap_frame(d, sym.ctypes.data, np.float32(n0), llr.ctypes.data)
        return llr
    finally:
        R.drx_demap_free(d)


class Oracle:
    def __init__(self, short_frame, rate, kldpc):
        self.L = oracle()
        self.h = self.L.oracle_ldpc_new(short_frame, rate.encode(), kldpc)
        assert self.h, rate
        self.kb = (kldpc + 7) // 8

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oracle_ldpc_free(self.h)

    def frames(self, llr, max_iter):
        """Per-frame decode: (bits (n, kb), iters, ok)."""
        llr = np.ascontiguousarray(llr, np.int8)
        n = llr.shape[0]
        out = np.zeros((n, self.kb), np.uint8)
        it = np.zeros(n, np.int32)
        ok = np.zeros(n, np.uint8)
        self.L.oracle_ldpc_decode_frames(self.h, llr.ctypes.data, out.ctypes.data, n, max_iter, it.ctypes.data, ok.ctypes.data)
        return out, it, ok

    def group(self, llr, max_iter):
        llr = np.ascontiguousarray(llr, np.int8)
        n = llr.shape[0]
        out = np.zeros((n, self.kb), np.uint8)
        r = self.L.oracle_ldpc_decode(self.h, llr.ctypes.data, out.ctypes.data, n, max_iter)
        return r, out

    def run(self, llr, T):
        """T plain updates, no early stop: posteriors (n, N) in codeword order."""
        llr = np.ascontiguousarray(llr, np.int8)
        n = llr.shape[0]
        post = np.zeros_like(llr)
        self.L.oracle_ldpc_run(self.h, llr.ctypes.data, n, T, post.ctypes.data)
        return post

    def nms_frames(self, S, B, llr, learned_max, fallback_max):
        """Learned decode per frame: (bits, iters, ok, fell_back)."""
        S = np.ascontiguousarray(S, np.uint8)
        B = np.ascontiguousarray(B, np.int16)
        assert S.shape == B.shape and S.shape[1] == NMS_DEGS
        llr = np.ascontiguousarray(llr, np.int8)
        n = llr.shape[0]
        out = np.zeros((n, self.kb), np.uint8)
        it = np.zeros(n, np.int32)
        ok = np.zeros(n, np.uint8)
        fb = np.zeros(n, np.uint8)
        self.L.oracle_nms_decode_frames(self.h, S.ctypes.data, B.ctypes.data, S.shape[0], llr.ctypes.data,
                                        out.ctypes.data, n, learned_max, fallback_max, it.ctypes.data,
                                        ok.ctypes.data, fb.ctypes.data)
        return out, it, ok, fb

    def nms_train(self, mode, llr, bits, alpha, beta, w, temp=0.5, grad=True):
        """Float model over T = alpha.shape[0] iterations: (loss, galpha,
        gbeta, final posteriors)."""
        llr = np.ascontiguousarray(llr, np.float32)
        bits = np.ascontiguousarray(bits, np.uint8)
        alpha = np.ascontiguousarray(alpha, np.float32)
        beta = np.ascontiguousarray(beta, np.float32)
        w = np.ascontiguousarray(w, np.float32)
        n = llr.shape[0]
        T = alpha.shape[0]
        ga = np.zeros_like(alpha)
        gb = np.zeros_like(beta)
        post = np.zeros_like(llr)
        loss = self.L.oracle_nms_train(self.h, mode, llr.ctypes.data, bits.ctypes.data, n, T, alpha.ctypes.data,
                                       beta.ctypes.data, w.ctypes.data, temp, ga.ctypes.data if grad else None,
                                       gb.ctypes.data if grad else None, post.ctypes.data)
        return loss, ga, gb, post

    def check(self, llr):
        llr = np.ascontiguousarray(llr, np.int8)
        return self.L.oracle_ldpc_check(self.h, llr.ctypes.data)


def check_node(deg, L, m, first):
    Lb = np.ascontiguousarray(L, np.int8)
    mb = np.ascontiguousarray(m, np.int8)
    Lo = np.zeros(deg, np.int8)
    mo = np.zeros(deg, np.int8)
    oracle().oracle_check_node(deg, Lb.ctypes.data, mb.ctypes.data, int(first), Lo.ctypes.data, mo.ctypes.data)
    return Lo, mo


def ref_builds():
    """Reference decoders built from the reference's own sources by
    oracle/Makefile (name -> CDLL); empty where they were not built.  A build
    whose SIMD level the host lacks is skipped."""
    flags = ""
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        pass
    out = {}
    for path in sorted(glob.glob(os.path.join(REF_DIR, "libdrxref_*.so"))):
        name = os.path.basename(path)[len("libdrxref_"):-3]
        if name == "v4" and "avx512bw" not in flags:
            continue
        if name == "v3" and "avx2" not in flags:
            continue
        L = C.CDLL(path)
        vp = C.c_void_p
        L.drx_ldpc_new.restype = vp
        L.drx_ldpc_new.argtypes = [vp, C.c_char_p]
        L.drx_ldpc_free.argtypes = [vp]
        L.drx_ldpc_decode.argtypes = [vp, vp, vp, C.c_int, C.c_int]
        L.drx_ldpc_lanes.restype = C.c_int
        L.drx_bch_new.restype = vp
        L.drx_bch_new.argtypes = [vp]
        L.drx_bch_free.argtypes = [vp]
        L.drx_bch_decode.argtypes = [vp, vp]
        out[name] = L
    return out


def ref_decode(L, fec, rate, llr, max_iter):
    """drx_ldpc_decode of the reference on up to its lane count of frames."""
    p = fec.as_struct()
    d = L.drx_ldpc_new(C.byref(p), rate.encode())
    assert d, rate
    try:
        llr = np.ascontiguousarray(llr, np.int8)
        n = llr.shape[0]
        kb = (fec.kldpc + 7) // 8
        out = np.zeros((n, kb), np.uint8)
        ins = (C.c_void_p * n)(*[llr[f].ctypes.data for f in range(n)])
        outs = (C.c_void_p * n)(*[out[f].ctypes.data for f in range(n)])
        it = L.drx_ldpc_decode(d, ins, outs, n, max_iter)
        return it, out
    finally:
        L.drx_ldpc_free(d)


_km = None


def kmodel():
    global _km
    if _km is None:
        L = C.CDLL(KMODEL_SO)
        L.kmodel_decode.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.kmodel_check_levels.argtypes = [C.c_int]
        _km = L
    return _km


def kmodel_decode(code, llr, max_iter, nbits):
    llr = np.ascontiguousarray(llr, np.int8)
    out = np.zeros((nbits + 7) // 8, np.uint8)
    ok = C.c_int()
    it = kmodel().kmodel_decode(code, llr.ctypes.data, max_iter, nbits, out.ctypes.data, C.byref(ok))
    return out, it, ok.value


def gen(short_frame, rate, esn0_db, nframes, seed, pls=-1, stream0=0):
    """nframes frames of one rate: (fec, llr (n, N), ref bits (n, kb))."""
    spec = [dict(short_frame=short_frame, rate=rate, pls=pls, esn0_db=esn0_db, stream=stream0 + f) for f in range(nframes)]
    llr, ref, layout = ldpcgpu.generate(spec, seed)
    if pls < 0:
        fec = ldpcgpu.fec_lookup(short_frame, rate)
    else:
        m = ldpcgpu.pls_lookup(pls)
        fec = ldpcgpu.fec_lookup(m["short_frame"], m["rate"])
    n = layout[0][1]
    kb = (fec.kldpc + 7) // 8
    return fec, llr.reshape(nframes, n), ref.reshape(nframes, kb)


def sha256(a):
    import hashlib

    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def codewords(short_frame, rate, ref):
    """LDPC codewords (n, N) of unpacked bits from the BCH codewords gen()
    returns (kldpc bits packed; kldpc = K for every rate but the
    shortened VLSNR codes, which are not used here)."""
    L = ldpcgpu.sim_lib()
    fec = ldpcgpu.fec_lookup(short_frame, rate)
    out = np.zeros((ref.shape[0], fec.nldpc), np.uint8)
    for f in range(ref.shape[0]):
        info = np.ascontiguousarray(np.unpackbits(ref[f])[:fec.kldpc], np.uint8)
        assert L.sim_ldpc_encode(short_frame, rate.encode(), info.ctypes.data, out[f].ctypes.data) == 0
    return out


def oracle_bch(short_frame, rate, cw):
    """Oracle BCH decode of one packed codeword (copy): (result, bytes)."""
    cw = np.array(cw, np.uint8, copy=True)
    r = oracle().oracle_bch_decode(short_frame, rate.encode(), cw.ctypes.data)
    return r, cw


def ref_bch(R, fec, rate, cw):
    """drx_bch_decode of a reference build on one packed codeword (copy)."""
    b = R.drx_bch_new(C.byref(fec.as_struct()))
    assert b, rate
    try:
        cw = np.array(cw, np.uint8, copy=True)
        return R.drx_bch_decode(b, cw.ctypes.data), cw
    finally:
        R.drx_bch_free(b)


def flip(cw, bits):
    """Packed codeword with the given bit indexes (MSB first) flipped."""
    out = np.array(cw, np.uint8, copy=True)
    for bit in bits:
        out[bit // 8] ^= 0x80 >> (bit % 8)
    return out


# ---- DVB-T2 FEC blocks -------------------------------------------------

T2_RATES = ["1_2", "3_5", "2_3", "3_4", "4_5", "5_6"]


def t2_geometry(cod, mod, fec_type):
    """(N, ncells, LDPC code name) of a PLP configuration (oracle)."""
    n, nc = C.c_int(), C.c_int()
    name = C.create_string_buffer(16)
    assert oracle().oracle_t2_geometry(cod, mod, fec_type, C.byref(n), C.byref(nc), name, 16) == 0
    return n.value, nc.value, name.value.decode()


def oracle_t2_demap(cod, mod, fec_type, rot, til, cells, csi):
    """Oracle front end: (llr (nblocks, N), err, sig)."""
    N, nc, _ = t2_geometry(cod, mod, fec_type)
    nb = cells.shape[0]
    llr = np.zeros((nb, N), np.int8)
    err = np.zeros(nb, np.float32)
    sig = np.zeros(nb, np.float32)
    cells = np.ascontiguousarray(cells, np.float32)
    csi = np.ascontiguousarray(csi, np.float32)
    assert oracle().oracle_t2_demap(cod, mod, fec_type, rot, til[0], til[1], nb, cells.ctypes.data,
                                    csi.ctypes.data, llr.ctypes.data, err.ctypes.data, sig.ctypes.data) == 0
    return llr, err, sig


def oracle_t2_job(cod, mod, fec_type, rot, til, cells, csi, max_iter, lanes):
    """The reference's FEC job restated with the oracle: front end, LDPC in
    calls of `lanes` blocks, BCH.  (codewords (nblocks, kldpc/8) after BCH,
    BCH results, err, sig)."""
    llr, err, sig = oracle_t2_demap(cod, mod, fec_type, rot, til, cells, csi)
    sf = fec_type == 0
    _, _, code = t2_geometry(cod, mod, fec_type)
    fec = ldpcgpu.fec_lookup(sf, T2_RATES[cod])
    d = oracle().oracle_ldpc_new(sf, code.encode(), fec.kldpc)
    assert d, code
    nb = llr.shape[0]
    kb = fec.kldpc // 8
    cw = np.zeros((nb, kb), np.uint8)
    try:
        for b0 in range(0, nb, lanes):
            part = np.ascontiguousarray(llr[b0:b0 + lanes])
            out = np.zeros((part.shape[0], kb), np.uint8)
            oracle().oracle_ldpc_decode(d, part.ctypes.data, out.ctypes.data, part.shape[0], max_iter)
            cw[b0:b0 + part.shape[0]] = out
    finally:
        oracle().oracle_ldpc_free(d)
    res = np.zeros(nb, np.int32)
    for b in range(nb):
        res[b], cw[b] = oracle_bch(sf, T2_RATES[cod], cw[b])
    return cw, res, err, sig


def t2_ref_builds():
    """The reference's T2 FEC job (oracle/_ref/libdrxt2ref_*.so), name ->
    (CDLL, LDPC lanes); empty where not built or the host lacks the SIMD
    level."""
    flags = ""
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        pass
    out = {}
    for name, need, lanes in (("v3", "avx2", 32), ("v4", "avx512bw", 64)):
        path = os.path.join(REF_DIR, f"libdrxt2ref_{name}.so")
        if not os.path.exists(path) or need not in flags:
            continue
        L = C.CDLL(path)
        L.t2ref_decode.argtypes = [C.c_int] * 7 + [C.c_void_p] * 2 + [C.c_int] + [C.c_void_p] * 5
        out[name] = (L, lanes)
    return out


def ref_t2_job(L, cod, mod, fec_type, rot, til, cells, csi, max_iter, with_llr=False):
    """decode_job of a reference build: (bb (nblocks, kbch/8), ok, err, sig,
    llr or None)."""
    N, nc, _ = t2_geometry(cod, mod, fec_type)
    fec = ldpcgpu.fec_lookup(fec_type == 0, T2_RATES[cod])
    nb = cells.shape[0]
    kb = fec.kbch // 8
    bb = np.zeros((nb, kb), np.uint8)
    ok = np.zeros(nb, np.uint8)
    err = np.zeros(nb, np.float32)
    sig = np.zeros(nb, np.float32)
    llr = np.zeros((nb, N), np.int8) if with_llr else None
    cells = np.ascontiguousarray(cells, np.float32)
    csi = np.ascontiguousarray(csi, np.float32)
    rc = L.t2ref_decode(cod, mod, fec_type, rot, til[0], til[1], nb, cells.ctypes.data, csi.ctypes.data, max_iter,
                        bb.ctypes.data, ok.ctypes.data, err.ctypes.data, sig.ctypes.data,
                        llr.ctypes.data if with_llr else None)
    assert rc == 0
    return bb, ok, err, sig, llr


def t2_random_cells(mod, ncells, nblocks, rng, snr_db=15.0):
    """Random constellation points (any labels) plus noise, and random
    channel state: exercises every soft-value branch."""
    m = 2 * (mod + 1)
    L = 2 ** (m // 2)
    lev = np.arange(-(L - 1), L, 2)
    norm = np.sqrt({2: 2.0, 4: 10.0, 6: 42.0, 8: 170.0}[m])
    x = (rng.choice(lev, (nblocks, ncells)) + 1j * rng.choice(lev, (nblocks, ncells))) / norm
    s = 10 ** (-snr_db / 20) / np.sqrt(2)
    x = x + s * (rng.standard_normal(x.shape) + 1j * rng.standard_normal(x.shape))
    cells = np.stack([x.real, x.imag], -1).astype(np.float32)
    csi = rng.uniform(0.2, 2.0, (nblocks, ncells)).astype(np.float32)
    return cells, csi


T2_ROT_DEG = [29.0, 16.8, 8.6, 3.576334375]


def t2_modulate(cod, mod, fec_type, rot, til, cw_bits, snr_db, rng):
    """Transmit side of the front end, for round trips: LDPC codewords
    (nblocks, N) of bits -> cells and channel state.  Each codeword bit goes
    to the (cell, plane) oracle_t2_bitsrc names; the planes of an axis are
    the Gray-coded PAM level whose max-log values have those signs (bit 0 the
    sign, bit j > 0 the sign of |u| - 2^(nb-j)); the point is rotated and its
    Q sent one cell later (cyclically); the cell interleaver puts cell i at
    (L0(i) + P) mod ncells; AWGN at snr_db."""
    N, nc, _ = t2_geometry(cod, mod, fec_type)
    m = 2 * (mod + 1)
    nb = m // 2
    src = np.zeros(N, np.uint16)
    assert oracle().oracle_t2_bitsrc(cod, mod, fec_type, src.ctypes.data) == 0
    perm = np.zeros(nc, np.int32)
    assert oracle().oracle_t2_cellperm(nc, perm.ctypes.data) == 0
    nblocks = cw_bits.shape[0]
    P = np.zeros(nblocks, np.int32)
    oracle().oracle_t2_shift(nblocks, til[0], til[1], nc, P.ctypes.data)
    norm = np.sqrt({2: 2.0, 4: 10.0, 6: 42.0, 8: 170.0}[m])
    th = np.deg2rad(T2_ROT_DEG[mod]) if rot else 0.0
    cells = np.zeros((nblocks, nc, 2), np.float32)
    for b in range(nblocks):
        planes = np.zeros(nc * m, np.uint8)
        planes[src.astype(np.int64)] = cw_bits[b]
        planes = planes.reshape(nc, m)
        axis = []
        for a in (0, 1):
            bits = planes[:, a::2]
            u = np.where(bits[:, nb - 1] == 0, 1.0, -1.0)
            for j in range(nb - 1, 0, -1):
                u = (u + 2.0 ** (nb - j)) * np.where(bits[:, j - 1] == 0, 1.0, -1.0)
            axis.append(u)
        x = (axis[0] + 1j * axis[1]) / norm * np.exp(1j * th)
        tx_i = x.real
        tx_q = np.roll(x.imag, 1) if rot else x.imag
        y = tx_i + 1j * tx_q
        s = 10 ** (-snr_db / 20) / np.sqrt(2)
        y = y + s * (rng.standard_normal(nc) + 1j * rng.standard_normal(nc))
        pos = (perm.astype(np.int64) + P[b]) % nc
        cells[b, pos, 0] = y.real
        cells[b, pos, 1] = y.imag
    csi = np.ones((nblocks, nc), np.float32)
    return cells, csi


def t2_codewords(cod, fec_type, nblocks, rng):
    """Random BBFRAMEs through the BCH and LDPC encoders: (info bytes
    (nblocks, kbch/8), LDPC codeword bits (nblocks, N))."""
    sf = fec_type == 0
    rate = T2_RATES[cod]
    fec = ldpcgpu.fec_lookup(sf, rate)
    _, _, code = t2_geometry(cod, 0, fec_type)
    L = ldpcgpu.sim_lib()
    info = rng.integers(0, 2, (nblocks, fec.kbch), dtype=np.uint8)
    bits = np.zeros((nblocks, fec.nldpc), np.uint8)
    for b in range(nblocks):
        bch = np.zeros(fec.kldpc, np.uint8)
        assert L.sim_bch_encode(sf, rate.encode(), np.ascontiguousarray(info[b]).ctypes.data, bch.ctypes.data) == 0
        assert L.sim_ldpc_encode(sf, code.encode(), bch.ctypes.data, bits[b].ctypes.data) == 0
    return np.packbits(info, axis=1), bits
