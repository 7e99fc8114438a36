"""GPU max-log demapper (ldpc_gpu_demap) against the oracle, bit for bit,
on every MODCOD; and demapping feeding the decoder on the device."""
import numpy as np
import pytest

from helpers import (codewords, demap_pls_codes, gen, ldpcgpu, modulate, oracle, oracle_demap)

pytestmark = pytest.mark.gpu
PLS = demap_pls_codes()


@pytest.fixture(scope="module")
def dec():
    d = ldpcgpu.Decoder(0)
    yield d
    d.close()


def _run(dec, pls, sym, n0):
    """sym (nframes, nsym, 2) float32, n0 (nframes,) -> LLRs (nframes, nldpc)."""
    nf, ns = sym.shape[0], sym.shape[1]
    assert ldpcgpu.demap_symbols(pls) == ns
    nldpc = 16200 if ldpcgpu.pls_lookup(pls)["short_frame"] else 64800
    d_sym = dec.malloc(sym.nbytes)
    d_n0 = dec.malloc(4 * nf)
    d_llr = dec.malloc(nf * nldpc)
    try:
        dec.upload(d_sym, np.ascontiguousarray(sym))
        dec.upload(d_n0, np.ascontiguousarray(n0, np.float32))
        ldpcgpu.demap(dec, pls, nf, d_sym, d_n0, d_llr)
        dec.synchronize()
        out = np.zeros((nf, nldpc), np.int8)
        dec.download(out, d_llr)
        return out
    finally:
        for p in (d_sym, d_n0, d_llr):
            dec.free(p)


@pytest.mark.parametrize("pls", PLS)
def test_demap_matches_oracle(dec, pls):
    print(f"# demap pls {pls:#04x}", flush=True)
    rng = np.random.default_rng(pls)
    ns = oracle().oracle_demap_symbols(pls)
    sym = np.stack([rng.normal(0, 0.8, (ns, 2)), rng.normal(0, 3.0, (ns, 2))]).astype(np.float32)
    sym[1, :8] = 0.0  # the origin, equidistant from symmetric points
    n0 = np.array([rng.uniform(0.02, 1.5), 1e-3], np.float32)
    got = _run(dec, pls, sym, n0)
    for f in range(2):
        want = oracle_demap(pls, sym[f], n0[f])
        assert (got[f] == want).all(), (hex(pls), f, int((got[f] != want).sum()))


def test_demap_rejects_bad_calls(dec):
    with pytest.raises(RuntimeError):
        ldpcgpu.demap(dec, 0x00, 1, 8, 8, 8)  # dummy PLS: no constellation
    with pytest.raises(RuntimeError):
        ldpcgpu.demap(dec, 0x30, 1, None, None, None)
    assert ldpcgpu.demap_symbols(0x00) == -1
    ldpcgpu.demap(dec, 0x30, 0, None, None, None)  # nothing to do


@pytest.mark.parametrize("pls,margin", [(0x30, 1.5), (0x48, 1.5), (0x60, 1.5), (0x12, 1.0), (0xb8, 1.5), (0xcc, 1.5)])
def test_demap_feeds_decoder(dec, pls, margin):
    """Encoder codewords, mapped and sent through AWGN on the host, demapped
    and decoded on the device: the decoded bits are the codewords'
    information bits."""
    print(f"# demap + decode pls {pls:#04x}", flush=True)
    m = ldpcgpu.pls_lookup(pls)
    short, rate = m["short_frame"], m["rate"]
    fec, _, ref = gen(short, rate, 100.0, 3, 90 + pls)
    cw = codewords(short, rate, ref)
    rng = np.random.default_rng(pls)
    esn0 = ldpcgpu.qef_db(pls) + margin
    syms, n0s = zip(*[modulate(pls, cw[f], esn0, rng) for f in range(3)])
    sym = np.stack(syms)
    llr = _run(dec, pls, sym, np.array(n0s, np.float32))
    for f in range(3):
        assert (llr[f] == oracle_demap(pls, sym[f], n0s[f])).all()
    N = fec.nldpc
    code = ldpcgpu.code_lookup(short, rate)
    frames = np.zeros(3, ldpcgpu.FRAME_DTYPE)
    kb = (fec.kldpc + 7) // 8
    for f in range(3):
        frames[f] = (code, 50, fec.kldpc, -1, f * N, f * kb)
    out, it, ok = dec.decode_host(frames, llr.reshape(-1), 3 * kb)
    assert ok.all(), it
    assert (out.reshape(3, kb) == ref).all()
