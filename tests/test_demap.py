"""Pins of the oracle's max-log demapper (oracle_demap_frame, restating
libdvbrx/src/demap.c:164-203 and :491-532):

* the reference demapper built from its own sources (oracle/_ref, the
  AVX-512 and AVX2 builds: the x87 scalar build rounds differently), on
  every MODCOD, frame for frame;
* the closed form of the QPSK path;
* max-log from its definition (squared distances in float64 to every
  point, minimum per label bit): within one LLR step;
* noiseless constellation points give LLR signs equal to the label bits,
  through the bit interleaver.
"""
import numpy as np
import pytest

from helpers import (demap_order, demap_points, demap_pls_codes, ldpcgpu, oracle, oracle_demap, ref_builds,
                     ref_demap)

REFS = {k: v for k, v in ref_builds().items() if k in ("v4", "v3")}
needs_ref = pytest.mark.skipif(not REFS, reason="oracle/_ref not built (needs the reference tree)")
PLS = demap_pls_codes()


def _frame(pls, seed, scale=0.8):
    rng = np.random.default_rng(seed)
    ns = oracle().oracle_demap_symbols(pls)
    return rng.normal(0, scale, (ns, 2)).astype(np.float32), np.float32(rng.uniform(0.02, 1.5))


def test_pls_codes():
    assert len(PLS) >= 100
    assert all(oracle().oracle_demap_symbols(p) > 0 for p in PLS)
    assert oracle().oracle_demap_symbols(0x00) == -1  # dummy frame: no constellation
    assert sum(oracle().oracle_demap_ring(p) == 1 for p in PLS) == 4


@needs_ref
@pytest.mark.parametrize("pls", PLS)
def test_oracle_demap_matches_reference(pls):
    for seed, scale in ((pls, 0.8), (pls + 1000, 3.0)):
        sym, n0 = _frame(pls, seed, scale)
        mine = oracle_demap(pls, sym, n0)
        for name, R in REFS.items():
            ref = ref_demap(R, pls, sym, n0)
            assert (ref == mine).all(), (name, hex(pls), int((ref != mine).sum()))


def test_qpsk_closed_form():
    pls = 0x10  # normal 1/2 QPSK
    sym, n0 = _frame(pls, 3)
    k = np.float32(np.float32(2.0) / n0)
    kq = np.float32(np.float32(k * np.float32(4.0)) * np.float32(0.70710678118654752440))
    want = np.rint(np.clip((kq * sym.reshape(-1)).astype(np.float32), -127, 127)).astype(np.int8)
    assert (oracle_demap(pls, sym, n0) == want).all()


# one MODCOD of every constellation (256APSK: ring and full-search forms)
ONE_EACH = [0x14, 0x16, 0x30, 0x32, 0x48, 0x4a, 0x60, 0x62, 0x8a, 0x94, 0xae, 0xb2, 0xb8, 0xba, 0xbe, 0xc8, 0xcc,
            0xce, 0xf6]


@pytest.mark.parametrize("pls", ONE_EACH)
def test_maxlog_definition(pls):
    """LLR_b = (min over label bit 1 of |y - p|^2 - min over bit 0) * 2 / n0,
    in float64: the oracle's float metric and rounding stay within one
    step, on noisy points (away from saturation)."""
    pts = demap_points(pls)
    nb = int(np.log2(len(pts)))
    ns = oracle().oracle_demap_symbols(pls)
    m = ldpcgpu.pls_lookup(pls)
    nldpc = 16200 if m["short_frame"] else 64800
    rng = np.random.default_rng(pls)
    lab = rng.integers(0, len(pts), ns)
    y = pts[lab] + (rng.normal(0, 0.05, ns) + 1j * rng.normal(0, 0.05, ns))
    n0 = np.float32(0.6)
    sym = np.stack([y.real, y.imag], 1).astype(np.float32)
    llr = oracle_demap(pls, sym, n0).astype(np.int64)
    ys = sym[:, 0].astype(np.float64) + 1j * sym[:, 1].astype(np.float64)
    d2 = np.abs(ys[:, None] - pts[None, :].astype(np.complex128)) ** 2
    bits = (np.arange(len(pts))[:, None] >> (nb - 1 - np.arange(nb))[None, :]) & 1
    order = demap_order(pls)
    for b in range(nb):
        m1 = d2[:, bits[:, b] == 1].min(1)
        m0 = d2[:, bits[:, b] == 0].min(1)
        want = np.clip((m1 - m0) * 2.0 / float(n0), -127, 127)
        if order is None:
            continue
        pos = order[b] * ns + np.arange(ns)
        ok = pos < nldpc
        assert np.abs(llr[pos[ok]] - want[ok]).max() <= 1.0, (hex(pls), b)


@pytest.mark.parametrize("pls", ONE_EACH)
def test_noiseless_points_give_label_bits(pls):
    pts = demap_points(pls)
    nb = int(np.log2(len(pts)))
    ns = oracle().oracle_demap_symbols(pls)
    m = ldpcgpu.pls_lookup(pls)
    nldpc = 16200 if m["short_frame"] else 64800
    rng = np.random.default_rng(7)
    lab = rng.integers(0, len(pts), ns)
    sym = np.stack([pts[lab].real, pts[lab].imag], 1).astype(np.float32)
    # small n0: the LLR of the nearest other point's distance is >= 1
    llr = oracle_demap(pls, sym, np.float32(1e-3))
    bits = (lab[:, None] >> (nb - 1 - np.arange(nb))[None, :]) & 1
    order = demap_order(pls)
    if order is None:  # QPSK: two bits a symbol in codeword order
        got = (llr.reshape(ns, 2) < 0).astype(np.int64)
        assert (got == bits).all()
        return
    assert sorted(order) == list(range(nb))
    # a few tables hold near-coincident points (S2X 256APSK 20/30 and 22/30:
    # 80 of the 256 points have a partner 1.4e-4 away that differs in the
    # last label bit): a bit that only such a pair
    # tells apart has no LLR; checked where the nearest point of the other
    # bit value is at least 0.03 away
    allb = (np.arange(len(pts))[:, None] >> (nb - 1 - np.arange(nb))[None, :]) & 1
    dist = np.abs(pts[:, None] - pts[None, :])
    for b in range(nb):
        far = np.array([dist[p, allb[:, b] != allb[p, b]].min() >= 0.03 for p in range(len(pts))])
        pos = order[b] * ns + np.arange(ns)
        ok = (pos < nldpc) & far[lab]
        assert ok.mean() > 0.6, (hex(pls), b)
        assert ((llr[pos[ok]] < 0) == (bits[ok, b] == 1)).all(), (hex(pls), b)
        assert (llr[pos[ok]] != 0).all()
