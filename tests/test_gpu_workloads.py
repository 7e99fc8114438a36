"""The benchmark configurations' own workloads through the HIP path,
checked against the oracle frame by frame.

* the mixed DVB-S2 batch (max-log MODCOD LLRs at QEF + 0.3 dB, 75 % normal
  frames, every S2 rate): a slice of the 8192-frame batch decoded in one
  ldpc_gpu_decode_batch call, every frame's bits, iteration count and
  syndrome flag equal to the oracle's; the whole batch against the sent
  bits as a property check;
* the DVB-S2X high-order batch (64/128/256APSK normal, QPSK/8PSK short,
  threshold + 0.2 dB) generated on the GPU and decoded through the
  autotuned graph plan, a subset per MODCOD against the oracle;
* the GPU channel those LLRs come from: its symbols demapped by the oracle
  give its LLRs, on every MODCOD of the S2X batch.
"""
import collections
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from helpers import Oracle, ldpcgpu, oracle

import workloads

pytestmark = pytest.mark.gpu


def _oracle_frames(frames_by_code, max_iter):
    """{(short, rate): (bits, it, ok)} from the oracle, codes in threads."""

    def run(key):
        llr = frames_by_code[key]
        fec = ldpcgpu.fec_lookup(*key)
        return key, Oracle(key[0], key[1], fec.kldpc).frames(llr, max_iter)

    with ThreadPoolExecutor(16) as ex:
        return dict(ex.map(run, frames_by_code))


def _decode(dec, frames, llr, out_bytes):
    d_llr = dec.malloc(llr.nbytes)
    d_out = dec.malloc(out_bytes)
    d_it = dec.malloc(4 * len(frames))
    d_ok = dec.malloc(len(frames))
    try:
        dec.upload(d_llr, llr)
        dec.decode_batch(frames, d_llr, d_out, d_it, d_ok)
        dec.synchronize()
        out = np.zeros(out_bytes, np.uint8)
        it = np.zeros(len(frames), np.int32)
        ok = np.zeros(len(frames), np.uint8)
        dec.download(out, d_out)
        dec.download(it, d_it)
        dec.download(ok, d_ok)
        return out, it, ok
    finally:
        for p in (d_llr, d_out, d_it, d_ok):
            dec.free(p)


def _modcod_key(pls):
    m = ldpcgpu.pls_lookup(pls)
    return (m["short_frame"], m["rate"])


def test_mixed_s2_batch_slice_vs_oracle():
    """1024 frames of the 8192-frame mixed batch (its first specs), one call."""
    specs = workloads.mixed_s2_specs(8192)[:1024]
    b = workloads.Batch("mixed_s2 slice", specs, 7, 50)
    with ldpcgpu.Decoder(0) as dec:
        out, it, ok = _decode(dec, b.frames, b.llr, b.ref.nbytes)
    by = collections.defaultdict(list)
    for f in range(b.nframes):
        by[_modcod_key(specs[f]["pls"])].append(f)
    res = _oracle_frames({k: np.stack([b.frame_llr(f) for f in fs]) for k, fs in by.items()}, 50)
    rates = set()
    for key, fs in by.items():
        o_bits, o_it, o_ok = res[key]
        for r, f in enumerate(fs):
            lo = b.layout[f][4]
            assert np.array_equal(out[lo:lo + o_bits.shape[1]], o_bits[r]), (key, f)
            assert (it[f], ok[f]) == (o_it[r], o_ok[r]), (key, f, it[f], o_it[r])
        rates.add(key)
    # the slice holds normal and short frames of many rates, and frames
    # that fail (the normal QPSK 1/4 MODCOD at QEF + 0.3 dB)
    assert len({k for k in rates if k[0] == 0}) >= 10 and len({k for k in rates if k[0] == 1}) >= 8
    assert (ok == 0).any() and it.max() == 50 and it.min() < 10


def test_mixed_s2_full_batch_properties():
    """The whole 8192-frame batch: every frame that meets its checks decodes
    to the sent bits; the frames that do not ran 50 iterations; frame
    errors outside the normal QPSK 1/4 MODCOD stay below 1e-3."""
    b = workloads.mixed_s2(8192)
    with ldpcgpu.Decoder(0) as dec:
        out, it, ok = _decode(dec, b.frames, b.llr, b.ref.nbytes)
    assert ((it >= 0) & (it <= 50)).all()
    assert (it[ok == 0] == 50).all()
    bad = np.zeros(b.nframes, bool)
    for f in range(b.nframes):
        bad[f] = not np.array_equal(b.frame_ref(f), out[b.layout[f][4]:b.layout[f][4] + (b.kldpc[f] + 7) // 8])
    assert not (bad & (ok == 1)).any()
    qpsk14 = np.array([_modcod_key(s["pls"]) == (0, "1_4") and ldpcgpu.pls_lookup(s["pls"])["constellation"] == "QPSK"
                       for s in b.specs])
    assert bad[~qpsk14].mean() < 1e-3


def test_s2x_high_order_plan_vs_oracle():
    """Six frames per MODCOD of the S2X high-order batch, generated on the
    GPU at threshold + 0.2 dB, decoded by the autotuned graph plan."""
    groups = [(pls, 6) for pls in workloads.S2X_NORMAL_HIGH + workloads.S2X_SHORT_LOW]
    with ldpcgpu.Decoder(0) as dec:
        b = workloads.GpuBatch(dec, groups, 0.2, seed=5)
        try:
            plan = ldpcgpu.Plan(dec, b.frames, b.d_llr, b.d_out, b.d_it, b.d_ok)
            try:
                plan.tune(1)
                plan.launch()
                dec.synchronize()
            finally:
                plan.close()
            out, it, ok = b.results()
            llr = np.zeros(b.llr_bytes, np.int8)
            dec.download(llr, b.d_llr)
        finally:
            b.free()
    by = collections.defaultdict(list)
    for pls, code, N, base, count, esn0 in b.layout:
        for f in range(count):
            by[_modcod_key(pls)].append((pls, llr[base + f * N:base + (f + 1) * N]))
    res = _oracle_frames({k: np.stack([x for _, x in v]) for k, v in by.items()}, 50)
    f = 0
    seen = collections.Counter()
    for pls, code, N, base, count, esn0 in b.layout:
        key = _modcod_key(pls)
        o_bits, o_it, o_ok = res[key]
        for r in range(count):
            i = seen[key] + r
            lo = int(b.frames["out_off"][f])
            assert np.array_equal(out[lo:lo + o_bits.shape[1]], o_bits[i]), (hex(pls), r)
            assert (it[f], ok[f]) == (o_it[i], o_ok[i]), (hex(pls), r, it[f], o_it[i])
            f += 1
        seen[key] += count


@pytest.mark.parametrize("pls", workloads.S2X_NORMAL_HIGH + workloads.S2X_SHORT_LOW)
def test_channel_llrs_are_the_oracle_demap_of_its_symbols(pls):
    from nms_train import codeword_pool

    m = ldpcgpu.pls_lookup(pls)
    fec = ldpcgpu.fec_lookup(m["short_frame"], m["rate"])
    pool = codeword_pool(m["short_frame"], m["rate"], 2, 3)
    nf, N = 3, fec.nldpc
    nsym = ldpcgpu.demap_symbols(pls)
    L = ldpcgpu.gpu_lib()
    with ldpcgpu.Decoder(0) as dec:
        ptrs = []

        def put(a):
            a = np.ascontiguousarray(a)
            p = dec.malloc(a.nbytes)
            ptrs.append(p)
            dec.upload(p, a)
            return p

        d_cw = put(pool)
        d_row = put(np.array([0, 1, 0], np.int32))
        d_e = put(np.full(nf, ldpcgpu.qef_db(pls) + 0.2, np.float32))
        d_sym = put(np.zeros(2 * nsym * nf, np.float32))
        d_n0 = put(np.zeros(nf, np.float32))
        d_llr = put(np.zeros(nf * N, np.int8))
        assert L.ldpc_gpu_channel_symbols(pls, N, d_cw, d_row, nf, d_e, 77, d_sym, d_n0, dec.stream) == 0
        ldpcgpu.channel(pls, N, d_cw, d_row, nf, d_e, 77, d_llr, stream=dec.stream)
        dec.synchronize()
        sym = np.zeros(2 * nsym * nf, np.float32)
        n0 = np.zeros(nf, np.float32)
        llr = np.zeros(nf * N, np.int8)
        dec.download(sym, d_sym)
        dec.download(n0, d_n0)
        dec.download(llr, d_llr)
        for p in ptrs:
            dec.free(p)
    O = oracle()
    for f in range(nf):
        ref = np.zeros(N, np.int8)
        s = np.ascontiguousarray(sym[2 * nsym * f:2 * nsym * (f + 1)])
        assert O.oracle_demap_frame(pls, s.ctypes.data, float(n0[f]), ref.ctypes.data) == 0
        assert np.array_equal(ref, llr[f * N:(f + 1) * N]), (hex(pls), f)
    # the symbols are the codeword's points with noise: hard decisions of
    # the LLRs agree with the sent bits far more often than chance (the
    # lowest S2X MODCODs run at a raw bit error rate near 25 %)
    assert ((llr[:N] < 0) == (pool[0] == 1)).mean() > 0.65
