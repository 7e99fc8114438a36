"""ldpc_gpu_decode_batch_static (the decode that allocates nothing) against
ldpc_gpu_decode_batch through the C-ABI: bits, iteration counts and
syndrome flags identical, on two calls in a row, the second with a smaller
batch at other offsets (the shared scratch keeps the first call's
contents)."""
import ctypes as C

import numpy as np
import pytest

from helpers import ldpcgpu

import workloads

pytestmark = pytest.mark.gpu


def _decode(dec, entry, frames, d_llr, out_bytes, nframes):
    """Run one entry point into fresh output buffers filled with garbage."""
    L = ldpcgpu.gpu_lib()
    d_out = dec.malloc(out_bytes)
    d_it = dec.malloc(4 * nframes)
    d_ok = dec.malloc(nframes)
    dec.upload(d_out, np.full(out_bytes, 0xA5, np.uint8))
    dec.upload(d_it, np.full(nframes, -7, np.int32))
    dec.upload(d_ok, np.full(nframes, 9, np.uint8))
    fr = np.ascontiguousarray(frames, ldpcgpu.FRAME_DTYPE)
    rc = getattr(L, entry)(dec.ctx, fr.ctypes.data, nframes, d_llr, d_out, d_it, d_ok)
    assert rc == 0, (entry, rc, dec.error())
    dec.synchronize()
    out = np.zeros(out_bytes, np.uint8)
    it = np.zeros(nframes, np.int32)
    ok = np.zeros(nframes, np.uint8)
    dec.download(out, d_out)
    dec.download(it, d_it)
    dec.download(ok, d_ok)
    for p in (d_out, d_it, d_ok):
        dec.free(p)
    return out, it, ok


def test_static_decode_matches_batch_decode_twice():
    b = workloads.mixed_s2(480, seed=21)
    with ldpcgpu.Decoder(0) as dec:
        d_llr = dec.malloc(b.llr.nbytes)
        dec.upload(d_llr, b.llr)
        n_out = b.ref.nbytes
        # first call: the whole mixed batch (normal and short, all S2 rates)
        ref = _decode(dec, "ldpc_gpu_decode_batch", b.frames, d_llr, n_out, b.nframes)
        got = _decode(dec, "ldpc_gpu_decode_batch_static", b.frames, d_llr, n_out, b.nframes)
        for r, g in zip(ref, got):
            assert np.array_equal(r, g)
        assert ref[1].max() > 1  # frames iterated
        # second call: fewer frames, in reverse order, written to other offsets
        sub = b.frames[::-1][:97].copy()
        kb = ((sub["nbits"] + 7) // 8).astype(np.uint64)
        sub["out_off"] = np.concatenate([[0], np.cumsum(kb)[:-1]])
        n2 = int(kb.sum())
        ref2 = _decode(dec, "ldpc_gpu_decode_batch", sub, d_llr, n2, len(sub))
        got2 = _decode(dec, "ldpc_gpu_decode_batch_static", sub, d_llr, n2, len(sub))
        for r, g in zip(ref2, got2):
            assert np.array_equal(r, g)
        dec.free(d_llr)


def test_static_decode_rejects_groups_and_oversize():
    L = ldpcgpu.gpu_lib()
    with ldpcgpu.Decoder(0) as dec:
        fr = np.zeros(2, ldpcgpu.FRAME_DTYPE)
        fr["code"] = ldpcgpu.code_lookup(1, "1_2")
        fr["max_iter"] = 10
        fr["group"] = 0
        d = dec.malloc(64)
        rc = L.ldpc_gpu_decode_batch_static(dec.ctx, fr.ctypes.data, 2, d, d, d, d)
        assert rc != 0 and "group" in dec.error()
        big = np.zeros(65537, ldpcgpu.FRAME_DTYPE)
        rc = L.ldpc_gpu_decode_batch_static(dec.ctx, big.ctypes.data, len(big), d, d, d, d)
        assert rc != 0
        dec.free(d)
