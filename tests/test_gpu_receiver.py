"""The receiver loop on the GPU (dvb-rx_amd/receiver.py): sixteen
receivers' frames demapped (or through the T2 front end), decoded and
BCH-corrected inside the timed loop, handed on in order.  Every sampled
frame's decoded bits, iteration count and BCH result equal the oracle's
chain on the same received symbols or cells: the oracle demapper (or T2
front end), the oracle decoder with the stream's iteration cap, the
oracle BCH decoder."""
import numpy as np
import pytest

from helpers import Oracle, ldpcgpu, oracle_bch, oracle_demap, oracle_t2_demap
import receiver  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("batch_max", [1, 16])
def test_receiver_frames_match_the_oracle(batch_max):
    print(f"# receiver loop 0.03 s, 2 contexts, batch_max {batch_max}", flush=True)
    decs = [ldpcgpu.Decoder(0), ldpcgpu.Decoder(0)]
    try:
        rec, tr, res = receiver.run(decs, duration=0.03, seed=23, capacity=512, batch_max=batch_max)
    finally:
        for d in decs:
            d.close()
    assert rec["frames"] == tr.n > 100
    assert rec["frame_errors_after_bch"] == 0
    rng = np.random.default_rng(1)
    for k, kd in enumerate(tr.kinds):
        sel = np.nonzero(tr.kind == k)[0]
        if len(sel) == 0:
            continue
        pick = rng.choice(sel, min(6, len(sel)), replace=False)
        fec = kd["fec"]
        o = Oracle(kd["short"], kd["rate"], fec.kldpc)
        for f in pick:
            p = tr.pos[f]
            if kd["t2"]:
                llr, _, _ = oracle_t2_demap(kd["cod"], kd["mod"], kd["ft"], 1, (1, 0), kd["cells"][p:p + 1],
                                            kd["csi"][p:p + 1])
                llr = llr[0]
            else:
                llr = oracle_demap(kd["pls"], kd["sym"][p], kd["n0"][p])
            bits, it, ok = o.frames(llr[None, :], kd["max_iter"])
            out, git, gok, gbch = res[int(f)]
            assert git == it[0] and gok == ok[0], (kd["name"], f)
            assert (out == bits[0]).all(), (kd["name"], f)
            want, _ = oracle_bch(kd["short"], kd["rate"], bits[0])
            assert gbch == want, (kd["name"], f)


def test_pinned_buffers_and_async_copies_round_trip():
    """ldpc_gpu_host_alloc / upload_async / download_async: bytes through
    the device and back, seen complete by ldpc_gpu_query."""
    dec = ldpcgpu.Decoder(0)
    try:
        a = dec.host_array((3, 1000), np.int32)
        b = dec.host_array((3, 1000), np.int32)
        d = dec.malloc(a.nbytes)
        try:
            a[...] = np.arange(a.size, dtype=np.int32).reshape(a.shape) * 7 - 11
            b[...] = 0
            dec.upload_async(d, a.ctypes.data, a.nbytes)
            dec.download_async(b.ctypes.data + 4000, d + 4000, 8000)
            dec.download_async(b.ctypes.data, d, 4000)
            while not dec.query():
                pass
            assert (a == b).all()
        finally:
            dec.free(d)
            dec.host_free(a)
            dec.host_free(b)
    finally:
        dec.close()
