"""GPU BCH decoder (ldpc_gpu_bch_decode) against the oracle's BCH decoder
(itself equal to the reference's drx_bch_decode, tests/test_bch.py):
results and corrected bytes, frame for frame, for error weights 0 .. t
(corrected), beyond t (-1, or the reference's miscorrection), in one batch
per FEC."""
import numpy as np
import pytest

from helpers import flip, gen, ldpcgpu, oracle_bch

pytestmark = pytest.mark.gpu
FECS = [(0, "1_4"), (0, "1_2"), (0, "2_3"), (0, "9_10"), (1, "1_4"), (1, "1_2"), (1, "8_9"), (0, "2_3_t2"),
        (0, "11_20"), (1, "11_45"), (0, "154_180")]


@pytest.fixture(scope="module")
def dec():
    d = ldpcgpu.Decoder(0)
    yield d
    d.close()


def _batch(short, rate, seed, pad=0):
    fec, _, ref = gen(short, rate, 1000.0, 1, seed)
    cw = ref[0]
    rng = np.random.default_rng(seed)
    weights = [0, 0, 1, 2, fec.t - 1, fec.t, fec.t, fec.t + 1, fec.t + 2, fec.t + 3, 2 * fec.t + 1, 40, 200]
    frames = [flip(cw, rng.choice(fec.kldpc, w, replace=False)) for w in weights]
    nb = len(cw)
    buf = np.zeros((len(frames), nb + pad), np.uint8)
    for f, x in enumerate(frames):
        buf[f, :nb] = x
    return fec, frames, buf


def _run(dec, short, rate, buf):
    n, stride = buf.shape
    d_cw = dec.malloc(buf.nbytes)
    d_res = dec.malloc(4 * n)
    try:
        dec.upload(d_cw, np.ascontiguousarray(buf))
        ldpcgpu.bch_decode(dec, short, rate, n, d_cw, stride, d_res)
        dec.synchronize()
        out = np.zeros_like(buf)
        res = np.zeros(n, np.int32)
        dec.download(out, d_cw)
        dec.download(res, d_res)
        return res, out
    finally:
        dec.free(d_cw)
        dec.free(d_res)


@pytest.mark.parametrize("short,rate", FECS)
def test_bch_matches_oracle(dec, short, rate):
    print(f"# bch {short} {rate}", flush=True)
    fec, frames, buf = _batch(short, rate, 31 + len(rate) + 7 * short, pad=5)
    res, out = _run(dec, short, rate, buf)
    nb = len(frames[0])
    for f, x in enumerate(frames):
        r, want = oracle_bch(short, rate, x)
        assert res[f] == r and (out[f, :nb] == want).all(), (f, int(res[f]), r)
        assert (out[f, nb:] == 0).all()  # the padding is not touched
    assert (res[:7] >= 0).all()  # up to t errors: corrected
    assert (res < 0).sum() >= 4


def test_bch_rejects_bad_calls(dec):
    with pytest.raises(RuntimeError):
        ldpcgpu.bch_decode(dec, 0, "7_7", 1, 8, 4050, 8)  # unknown rate
    with pytest.raises(RuntimeError):
        ldpcgpu.bch_decode(dec, 0, "1_2", 1, 8, 100, 8)  # stride below the codeword
    ldpcgpu.bch_decode(dec, 0, "1_2", 0, None, 4050, None)  # nothing to do
