"""DVB-T2 FEC blocks on the GPU (ldpc_gpu_t2_demap, ldpc_gpu_t2_decode)
against the oracle: soft values and MER sums bit for bit on every PLP
configuration, and the whole job (front end, LDPC, BCH) identical to the
oracle's restatement of the reference's FEC thread and, where the
reference's own build is present, to it."""
import itertools
import json
import os
import sys

import numpy as np
import pytest

import helpers as H
from helpers import ldpcgpu

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))

pytestmark = pytest.mark.gpu
CONFIGS = list(itertools.product([1, 0], range(4), range(6)))  # fec_type, mod, cod


@pytest.fixture(scope="module")
def dec():
    d = ldpcgpu.Decoder(0)
    yield d
    d.close()


def _plp(cod, mod, fec_type, rot, til):
    return ldpcgpu.T2Plp(cod, mod, fec_type, rot, til[0], til[1])


class _Bufs:
    def __init__(self, dec, sizes):
        self.dec = dec
        self.p = {k: dec.malloc(max(v, 8)) for k, v in sizes.items()}

    def __getitem__(self, k):
        return self.p[k]

    def close(self):
        for v in self.p.values():
            self.dec.free(v)


def gpu_demap(dec, cod, mod, fec_type, rot, til, cells, csi):
    plp = _plp(cod, mod, fec_type, rot, til)
    N, K, kbch, nc = ldpcgpu.t2_geometry(plp)
    nb = cells.shape[0]
    assert cells.shape == (nb, nc, 2) and csi.shape == (nb, nc)
    b = _Bufs(dec, dict(cells=cells.nbytes, csi=csi.nbytes, llr=nb * N, err=4 * nb, sig=4 * nb))
    try:
        dec.upload(b["cells"], np.ascontiguousarray(cells, np.float32))
        dec.upload(b["csi"], np.ascontiguousarray(csi, np.float32))
        ldpcgpu.t2_demap(dec, plp, nb, b["cells"], b["csi"], b["llr"], b["err"], b["sig"])
        dec.synchronize()
        llr = np.zeros((nb, N), np.int8)
        err = np.zeros(nb, np.float32)
        sig = np.zeros(nb, np.float32)
        dec.download(llr, b["llr"])
        dec.download(err, b["err"])
        dec.download(sig, b["sig"])
        return llr, err, sig
    finally:
        b.close()


def gpu_job(dec, cod, mod, fec_type, rot, til, cells, csi, max_iter, lanes):
    plp = _plp(cod, mod, fec_type, rot, til)
    N, K, kbch, nc = ldpcgpu.t2_geometry(plp)
    nb = cells.shape[0]
    kb = K // 8
    b = _Bufs(dec, dict(cells=cells.nbytes, csi=csi.nbytes, llr=nb * N, cw=nb * kb, it=4 * nb, syn=nb, bch=4 * nb,
                        err=4 * nb, sig=4 * nb))
    try:
        dec.upload(b["cells"], np.ascontiguousarray(cells, np.float32))
        dec.upload(b["csi"], np.ascontiguousarray(csi, np.float32))
        ldpcgpu.t2_decode(dec, plp, nb, b["cells"], b["csi"], max_iter, lanes, b["llr"], b["cw"], b["it"], b["syn"],
                          b["bch"], b["err"], b["sig"])
        dec.synchronize()
        cw = np.zeros((nb, kb), np.uint8)
        res = np.zeros(nb, np.int32)
        err = np.zeros(nb, np.float32)
        sig = np.zeros(nb, np.float32)
        dec.download(cw, b["cw"])
        dec.download(res, b["bch"])
        dec.download(err, b["err"])
        dec.download(sig, b["sig"])
        return cw, res, err, sig
    finally:
        b.close()


def _ids(c):
    return f"{'normal' if c[0] else 'short'}-m{2 * (c[1] + 1)}-{H.T2_RATES[c[2]]}"


@pytest.mark.parametrize("cfg", CONFIGS, ids=[_ids(c) for c in CONFIGS])
def test_t2_demap_matches_oracle(dec, cfg):
    fec_type, mod, cod = cfg
    N, nc, _ = H.t2_geometry(cod, mod, fec_type)
    rng = np.random.default_rng(2000 + 37 * fec_type + 7 * mod + cod)
    for rot, til in itertools.product([0, 1], [(0, 0), (3, 0)]):
        print(f"# t2_demap {_ids(cfg)} rot {rot} til {til}", flush=True)
        cells, csi = H.t2_random_cells(mod, nc, 5, rng, snr_db=rng.uniform(5, 25))
        llr, err, sig = gpu_demap(dec, cod, mod, fec_type, rot, til, cells, csi)
        want, werr, wsig = H.oracle_t2_demap(cod, mod, fec_type, rot, til, cells, csi)
        assert np.array_equal(llr, want), (rot, til, int((llr != want).sum()))
        assert np.array_equal(err.view(np.int32), werr.view(np.int32)), (rot, til, err, werr)
        assert np.array_equal(sig.view(np.int32), wsig.view(np.int32)), (rot, til, sig, wsig)


@pytest.mark.parametrize("mod,til", [(3, (10, 0)), (0, (7, 0)), (2, (4, 1))])
def test_t2_demap_many_blocks(dec, mod, til):
    """A long PLP: 203 normal blocks over several TI-blocks (each restarting
    the cell deinterleaver's shift), all equal to the oracle."""
    print(f"# t2_demap 203 blocks mod {mod} til {til}", flush=True)
    cod, fec_type, rot = 3, 1, 1
    N, nc, _ = H.t2_geometry(cod, mod, fec_type)
    rng = np.random.default_rng(9 + mod)
    cells, csi = H.t2_random_cells(mod, nc, 203, rng, snr_db=12.0)
    llr, err, sig = gpu_demap(dec, cod, mod, fec_type, rot, til, cells, csi)
    want, werr, wsig = H.oracle_t2_demap(cod, mod, fec_type, rot, til, cells, csi)
    assert np.array_equal(llr, want)
    assert np.array_equal(err.view(np.int32), werr.view(np.int32))
    assert np.array_equal(sig.view(np.int32), wsig.view(np.int32))


JOBS = [(0, 0, 0, 0, 20.0), (0, 1, 1, 1, 13.0), (0, 3, 2, 1, 18.5), (1, 2, 2, 1, 16.0), (1, 3, 5, 0, 30.0),
        (0, 2, 3, 1, 15.0)]


@pytest.mark.parametrize("job", JOBS, ids=[f"{_ids(j[:3])}-rot{j[3]}-{j[4]}dB" for j in JOBS])
def test_t2_job_matches_oracle_and_reference(dec, job):
    """Encoded BBFRAMEs through the transmit model: the GPU job returns what
    the oracle's job does (LDPC calls of 32 blocks as the reference's AVX2
    build makes them), and, where present, what the reference's build does
    with its own lane count."""
    fec_type, mod, cod, rot, snr = job
    print(f"# t2_decode {_ids(job[:3])} rot {rot} at {snr} dB", flush=True)
    rng = np.random.default_rng(int(snr * 10) + cod)
    nblocks = 40
    til = (3, 0)
    info, bits = H.t2_codewords(cod, fec_type, nblocks, rng)
    cells, csi = H.t2_modulate(cod, mod, fec_type, rot, til, bits, snr, rng)
    cw, res, err, sig = gpu_job(dec, cod, mod, fec_type, rot, til, cells, csi, 30, 32)
    wcw, wres, werr, wsig = H.oracle_t2_job(cod, mod, fec_type, rot, til, cells, csi, 30, 32)
    assert np.array_equal(res, wres), (res, wres)
    assert np.array_equal(cw, wcw)
    assert np.array_equal(err.view(np.int32), werr.view(np.int32))
    kb = info.shape[1]
    assert np.array_equal(cw[res >= 0, :kb], info[res >= 0])
    for name, (L, lanes) in H.t2_ref_builds().items():
        gcw, gres, _, _ = gpu_job(dec, cod, mod, fec_type, rot, til, cells, csi, 30, lanes)
        bb, ok, _, _, _ = H.ref_t2_job(L, cod, mod, fec_type, rot, til, cells, csi, 30)
        assert np.array_equal(ok, (gres >= 0).astype(np.uint8)), name
        assert np.array_equal(bb, gcw[:, :kb]), name


def test_t2_job_independent_blocks(dec):
    """lanes 0: every block stops on its own; the result equals the oracle
    decoding each block alone."""
    fec_type, mod, cod, rot = 0, 1, 0, 1
    rng = np.random.default_rng(3)
    info, bits = H.t2_codewords(cod, fec_type, 12, rng)
    cells, csi = H.t2_modulate(cod, mod, fec_type, rot, (0, 0), bits, 7.0, rng)
    cw, res, _, _ = gpu_job(dec, cod, mod, fec_type, rot, (0, 0), cells, csi, 30, 0)
    wcw, wres, _, _ = H.oracle_t2_job(cod, mod, fec_type, rot, (0, 0), cells, csi, 30, 1)
    assert np.array_equal(res, wres) and np.array_equal(cw, wcw)


def test_t2_rejects_bad_calls(dec):
    with pytest.raises(ValueError):
        ldpcgpu.t2_geometry(_plp(6, 0, 1, 0, (0, 0)))
    with pytest.raises(RuntimeError):
        ldpcgpu.t2_demap(dec, _plp(0, 4, 1, 0, (0, 0)), 1, 8, 8, 8, 8, 8)
    with pytest.raises(RuntimeError):
        ldpcgpu.t2_demap(dec, _plp(0, 1, 1, 0, (0, 0)), 1, None, None, None, None, None)
    ldpcgpu.t2_demap(dec, _plp(0, 1, 1, 0, (0, 0)), 0, None, None, None, None, None)  # nothing to do
    # 2 normal 256-QAM blocks per TI-block would need more shifts than the
    # 8100-cell interleaver's counter gives for 16384 blocks
    with pytest.raises(RuntimeError):
        ldpcgpu.t2_demap(dec, _plp(0, 3, 1, 0, (1, 0)), 16384, 8, 8, 8, 8, 8)


@pytest.mark.parametrize("nblocks,til,mod,ft", [(1, (0, 0), 3, 1), (1, (4, 0), 0, 0), (5, (8, 0), 2, 1),
                                                 (7, (3, 1), 3, 0), (2, (1, 0), 1, 0)])
def test_t2_demap_edge_shapes(dec, nblocks, til, mod, ft):
    """One block; fewer blocks than TI-blocks (some TI-blocks empty); a
    type-1 interleaver (no shift); 256-QAM short frames (one cell per
    demultiplexer word): equal to the oracle."""
    print(f"# t2_demap edge nblocks {nblocks} til {til} mod {mod} fec_type {ft}", flush=True)
    cod = 4
    N, nc, _ = H.t2_geometry(cod, mod, ft)
    rng = np.random.default_rng(nblocks * 31 + mod)
    cells, csi = H.t2_random_cells(mod, nc, nblocks, rng, snr_db=9.0)
    for rot in (0, 1):
        llr, err, sig = gpu_demap(dec, cod, mod, ft, rot, til, cells, csi)
        want, werr, wsig = H.oracle_t2_demap(cod, mod, ft, rot, til, cells, csi)
        assert np.array_equal(llr, want), rot
        assert np.array_equal(err.view(np.int32), werr.view(np.int32)), rot
        assert np.array_equal(sig.view(np.int32), wsig.view(np.int32)), rot


def test_t2_saturated_and_zero_state(dec):
    """A block scaled 40x (its own noise estimate grows with it: every soft
    value rounds to 0), a 30 dB block with exact zeros (soft values
    saturate at +-127), and channel state 0 (the 1e-12 guard keeps the
    weights finite): equal to the oracle."""
    cod, mod, ft = 3, 2, 0
    N, nc, _ = H.t2_geometry(cod, mod, ft)
    rng = np.random.default_rng(8)
    cells, csi = H.t2_random_cells(mod, nc, 3, rng, snr_db=30.0)
    cells[0] *= 40.0
    cells[1, :100] = 0.0
    csi[2, ::7] = 0.0
    for rot in (0, 1):
        llr, err, sig = gpu_demap(dec, cod, mod, ft, rot, (0, 0), cells, csi)
        want, werr, wsig = H.oracle_t2_demap(cod, mod, ft, rot, (0, 0), cells, csi)
        assert np.array_equal(llr, want), rot
        assert np.array_equal(err.view(np.int32), werr.view(np.int32)), rot
        assert np.array_equal(sig.view(np.int32), wsig.view(np.int32)), rot
        assert (llr[0] == 0).all()
        if rot == 0:
            assert (np.abs(llr[1].astype(int)) == 127).mean() > 0.3


GOLDEN = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "t2_golden.json")))


@pytest.mark.parametrize("k", range(len(GOLDEN["front_end"])))
def test_t2_demap_matches_golden(dec, k):
    """The GPU front end against the reference's own outputs (digests in
    tests/golden/t2_golden.json), inputs regenerated from their seeds."""
    import make_t2_golden as G

    g = GOLDEN["front_end"][k]
    cod, mod, ft, rot, tl, tt, nb, snr, seed = g["case"]
    print(f"# t2_demap golden {g['case']}", flush=True)
    cells, csi = G.front_inputs(g["case"])
    llr, err, sig = gpu_demap(dec, cod, mod, ft, rot, (tl, tt), cells, csi)
    assert G.sha(llr) == g["llr_sha256"] and G.sha(err) == g["err_sha256"] and G.sha(sig) == g["sig_sha256"]


@pytest.mark.parametrize("k", range(len(GOLDEN["jobs"])))
def test_t2_job_matches_golden(dec, k):
    import make_t2_golden as G

    g = GOLDEN["jobs"][k]
    cod, mod, ft, rot, tl, tt, nb, snr, seed = g["case"]
    print(f"# t2_decode golden {g['case']}", flush=True)
    info, cells, csi = G.job_inputs(g["case"])
    cw, res, err, sig = gpu_job(dec, cod, mod, ft, rot, (tl, tt), cells, csi, g["max_iter"], g["lanes"])
    kb = info.shape[1]
    assert G.sha(np.ascontiguousarray(cw[:, :kb])) == g["bb_sha256"]
    assert (res >= 0).astype(int).tolist() == g["ok"]
    assert G.sha(err) == g["err_sha256"] and G.sha(sig) == g["sig_sha256"]


@pytest.mark.parametrize("k", range(len(GOLDEN["jobs"])))
def test_t2_decode_host_matches_golden(dec, k):
    """ldpc_gpu_t2_decode_host, decode_job's own interface on host memory:
    BBFRAMEs, ok flags and MER sums equal to the reference's (golden
    digests, and the reference build where present)."""
    import make_t2_golden as G

    g = GOLDEN["jobs"][k]
    cod, mod, ft, rot, tl, tt, nb, snr, seed = g["case"]
    print(f"# t2_decode_host golden {g['case']}", flush=True)
    info, cells, csi = G.job_inputs(g["case"])
    plp = _plp(cod, mod, ft, rot, (tl, tt))
    bb, ok, err, sig = ldpcgpu.t2_decode_host(dec, plp, cells, csi, g["max_iter"], g["lanes"])
    assert G.sha(bb) == g["bb_sha256"] and ok.tolist() == g["ok"]
    assert G.sha(err) == g["err_sha256"] and G.sha(sig) == g["sig_sha256"]
    for name, (L, lanes) in H.t2_ref_builds().items():
        bb2, ok2, err2, sig2 = ldpcgpu.t2_decode_host(dec, plp, cells, csi, 30, lanes)
        rbb, rok, rerr, rsig, _ = H.ref_t2_job(L, cod, mod, ft, rot, (tl, tt), cells, csi, 30)
        assert np.array_equal(bb2, rbb) and np.array_equal(ok2, rok), name
        assert np.array_equal(err2.view(np.int32), rerr.view(np.int32)), name
