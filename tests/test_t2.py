"""The oracle's DVB-T2 FEC block front end (oracle/ldpc_oracle.c, after
libdvbrx/src/dvbt2.c:1953-2082 and :2188-2400) pinned three ways:

- against the reference's own decode_job, compiled from its sources by
  oracle/Makefile (oracle/_ref/libdrxt2ref_*.so), on random cells of every
  PLP configuration: soft values, MER sums, BBFRAMEs and BCH results
  identical;
- structure: L0 is a permutation, the bit deinterleaver map is a bijection,
  QPSK carries no bit interleaving, each TI-block's shifts are distinct;
- round trips: random BBFRAMEs through the BCH and LDPC encoders and a
  transmit model (tests/helpers.t2_modulate) come back from the oracle's
  job and the reference's.
"""
import itertools
import json
import os
import sys

import numpy as np
import pytest

import helpers as H

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))

CONFIGS = list(itertools.product([1, 0], range(4), range(6)))  # fec_type, mod, cod
REFS = H.t2_ref_builds()


def _ids(c):
    return f"{'normal' if c[0] else 'short'}-m{2 * (c[1] + 1)}-{H.T2_RATES[c[2]]}"


@pytest.mark.parametrize("cfg", CONFIGS, ids=[_ids(c) for c in CONFIGS])
def test_bit_deinterleaver_is_a_bijection(cfg):
    fec_type, mod, cod = cfg
    N, nc, code = H.t2_geometry(cod, mod, fec_type)
    assert N == (16200 if fec_type == 0 else 64800) and nc * 2 * (mod + 1) == N
    assert code == ("2_3_t2" if (fec_type, cod) == (1, 2) else "3_5_t2" if (fec_type, cod) == (0, 1)
                    else H.T2_RATES[cod])
    src = np.zeros(N, np.uint16)
    assert H.oracle().oracle_t2_bitsrc(cod, mod, fec_type, src.ctypes.data) == 0
    assert np.array_equal(np.sort(src.astype(np.int64)), np.arange(N))
    if mod == 0:
        assert np.array_equal(src, np.arange(N))


@pytest.mark.parametrize("ncells", [2025, 2700, 3240, 4050, 5400, 8100, 10800, 16200, 32400])
def test_cell_permutation(ncells):
    perm = np.zeros(ncells, np.int32)
    assert H.oracle().oracle_t2_cellperm(ncells, perm.ctypes.data) == 0
    assert np.array_equal(np.sort(perm), np.arange(ncells))
    # EN 302 755 clause 6.5: L0(0) = 0 and the second entry has only the
    # toggled top bit set
    nd = int(np.ceil(np.log2(ncells)))
    assert perm[0] == 0 and perm[1] == 1 << (nd - 1)


@pytest.mark.parametrize("nblocks,nti", [(10, 3), (7, 7), (5, 8), (40, 1)])
def test_ti_block_shifts(nblocks, nti):
    nc = 8100
    P = np.zeros(nblocks, np.int32)
    H.oracle().oracle_t2_shift(nblocks, nti, 0, nc, P.ctypes.data)
    sizes = [nblocks // nti + (s >= nti - nblocks % nti) for s in range(nti)]
    assert sum(sizes) == nblocks
    i = 0
    for n in sizes:
        part = P[i:i + n]
        # first block of a TI-block: counter 0 -> shift 0; then the
        # bit-reversed counter (14 bits for 8100 cells): 1 -> 8192 (skipped,
        # >= 8100), 2 -> 4096, 3 -> 12288 (skipped), 4 -> 2048
        assert list(part[:3]) == [0, 4096, 2048][:n]
        assert len(set(part.tolist())) == n and (part < nc).all()
        i += n
    P1 = np.full(nblocks, -1, np.int32)
    H.oracle().oracle_t2_shift(nblocks, nti, 1, nc, P1.ctypes.data)
    assert (P1 == 0).all()  # only a type-0 time interleaver restarts the shift


@pytest.mark.skipif(not REFS, reason="reference T2 build (oracle/_ref) not present")
@pytest.mark.parametrize("cfg", CONFIGS, ids=[_ids(c) for c in CONFIGS])
def test_front_end_matches_reference(cfg):
    """Random cells, rotation off and on, with and without a type-0 time
    interleaver: soft values and MER sums bit-identical to every reference
    build present."""
    fec_type, mod, cod = cfg
    N, nc, _ = H.t2_geometry(cod, mod, fec_type)
    rng = np.random.default_rng(1000 + 37 * fec_type + 7 * mod + cod)
    for rot, til in itertools.product([0, 1], [(0, 0), (3, 0)]):
        cells, csi = H.t2_random_cells(mod, nc, 5, rng, snr_db=rng.uniform(5, 25))
        llr, err, sig = H.oracle_t2_demap(cod, mod, fec_type, rot, til, cells, csi)
        for name, (L, lanes) in REFS.items():
            _, _, rerr, rsig, rllr = H.ref_t2_job(L, cod, mod, fec_type, rot, til, cells, csi, 1, with_llr=True)
            assert np.array_equal(rllr, llr), (name, rot, til, np.argwhere(rllr != llr)[:5])
            assert np.array_equal(rerr.view(np.int32), err.view(np.int32)), (name, rot, til)
            assert np.array_equal(rsig.view(np.int32), sig.view(np.int32)), (name, rot, til)


ROUND_TRIP = [(0, 0, 0, 0), (0, 1, 1, 1), (0, 3, 2, 1), (0, 3, 5, 0), (1, 1, 2, 1), (1, 2, 1, 0)]


@pytest.mark.parametrize("cfg", ROUND_TRIP, ids=[f"{_ids(c[:3])}-rot{c[3]}" for c in ROUND_TRIP])
def test_round_trip_job(cfg):
    """BBFRAMEs -> BCH/LDPC encoders -> transmit model -> AWGN well above
    threshold -> the oracle's job returns every BBFRAME with BCH success,
    and the reference's job gives the same BBFRAMEs, results and MER sums."""
    fec_type, mod, cod, rot = cfg
    rng = np.random.default_rng(77 + sum(cfg))
    nblocks = 3
    til = (2, 0)
    info, bits = H.t2_codewords(cod, fec_type, nblocks, rng)
    cells, csi = H.t2_modulate(cod, mod, fec_type, rot, til, bits, 18.0 + 6 * mod, rng)
    cw, res, err, sig = H.oracle_t2_job(cod, mod, fec_type, rot, til, cells, csi, 30, 32)
    kb = info.shape[1]
    assert (res >= 0).all(), res
    assert np.array_equal(cw[:, :kb], info)
    for name, (L, lanes) in REFS.items():
        cw_l, res_l, _, _ = H.oracle_t2_job(cod, mod, fec_type, rot, til, cells, csi, 30, lanes)
        bb, ok, rerr, rsig, _ = H.ref_t2_job(L, cod, mod, fec_type, rot, til, cells, csi, 30)
        assert np.array_equal(bb, cw_l[:, :kb]), name
        assert np.array_equal(ok, (res_l >= 0).astype(np.uint8)), name
        assert np.array_equal(rerr.view(np.int32), err.view(np.int32)), name
        assert np.array_equal(rsig.view(np.int32), sig.view(np.int32)), name


def test_noisy_job_matches_reference():
    """Near threshold some blocks fail: the reference's job and the
    oracle's (its LDPC calls grouped by the build's lanes) agree on every
    BBFRAME and BCH result."""
    if not REFS:
        pytest.skip("reference T2 build (oracle/_ref) not present")
    fec_type, mod, cod, rot = 0, 2, 3, 1
    rng = np.random.default_rng(5)
    nblocks = 6
    info, bits = H.t2_codewords(cod, fec_type, nblocks, rng)
    cells, csi = H.t2_modulate(cod, mod, fec_type, rot, (0, 0), bits, 15.0, rng)
    for name, (L, lanes) in REFS.items():
        cw, res, _, _ = H.oracle_t2_job(cod, mod, fec_type, rot, (0, 0), cells, csi, 15, lanes)
        assert (res >= 0).any() and (res < 0).any(), res  # the waterfall: some blocks fail
        bb, ok, _, _, _ = H.ref_t2_job(L, cod, mod, fec_type, rot, (0, 0), cells, csi, 15)
        assert np.array_equal(bb, cw[:, :bb.shape[1]]), name
        assert np.array_equal(ok, (res >= 0).astype(np.uint8)), name


def test_unsupported_plp():
    for cod, mod, ft in [(6, 0, 1), (0, 4, 1), (0, 0, 2), (-1, 0, 0)]:
        assert H.oracle().oracle_t2_geometry(cod, mod, ft, None, None, None, 0) == -1


@pytest.mark.parametrize("til", [(0, 0), (3, 0), (5, 1), (7, 0)])
def test_library_maps_match_oracle(til):
    """The maps the HIP library's kernel uses (ldpc_gpu_t2_maps, computed
    on the host: no device needed) equal the oracle's, for every PLP
    configuration."""
    for fec_type, mod, cod in CONFIGS:
        plp = H.ldpcgpu.T2Plp(cod, mod, fec_type, 1, til[0], til[1])
        perm, src, shift = H.ldpcgpu.t2_maps(plp, 23)
        N, nc, _ = H.t2_geometry(cod, mod, fec_type)
        assert H.ldpcgpu.t2_geometry(plp)[0] == N and H.ldpcgpu.t2_geometry(plp)[3] == nc
        op = np.zeros(nc, np.int32)
        H.oracle().oracle_t2_cellperm(nc, op.ctypes.data)
        osrc = np.zeros(N, np.uint16)
        H.oracle().oracle_t2_bitsrc(cod, mod, fec_type, osrc.ctypes.data)
        osh = np.zeros(23, np.int32)
        H.oracle().oracle_t2_shift(23, til[0], til[1], nc, osh.ctypes.data)
        assert np.array_equal(perm, op) and np.array_equal(src, osrc) and np.array_equal(shift, osh), \
            (fec_type, mod, cod)


def test_saturated_and_zero_state_match_reference():
    """The same extreme cells as the GPU test: saturating soft values, exact
    zeros, channel state 0 -- the oracle equals the reference's build."""
    if not REFS:
        pytest.skip("reference T2 build (oracle/_ref) not present")
    cod, mod, ft = 3, 2, 0
    N, nc, _ = H.t2_geometry(cod, mod, ft)
    rng = np.random.default_rng(8)
    cells, csi = H.t2_random_cells(mod, nc, 3, rng, snr_db=30.0)
    cells[0] *= 40.0
    cells[1, :100] = 0.0
    csi[2, ::7] = 0.0
    for rot in (0, 1):
        llr, err, sig = H.oracle_t2_demap(cod, mod, ft, rot, (0, 0), cells, csi)
        for name, (L, lanes) in REFS.items():
            _, _, rerr, rsig, rllr = H.ref_t2_job(L, cod, mod, ft, rot, (0, 0), cells, csi, 1, with_llr=True)
            assert np.array_equal(rllr, llr), (name, rot)
            assert np.array_equal(rerr.view(np.int32), err.view(np.int32)), (name, rot)
            assert np.array_equal(rsig.view(np.int32), sig.view(np.int32)), (name, rot)


GOLDEN = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "t2_golden.json")))


@pytest.mark.parametrize("k", range(len(GOLDEN["front_end"])))
def test_oracle_front_end_matches_golden(k):
    """The reference's soft values and MER sums, as digests written by
    tests/golden/make_t2_golden.py from its own decode_job: the oracle
    reproduces them from the same seeded cells (no reference build
    needed)."""
    import make_t2_golden as G

    g = GOLDEN["front_end"][k]
    cod, mod, ft, rot, tl, tt, nb, snr, seed = g["case"]
    cells, csi = G.front_inputs(g["case"])
    llr, err, sig = H.oracle_t2_demap(cod, mod, ft, rot, (tl, tt), cells, csi)
    assert G.sha(llr) == g["llr_sha256"] and G.sha(err) == g["err_sha256"] and G.sha(sig) == g["sig_sha256"]


@pytest.mark.parametrize("k", range(len(GOLDEN["jobs"])))
def test_oracle_job_matches_golden(k):
    import make_t2_golden as G

    g = GOLDEN["jobs"][k]
    cod, mod, ft, rot, tl, tt, nb, snr, seed = g["case"]
    info, cells, csi = G.job_inputs(g["case"])
    cw, res, err, sig = H.oracle_t2_job(cod, mod, ft, rot, (tl, tt), cells, csi, g["max_iter"], g["lanes"])
    kb = info.shape[1]
    assert G.sha(np.ascontiguousarray(cw[:, :kb])) == g["bb_sha256"]
    assert (res >= 0).astype(int).tolist() == g["ok"]
    assert G.sha(err) == g["err_sha256"] and G.sha(sig) == g["sig_sha256"]
