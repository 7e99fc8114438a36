"""Golden fixtures (tests/golden/): the frame source, the CPU oracle and the
HIP decoder against the reference decoder's own outputs.

golden.json holds, per set of seeded frames, the SHA-256 of every frame's
LLRs, the reference's per-frame decode (nframes = 1) and its batch decode
of the whole set (one shared iteration count); vectors.npz the full LLRs
and decoded bits of two frames of two sets.  tests/golden/make_golden.py
wrote them from the reference decoder built from its own sources.
"""
import json
import os

import numpy as np
import pytest

from helpers import Oracle, gen, ldpcgpu, sha256

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "golden.json")))
VECTORS = np.load(os.path.join(HERE, "golden", "vectors.npz"))
SETS = {s["name"]: s for s in GOLDEN["sets"]}


def _frames(s):
    fec, llr, ref = gen(s["short_frame"], s["rate"], s["esn0_db"], s["frames"], s["seed"], pls=s["pls"])
    assert fec.kldpc == s["kldpc"]
    return fec, llr, ref


def test_fixture_coverage():
    """The sets cover converging, non-converging and batch-stop cases."""
    results = [g["result"] for s in GOLDEN["sets"] for g in s["groups"]]
    assert -1 in results and any(r > 0 for r in results)
    oks = [p["syndrome_ok"] for s in GOLDEN["sets"] for p in s["per_frame"]]
    assert 0 in oks and 1 in oks
    assert {s["short_frame"] for s in GOLDEN["sets"]} == {0, 1}
    assert any(s["rate"].endswith("_t2") for s in GOLDEN["sets"])


@pytest.mark.parametrize("name", sorted(SETS))
def test_generator_matches_fixture(name):
    """The frame source is deterministic: same seed, same LLRs."""
    s = SETS[name]
    _, llr, _ = _frames(s)
    assert [sha256(llr[f]) for f in range(s["frames"])] == [p["llr_sha256"] for p in s["per_frame"]]


@pytest.mark.parametrize("name", sorted(SETS))
def test_oracle_per_frame_matches_reference(name):
    s = SETS[name]
    fec, llr, ref = _frames(s)
    o = Oracle(s["short_frame"], s["rate"], fec.kldpc)
    bits, it, ok = o.frames(llr, s["max_iter"])
    for f, p in enumerate(s["per_frame"]):
        assert (int(it[f]), int(ok[f])) == (p["iterations"], p["syndrome_ok"]), (name, f)
        assert sha256(bits[f]) == p["bits_sha256"], (name, f)
        assert bool((bits[f] == ref[f]).all()) == p["bits_equal_sent"], (name, f)


@pytest.mark.parametrize("name", sorted(SETS))
def test_oracle_batch_matches_reference(name):
    s = SETS[name]
    fec, llr, _ = _frames(s)
    o = Oracle(s["short_frame"], s["rate"], fec.kldpc)
    for g in s["groups"]:
        r, bits = o.group(llr[g["first"]:g["first"] + g["count"]], s["max_iter"])
        assert r == g["result"], name
        assert [sha256(b) for b in bits] == g["bits_sha256"], name


@pytest.mark.parametrize("name", [k[:-4] for k in VECTORS.files if k.endswith("_llr")])
def test_vectors(name):
    """Full-content vectors: the stored LLRs are the generator's, and the
    oracle's batch decode of the two frames gives the stored bits."""
    s = SETS[name]
    fec, llr, _ = _frames(s)
    assert (VECTORS[name + "_llr"] == llr[:2]).all()
    o = Oracle(s["short_frame"], s["rate"], fec.kldpc)
    r, bits = o.group(VECTORS[name + "_llr"], s["max_iter"])
    assert r == int(VECTORS[name + "_group2_result"][0])
    assert (bits == VECTORS[name + "_group2_bits"]).all()


# ---------------------------------------------------------------- GPU ----

@pytest.fixture(scope="module")
def dec():
    return ldpcgpu.Decoder(0)


def _decode(dec, s, llr, group):
    code = ldpcgpu.code_lookup(s["short_frame"], s["rate"])
    n = llr.shape[1]
    kb = (s["kldpc"] + 7) // 8
    frames = np.zeros(len(llr), ldpcgpu.FRAME_DTYPE)
    frames["code"] = code
    frames["max_iter"] = s["max_iter"]
    frames["nbits"] = s["kldpc"]
    frames["group"] = group
    frames["llr_off"] = np.arange(len(llr)) * n
    frames["out_off"] = np.arange(len(llr)) * kb
    return dec.decode_host(frames, llr.reshape(-1), len(llr) * kb)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(SETS))
def test_gpu_per_frame_matches_reference(dec, name):
    s = SETS[name]
    _, llr, _ = _frames(s)
    out, it, ok = _decode(dec, s, llr, -1)
    out = out.reshape(len(llr), -1)
    for f, p in enumerate(s["per_frame"]):
        assert (int(it[f]), int(ok[f])) == (p["iterations"], p["syndrome_ok"]), (name, f)
        assert sha256(out[f]) == p["bits_sha256"], (name, f)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(SETS))
def test_gpu_batch_matches_reference(dec, name):
    s = SETS[name]
    _, llr, _ = _frames(s)
    for g in s["groups"]:
        gl = llr[g["first"]:g["first"] + g["count"]]
        out, it, ok = _decode(dec, s, gl, 0)
        out = out.reshape(len(gl), -1)
        r = int(it[0]) if ok.all() else -1
        assert r == g["result"], name
        assert (it == it[0]).all()
        assert [sha256(b) for b in out] == g["bits_sha256"], name
