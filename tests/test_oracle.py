"""Pins of the CPU oracle (oracle/ldpc_oracle.c) beyond the golden sets:

* one check node against offset min-sum written from its definition (the
  minimum and the sign product taken over the other edges by brute force),
  on random and extreme int8 inputs, every degree the codes use;
* whole-frame invariants the mathematics fixes: a codeword decodes in zero
  iterations to its own bits, one flipped sign fails the syndrome and is
  corrected, all-zero LLRs never converge (a zero LLR fails every check),
  a batch stops no earlier than its slowest frame;
* the reference decoder built from its own sources (oracle/_ref), on every
  code, where that build exists.
"""
import numpy as np
import pytest

from helpers import ALL_RATES, Oracle, check_node, esn0_for, gen, ldpcgpu, ref_builds, ref_decode


def sat8(v):
    return max(-128, min(127, v))


def textbook_check_node(L, m, first):
    """Offset min-sum with offset 1 on saturating int8, from its
    definition: x_e = L_e - m_e; the new message of edge e has the sign
    product of the other edges' x (0 counts as positive) and magnitude
    max(min(min over the other edges of |x|, 127) - 1, 0); the posterior is
    x_e + message (saturated), the stored message is clamped to [-32, 31]."""
    d = len(L)
    x = [sat8(int(L[e]) - (0 if first else int(m[e]))) for e in range(d)]
    Lo, mo = [], []
    for e in range(d):
        others = [x[f] for f in range(d) if f != e]
        mag = max(min(min(abs(v) for v in others), 127) - 1, 0)
        neg = sum(1 for v in others if v < 0) % 2
        o = -mag if neg else mag
        Lo.append(sat8(x[e] + o))
        mo.append(max(-32, min(31, o)))
    return np.array(Lo, np.int8), np.array(mo, np.int8)


DEGREES = list(range(3, 33))


@pytest.mark.parametrize("deg", DEGREES)
@pytest.mark.parametrize("first", [0, 1])
def test_check_node_random(deg, first):
    rng = np.random.default_rng(1000 * deg + first)
    for trial in range(60):
        if trial % 3 == 0:
            L = rng.integers(-128, 128, deg)
        elif trial % 3 == 1:
            L = rng.integers(-6, 7, deg)  # many ties and zeros
        else:
            L = rng.choice([-128, -127, -1, 0, 1, 126, 127], deg)
        m = rng.integers(-32, 32, deg)
        got = check_node(deg, L, m, first)
        want = textbook_check_node(L, m, first)
        assert (got[0] == want[0]).all() and (got[1] == want[1]).all(), (deg, first, L, m)


def test_check_node_exhaustive_small():
    """Degree 3 over a grid of edge values (incl. -128, saturation, zero)."""
    vals = [-128, -127, -100, -33, -2, -1, 0, 1, 2, 31, 32, 100, 127]
    msgs = [-32, -1, 0, 31]
    for a in vals:
        for b in vals:
            for c in vals:
                for ma in msgs:
                    L = np.array([a, b, c])
                    m = np.array([ma, 0, -ma if ma > -32 else 5])
                    got = check_node(3, L, m, 0)
                    want = textbook_check_node(L, m, 0)
                    assert (got[0] == want[0]).all() and (got[1] == want[1]).all(), (L, m)


def _planted(L, m, first, no_offset=False, self_min=False, no_clamp=False, no_cap=False):
    """textbook_check_node with one plausible slip planted."""
    d = len(L)
    x = [sat8(int(L[e]) - (0 if first else int(m[e]))) for e in range(d)]
    Lo, mo = [], []
    for e in range(d):
        pool = x if self_min else [x[f] for f in range(d) if f != e]
        mn = min(abs(v) for v in pool)
        mag = max((mn if no_cap else min(mn, 127)) - (0 if no_offset else 1), 0)
        o = -mag if sum(1 for f in range(d) if f != e and x[f] < 0) % 2 else mag
        Lo.append(sat8(x[e] + o))
        mo.append(o if no_clamp else max(-32, min(31, o)))
    return np.array(Lo, np.int64), np.array(mo, np.int64)


def test_check_node_catches_planted_mistakes():
    """Designed inputs on which each plausible slip (no offset, minimum
    taken over all edges, message not clamped, |-128| not capped at 127)
    changes the result: the oracle matches the definition and none of the
    slips."""
    cases = [
        ([10, -4, 7, 20, -15], [0] * 5, 1),      # unique minimum, mixed signs
        ([100, 90, -80, 70], [3, -2, 1, 0], 0),  # magnitudes above the clamp
        ([-128, -128, -60], [0, 0, 0], 1),       # |-128| = 128
    ]
    slips = [dict(no_offset=True), dict(self_min=True), dict(no_clamp=True), dict(no_cap=True)]
    caught = set()
    for L, m, first in cases:
        got = check_node(len(L), L, m, first)
        want = textbook_check_node(L, m, first)
        assert (got[0] == want[0]).all() and (got[1] == want[1]).all()
        for i, slip in enumerate(slips):
            bad = _planted(L, m, first, **slip)
            if not ((bad[0] == got[0]).all() and (bad[1] == got[1]).all()):
                caught.add(i)
    assert caught == set(range(len(slips)))


CODES = [(s, r) for s in (0, 1) for r in ALL_RATES[s]]


def _codeword(short, rate, seed):
    """A noiseless frame: every LLR +-127 from a valid codeword."""
    fec, llr, ref = gen(short, rate, 1000.0, 1, seed)
    return fec, llr[0], ref[0]


@pytest.mark.parametrize("short,rate", CODES)
def test_codeword_invariants(short, rate):
    fec, llr, ref = _codeword(short, rate, 77)
    o = Oracle(short, rate, fec.kldpc)
    assert o.check(llr) == 1
    bits, it, ok = o.frames(llr[None], 10)
    assert it[0] == 0 and ok[0] == 1 and (bits[0] == ref).all()
    # one flipped information sign: the checks of that variable fail, and
    # messages of magnitude 126 from its intact neighbours restore it
    bad = llr.copy()
    bad[5] = -bad[5]
    assert o.check(bad) == 0
    bits, it, ok = o.frames(bad[None], 10)
    assert ok[0] == 1 and 1 <= it[0] <= 2 and (bits[0] == ref).all()


def test_all_zero_llrs_never_converge():
    fec = ldpcgpu.fec_lookup(1, "1_2")
    o = Oracle(1, "1_2", fec.kldpc)
    llr = np.zeros((2, fec.nldpc), np.int8)
    assert o.check(llr[0]) == 0
    bits, it, ok = o.frames(llr, 7)
    assert (it == 7).all() and (ok == 0).all() and (bits == 0).all()
    r, bits = o.group(llr, 7)
    assert r == -1 and (bits == 0).all()


def test_all_positive_is_the_zero_codeword():
    fec = ldpcgpu.fec_lookup(0, "3_5")
    o = Oracle(0, "3_5", fec.kldpc)
    llr = np.full((1, fec.nldpc), 1, np.int8)
    r, bits = o.group(llr, 5)
    assert r == 0 and (bits == 0).all()


def test_group_stops_together():
    """A batch iterates until every frame holds its checks: its count is
    at least each frame's own count, and a frame alone in the batch gives
    its own count."""
    fec, llr, _ = gen(1, "1_2", 1.6, 6, 31, pls=-1)
    o = Oracle(1, "1_2", fec.kldpc)
    _, it, ok = o.frames(llr, 40)
    assert ok.all()
    r, _ = o.group(llr, 40)
    assert r >= it.max()
    r1, _ = o.group(llr[:1], 40)
    assert r1 == it[0]


# ------------------------------------------ reference builds (oracle/_ref) --

REFS = ref_builds()
needs_ref = pytest.mark.skipif(not REFS, reason="oracle/_ref not built (needs the reference tree)")


@needs_ref
@pytest.mark.parametrize("short,rate", CODES)
def test_oracle_matches_reference_every_code(short, rate):
    """Noisy frames near each code's threshold, per frame and as a batch,
    against every reference build the host can run."""
    fec, llr, _ = gen(short, rate, esn0_for(short, rate), 3, 500 + ALL_RATES[short].index(rate))
    o = Oracle(short, rate, fec.kldpc)
    mi = 12
    bits, it, ok = o.frames(llr, mi)
    r, gbits = o.group(llr, mi)
    for name, R in REFS.items():
        for f in range(len(llr)):
            ri, rb = ref_decode(R, fec, rate, llr[f:f + 1], mi)
            assert (mi if ri < 0 else ri) == it[f] and int(ri >= 0) == ok[f], (name, f)
            assert (rb[0] == bits[f]).all(), (name, f)
        ri, rb = ref_decode(R, fec, rate, llr, mi)
        assert ri == r and (rb == gbits).all(), name


@needs_ref
@pytest.mark.parametrize("short,rate", [(0, "1_2"), (0, "3_4"), (0, "9_10"), (1, "1_4"), (1, "8_9"), (0, "2_3_t2")])
def test_fixed_iteration_run_matches_reference(short, rate):
    """oracle_ldpc_run (T updates, no early stop) against the reference
    build: well below threshold no frame converges, so the reference runs
    exactly T updates and its output bits are the signs of the oracle's
    posteriors after T."""
    fec, llr, _ = gen(short, rate, esn0_for(short, rate) - 3.0, 3, 900 + len(rate))
    o = Oracle(short, rate, fec.kldpc)
    T = 5
    post = o.run(llr, T)
    R = next(iter(REFS.values()))
    for f in range(len(llr)):
        ri, rb = ref_decode(R, fec, rate, llr[f:f + 1], T)
        assert ri == -1, "frame converged: the case needs a lower Es/N0"
        mine = np.packbits((post[f, :fec.kldpc] < 0).astype(np.uint8))
        assert (mine == rb[0]).all(), f
