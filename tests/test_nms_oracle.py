"""Pins of the learned (neural) offset min-sum in the CPU oracle.

* the learned check node against its definition written out by brute
  force, and equal to the reference's node at S = 16, B = 1;
* the float training model: at alpha = beta = 1 on integer LLRs its
  forward is the reference decoder (every posterior, no early stop); in
  quantised mode its forward is the integer learned decoder;
* the gradient of the float model against central finite differences,
  for every parameter the code uses, with planted slips shown to fail;
* the learned decode with plain fallback: frames that fail the learned
  iterations are decoded again by the plain decoder.
"""
import numpy as np
import pytest

from helpers import NMS_DEGS, Oracle, check_node, codewords, gen, oracle


def sat8(v):
    return max(-128, min(127, v))


def textbook_nms_node(L, m, first, S, B):
    d = len(L)
    x = [sat8(int(L[e]) - (0 if first else int(m[e]))) for e in range(d)]
    Lo, mo = [], []
    for e in range(d):
        others = [x[f] for f in range(d) if f != e]
        mn = min(min(abs(v) for v in others), 127)
        mag = max(0, min(127, (mn * S - B) >> 4))  # Python's >> floors, as the arithmetic shift
        o = -mag if sum(1 for v in others if v < 0) % 2 else mag
        Lo.append(sat8(x[e] + o))
        mo.append(max(-32, min(31, o)))
    return np.array(Lo, np.int8), np.array(mo, np.int8)


def nms_node(deg, L, m, first, S, B):
    Lb = np.ascontiguousarray(L, np.int8)
    mb = np.ascontiguousarray(m, np.int8)
    Lo = np.zeros(deg, np.int8)
    mo = np.zeros(deg, np.int8)
    oracle().oracle_nms_check_node(deg, Lb.ctypes.data, mb.ctypes.data, int(first), S, B, Lo.ctypes.data,
                                   mo.ctypes.data)
    return Lo, mo


@pytest.mark.parametrize("deg", [3, 5, 8, 11, 16, 30])
def test_learned_node_definition(deg):
    rng = np.random.default_rng(deg)
    for trial in range(200):
        L = rng.integers(-128, 128, deg)
        m = rng.integers(-32, 32, deg)
        S = int(rng.integers(0, 64))
        B = int(rng.integers(-1024, 1024)) if trial % 3 == 0 else int(rng.integers(-40, 80))
        first = trial % 4 == 0
        got = nms_node(deg, L, m, first, S, B)
        want = textbook_nms_node(L, m, first, S, B)
        assert (got[0] == want[0]).all() and (got[1] == want[1]).all(), (L, m, S, B)
        if trial % 5 == 0:
            ref = check_node(deg, L, m, first)
            one = nms_node(deg, L, m, first, 16, 16)
            assert (ref[0] == one[0]).all() and (ref[1] == one[1]).all()


def test_quantiser():
    L = oracle()
    assert L.oracle_nms_scale_q(1.0) == 16 and L.oracle_nms_offset_q(1.0) == 16
    assert L.oracle_nms_scale_q(0.78) == 12 and L.oracle_nms_scale_q(-1.0) == 0 and L.oracle_nms_scale_q(40.0) == 63
    assert L.oracle_nms_offset_q(0.5 / 16) == 0 and L.oracle_nms_offset_q(1.5 / 16) == 2  # round half to even
    assert L.oracle_nms_offset_q(-300.0) == -1024 and L.oracle_nms_offset_q(0.7) == 11


def _frames(short, rate, esn0, n, seed):
    fec, llr, ref = gen(short, rate, esn0, n, seed)
    return fec, llr, codewords(short, rate, ref)


@pytest.mark.parametrize("mode", [0, 1])
def test_float_model_at_unit_parameters_is_the_reference(mode):
    fec, llr, cw = _frames(1, "1_2", 1.0, 2, 71)
    o = Oracle(1, "1_2", fec.kldpc)
    T = 4
    ones = np.ones((T, NMS_DEGS), np.float32)
    loss, _, _, post = o.nms_train(mode, llr.astype(np.float32), cw, ones, ones, np.ones(T), grad=False)
    ref = o.run(llr, T)
    assert (post == ref.astype(np.float32)).all()
    assert np.isfinite(loss) and loss > 0


def test_quantised_forward_is_the_integer_learned_decoder():
    """At low Es/N0 no frame converges in T iterations, so the learned
    decoder runs exactly T iterations: its bits must be the signs of the
    quantised model's posteriors."""
    fec, llr, cw = _frames(1, "1_2", -1.0, 3, 72)
    o = Oracle(1, "1_2", fec.kldpc)
    T = 5
    rng = np.random.default_rng(3)
    alpha = rng.uniform(0.6, 1.2, (T, NMS_DEGS)).astype(np.float32)
    beta = rng.uniform(-0.5, 2.5, (T, NMS_DEGS)).astype(np.float32)
    S = np.array([[o.L.oracle_nms_scale_q(float(a)) for a in r] for r in alpha], np.uint8)
    B = np.array([[o.L.oracle_nms_offset_q(float(b)) for b in r] for r in beta], np.int16)
    bits, it, ok, fb = o.nms_frames(S, B, llr, T, 0)
    assert (it == T).all() and not ok.any() and not fb.any()
    _, _, _, post = o.nms_train(1, llr.astype(np.float32), cw, alpha, beta, np.ones(T), grad=False)
    assert (post == np.rint(post)).all()
    mine = np.packbits((post[:, :fec.kldpc] < 0).astype(np.uint8), axis=1)
    assert (mine == bits).all()


def _fd_setup():
    fec, llr, cw = _frames(1, "1_2", 0.8, 2, 73)
    rng = np.random.default_rng(4)
    # continuous inputs: no ties among magnitudes, no saturation
    llrf = (llr.astype(np.float32) * 0.5 + rng.uniform(-0.5, 0.5, llr.shape)).astype(np.float32)
    T = 3
    alpha = rng.uniform(0.7, 0.95, (T, NMS_DEGS)).astype(np.float32)
    beta = rng.uniform(0.2, 0.9, (T, NMS_DEGS)).astype(np.float32)
    w = np.array([0.3, 0.5, 1.0], np.float32)
    return Oracle(1, "1_2", fec.kldpc), llrf, cw, alpha, beta, w


def test_gradient_matches_finite_differences():
    o, llrf, cw, alpha, beta, w = _fd_setup()
    loss, ga, gb, _ = o.nms_train(0, llrf, cw, alpha, beta, w)
    used = sorted({(t, d) for t in range(alpha.shape[0]) for d in range(NMS_DEGS) if ga[t, d] != 0})
    assert len(used) >= 2 * alpha.shape[0]  # two check degrees in this code
    h = 2e-3
    for t, d in used:
        for name, P, G in (("alpha", alpha, ga), ("beta", beta, gb)):
            P2 = P.copy()
            P2[t, d] += h
            lp = o.nms_train(0, llrf, cw, *(([P2, beta] if name == "alpha" else [alpha, P2])), w, grad=False)[0]
            P2[t, d] -= 2 * h
            lm = o.nms_train(0, llrf, cw, *(([P2, beta] if name == "alpha" else [alpha, P2])), w, grad=False)[0]
            fd = (lp - lm) / (2 * h)
            assert abs(fd - G[t, d]) <= 0.03 * abs(fd) + 2e-6, (name, t, d, fd, G[t, d])


def test_gradient_check_catches_planted_slips():
    """Perturbing the analytic gradient as a plausible slip would (sign
    of the offset term, a dropped factor, the last iteration's loss
    weight ignored) breaks the agreement with finite differences."""
    o, llrf, cw, alpha, beta, w = _fd_setup()
    loss, ga, gb, _ = o.nms_train(0, llrf, cw, alpha, beta, w)
    t, d = max(((t, d) for t in range(3) for d in range(NMS_DEGS)), key=lambda k: abs(gb[k]))
    h = 2e-3
    P2 = beta.copy()
    P2[t, d] += h
    lp = o.nms_train(0, llrf, cw, alpha, P2, w, grad=False)[0]
    P2[t, d] -= 2 * h
    lm = o.nms_train(0, llrf, cw, alpha, P2, w, grad=False)[0]
    fd = (lp - lm) / (2 * h)
    assert abs(fd - gb[t, d]) <= 0.03 * abs(fd)
    assert abs(fd - (-gb[t, d])) > 0.5 * abs(fd)  # sign slip
    w2 = w.copy()
    w2[-1] = 0
    _, _, gb2, _ = o.nms_train(0, llrf, cw, alpha, beta, w2)
    assert abs(fd - gb2[t, d]) > 0.1 * abs(fd)  # a loss term dropped


def test_learned_decode_with_fallback():
    fec, llr, _ = gen(0, "1_2", 1.5, 8, 74)
    o = Oracle(0, "1_2", fec.kldpc)
    S = np.full((10, NMS_DEGS), 16, np.uint8)
    B = np.full((10, NMS_DEGS), 16, np.int16)
    # unit parameters + fallback: the learned phase is the plain decoder,
    # so a frame either converges within 10 or restarts and runs plain
    bits, it, ok, fb = o.nms_frames(S, B, llr, 10, 50)
    pbits, pit, pok = o.frames(llr, 50)
    for f in range(len(llr)):
        if pit[f] <= 10 and pok[f]:
            assert not fb[f] and it[f] == pit[f] and (bits[f] == pbits[f]).all()
        else:
            assert fb[f] and it[f] == 10 + pit[f] and ok[f] == pok[f] and (bits[f] == pbits[f]).all()
    # without fallback the failing frames stop at the learned cap
    _, it2, ok2, fb2 = o.nms_frames(S, B, llr, 4, 0)
    assert not fb2.any() and (it2 <= 4).all()
