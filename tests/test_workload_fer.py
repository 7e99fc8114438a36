"""Frame errors of the mixed DVB-S2 workload's QPSK 1/4 frames are the
reference decoder's own.

The mixed workload sends every MODCOD at its quasi-error-free Es/N0 + 0.3
dB (pl.c:68-87).  DVB-S2 normal QPSK 1/4 (PLS 0x04, -2.05 dB) is the one
MODCOD whose frames mostly fail there, on the GPU as in the oracle.  These
tests show on the CPU that the reference decoder itself (libdvbrx ldpc.c
built from its sources) fails the same frames, frame for frame with the
oracle, so the frame error rate the benchmark reports for that MODCOD
belongs to the reference's fixed-point chain (the demapper's LLR scale,
demap.c:32-34, and the offset of 1 per check, ldpc.c:353), not to this
decoder; and that the same frames decode with a larger margin.
"""
import pytest

from helpers import Oracle, gen, ldpcgpu, ref_builds, ref_decode

PLS = 0x04  # DVB-S2 normal 1/4 QPSK
NF = 16


def frames(margin_db):
    return gen(0, "1_4", ldpcgpu.qef_db(PLS) + margin_db, NF, 7, pls=PLS, stream0=100)


@pytest.mark.skipif(not ref_builds(), reason="the reference decoder is not built here (oracle/_ref)")
def test_reference_fails_the_qpsk_1_4_frames_of_the_workload():
    fec, llr, ref = frames(0.3)
    R = next(iter(ref_builds().values()))
    o_bits, o_it, o_ok = Oracle(0, "1_4", fec.kldpc).frames(llr, 50)
    errors = 0
    for f in range(NF):
        it, bits = ref_decode(R, fec, "1_4", llr[f:f + 1], 50)
        # the reference's result for a batch of one: its iterations, or
        # -1 when the frame still fails a check after 50
        assert (it >= 0) == bool(o_ok[f]), f
        assert (bits[0] == o_bits[f]).all(), f
        # a frame error as the benchmark counts it: decoded bits that are
        # not the sent ones
        errors += int((bits[0] != ref[f]).any())
    assert errors >= NF // 2, errors


def test_the_same_frames_decode_with_a_larger_margin():
    fec, llr, ref = frames(1.0)
    bits, it, ok = Oracle(0, "1_4", fec.kldpc).frames(llr, 50)
    assert ok.all() and (bits == ref).all()
