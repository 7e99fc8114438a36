"""The streaming loop on the GPU (dvb-rx_amd/streaming.py): both schedulers
decode every frame of a short trace, and decoding is deterministic, so the
live scheduler with two contexts (frames renumbered into per-launch-kind
queues) must report the same iteration counts and frame errors as the
one-batch-at-a-time scheduler.  The trace's DVB-T2 blocks come through the
GPU T2 front end; every frame is error-free after BCH."""
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dvb-rx_amd"))

import ldpcgpu  # noqa: E402
import streaming  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("learned", [False, True])
def test_live_matches_simulated(learned):
    print(f"# streaming live vs simulated, learned={learned}", flush=True)
    dec = ldpcgpu.Decoder(0)
    dec2 = ldpcgpu.Decoder(0)
    try:
        tabs = streaming.learned_tables() if learned else None
        sim = streaming.run(dec, 0.05, 1.0, learned, tabs, seed=17)
        live = streaming.run(dec, 0.05, 1.0, learned, tabs, seed=17, lanes=[dec, dec2])
    finally:
        dec2.close()
        dec.close()
    assert sim["frames"] == live["frames"] > 100
    assert live["avg_iterations"] == sim["avg_iterations"]
    # the T2 256-QAM frames leave residual LDPC errors that BCH corrects
    assert live["frame_errors_ldpc"] == sim["frame_errors_ldpc"]
    assert live["frame_errors_after_bch"] == sim["frame_errors_after_bch"] == 0
    assert live["latency_ms"]["max"] < 1000 and live["latency_ms"]["p50"] > 0
