"""Launch shapes (ldpc_gpu_set_launch_shape, ldpc_gpu_plan_shape): the
kernel instances built for two workgroups per CU decode short frames to the
same bits, iteration counts and syndrome flags as the one-workgroup build
and as the oracle; grid scales change nothing either; the plan autotuner's
per-launch search keeps replays equal to the eager decode."""
import numpy as np
import pytest

from helpers import Oracle, gen, ldpcgpu

pytestmark = pytest.mark.gpu

# short codes on the instances with a two-workgroup build (layer degree
# bounds 2, 4 and 106), one short code without one (10) and a normal code on
# instance 2, near their thresholds so that iteration counts spread
SETS = [(1, "1_4", -2.0), (1, "1_3", -1.0), (1, "2_5", -0.2), (1, "1_2", 1.3), (1, "3_5", 2.6), (0, "1_4", -2.3)]
OCC2 = (2, 4, 6, 106)
MAX_ITER = 40


def _batch(seed, per):
    rows, frames, parts = [], [], []
    off = 0
    for k, (short, rate, esn0) in enumerate(SETS):
        fec, llr, _ = gen(short, rate, esn0, per, seed + k)
        code = ldpcgpu.code_lookup(short, rate)
        parts.append((short, rate, fec, llr, len(frames), off))
        for f in range(per):
            slot = np.zeros(64800, np.int8)
            slot[:llr.shape[1]] = llr[f]
            rows.append(slot)
            frames.append((code, MAX_ITER, fec.kldpc, -1, 64800 * len(frames), off))
            off += (fec.kldpc + 7) // 8
    return np.stack(rows), np.array(frames, dtype=ldpcgpu.FRAME_DTYPE), off, parts


class Bufs:
    def __init__(self, dec, llr, nout, nf):
        self.dec = dec
        self.llr = dec.malloc(llr.nbytes)
        self.out = dec.malloc(nout)
        self.it = dec.malloc(4 * nf)
        self.ok = dec.malloc(nf)
        self.nout, self.nf = nout, nf
        dec.upload(self.llr, llr)

    def clear(self):
        self.dec.upload(self.out, np.zeros(self.nout, np.uint8))
        self.dec.upload(self.it, np.full(self.nf, -7, np.int32))
        self.dec.upload(self.ok, np.full(self.nf, 9, np.uint8))

    def results(self):
        out = np.zeros(self.nout, np.uint8)
        it = np.zeros(self.nf, np.int32)
        ok = np.zeros(self.nf, np.uint8)
        self.dec.download(out, self.out)
        self.dec.download(it, self.it)
        self.dec.download(ok, self.ok)
        return out, it, ok

    def free(self):
        for p in (self.llr, self.out, self.it, self.ok):
            self.dec.free(p)


@pytest.fixture()
def dec():
    d = ldpcgpu.Decoder(0)
    yield d
    d.close()


def _same(want, got, what):
    for name, w, g in zip(("bits", "iterations", "syndrome"), want, got):
        assert (w == g).all(), (what, name, np.flatnonzero(w != g)[:8])


def test_launch_shape_rejects_what_is_not_built(dec):
    for args in ((999, 1, 1, 1.0), (10, 1, 2, 1.0), (2, 1, 3, 1.0), (2, 1, 1, 0.0), (2, 1, 1, 1.5), (2, 2, 1, 1.0)):
        with pytest.raises(RuntimeError):
            dec.set_launch_shape(*args)
    for di in OCC2:
        dec.set_launch_shape(di, 1, 2, 1.0)


def test_two_workgroups_per_cu_match_oracle_and_one(dec):
    llr, fr, nout, parts = _batch(700, 24)
    b = Bufs(dec, llr, nout, len(fr))
    try:
        b.clear()
        dec.decode_batch(fr, b.llr, b.out, b.it, b.ok)
        dec.synchronize()
        want = b.results()
        # the oracle on every frame of the batch
        for short, rate, fec, fl, f0, o0 in parts:
            bits, it, ok = Oracle(short, rate, fec.kldpc).frames(fl, MAX_ITER)
            kb = bits.shape[1]
            for f in range(fl.shape[0]):
                assert (want[0][o0 + f * kb:o0 + (f + 1) * kb] == bits[f]).all(), (rate, f)
            assert (want[1][f0:f0 + fl.shape[0]] == it).all(), rate
            assert (want[2][f0:f0 + fl.shape[0]] == ok).all(), rate
        assert want[1].max() > want[1].min() + 5, "iteration counts should spread"
        for occ, scale in ((1, 1.0), (2, 1.0), (2, 0.5), (1, 0.5), (2, 0.75)):
            for di in OCC2:
                dec.set_launch_shape(di, 1, occ, scale)
            dec.set_launch_shape(2, 0, 1, scale)
            b.clear()
            dec.decode_batch(fr, b.llr, b.out, b.it, b.ok)
            dec.synchronize()
            _same(want, b.results(), ("eager", occ, scale))
            dis = {(d, n) for d, n, _ in dec.last_launches()}
            assert {d for d, _ in dis} >= {2, 4, 106, 10}
            b.clear()
            dec._check(dec.L.ldpc_gpu_decode_batch_static(dec.ctx, fr.ctypes.data, len(fr), b.llr, b.out, b.it, b.ok),
                       "decode_batch_static")
            dec.synchronize()
            _same(want, b.results(), ("static", occ, scale))
    finally:
        b.free()


def test_plan_per_launch_tuning_matches_eager(dec):
    llr, fr, nout, _ = _batch(900, 40)
    b = Bufs(dec, llr, nout, len(fr))
    ref = Bufs(dec, llr, nout, len(fr))
    plan = None
    try:
        ref.clear()
        dec.decode_batch(fr, ref.llr, ref.out, ref.it, ref.ok)
        dec.synchronize()
        want = ref.results()
        plan = ldpcgpu.Plan(dec, fr, b.llr, b.out, b.it, b.ok)
        plan.tune(2)
        info = plan.info()
        assert -1 <= info["variant"] < len(info["variants"])
        shape = info["launches"]
        assert sum(s["frames"] for s in shape) == len(fr)
        for s in shape:
            assert s["occ"] in (1, 2) and (s["occ"] == 1 or (s["di"] in OCC2 and s["short"])), s
            assert s["grid_scale"] in (1.0, 0.75, 0.5) and 1 <= s["grid"] <= (s["frames"] + 1) // 2, s
        # the default shape (two workgroups per CU on the short launches
        # built for it), then the tuned one again
        plan2 = ldpcgpu.Plan(dec, fr, b.llr, b.out, b.it, b.ok)
        try:
            occ2 = [s for s in plan2.shape() if s["occ"] == 2]
            assert {s["di"] for s in occ2} == {2, 4, 106} and all(s["short"] for s in occ2)
            for p in (plan2, plan, plan2):
                for rep in range(2):
                    b.clear()
                    p.launch()
                    dec.synchronize()
                    _same(want, b.results(), ("plan", rep))
                    dec.decode_batch(fr, ref.llr, ref.out, ref.it, ref.ok)
                    dec.synchronize()
        finally:
            plan2.close()
    finally:
        if plan:
            plan.close()
        b.free()
        ref.free()


def test_launch_stream_counts_give_the_same_results(dec):
    """ldpc_gpu_set_concurrent: serial, 2, 3 (the default) and 4 forked
    streams decode the batch identically; out-of-range counts rejected."""
    llr, fr, nout, _ = _batch(1100, 12)
    b = Bufs(dec, llr, nout, len(fr))
    try:
        b.clear()
        dec.decode_batch(fr, b.llr, b.out, b.it, b.ok)
        dec.synchronize()
        want = b.results()
        for n in (0, 2, 4, 1, 3):
            dec.set_concurrent(n)
            b.clear()
            dec.decode_batch(fr, b.llr, b.out, b.it, b.ok)
            dec.synchronize()
            _same(want, b.results(), ("streams", n))
        for n in (-1, 5):
            with pytest.raises(RuntimeError):
                dec.set_concurrent(n)
    finally:
        dec.set_concurrent(True)
        b.free()
