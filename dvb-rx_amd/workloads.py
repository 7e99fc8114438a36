"""Synthetic frame batches of the benchmark configurations.

Every frame is random information bits from a seed, BCH + LDPC encoded by
lib/libdvbs2sim.so, mapped, sent through AWGN and demapped to int8 LLRs
(see include/dvbs2_sim.h).

  single_code  : n frames of one MODCOD (DVB-S2 normal 1/2 QPSK at 1.5 dB,
                 64 frames, is the reference's CPU-runnable case)
  mixed_s2     : 8192 frames, 75 % normal / 25 % short, rates uniform over
                 the DVB-S2 sets (normal 1/4..9/10, short 1/4..8/9), the
                 modulation uniform among the MODCODs of the rate (QPSK,
                 8PSK, 16APSK, 32APSK), Es/N0 = QEF + 0.3 dB, max 50
                 iterations, per-frame early stop
"""
import numpy as np

import ldpcgpu

S2_NORMAL = ["1_4", "1_3", "2_5", "1_2", "3_5", "2_3", "3_4", "4_5", "5_6", "8_9", "9_10"]
S2_SHORT = ["1_4", "1_3", "2_5", "1_2", "3_5", "2_3", "3_4", "4_5", "5_6", "8_9"]


def s2_modcods():
    """DVB-S2 (not S2X) PLS codes by (short_frame, rate)."""
    out = {}
    for pls in range(0, 0x80, 2):
        try:
            m = ldpcgpu.pls_lookup(pls)
        except KeyError:
            continue
        out.setdefault((m["short_frame"], m["rate"]), []).append(pls)
    return out


class Batch:
    """A generated batch: LLRs, reference BCH codewords, frame table."""

    def __init__(self, name, specs, seed, max_iter, group=-1):
        self.name = name
        self.specs = specs
        self.seed = seed
        self.max_iter = max_iter
        self.llr, self.ref, self.layout = ldpcgpu.generate(specs, seed)
        self.frames = ldpcgpu.frames_for(self.layout, max_iter, group)
        self.nframes = len(specs)
        self.kldpc = np.array([l[2] for l in self.layout], np.int64)
        self.n = np.array([l[1] for l in self.layout], np.int64)
        self.codes = np.array([l[0] for l in self.layout], np.int64)
        self.info_bits = int(self.kldpc.sum())

    def frame_llr(self, f):
        code, n, kldpc, lo, ro = self.layout[f]
        return self.llr[lo:lo + n]

    def frame_ref(self, f):
        code, n, kldpc, lo, ro = self.layout[f]
        return self.ref[ro:ro + (kldpc + 7) // 8]

    def edges(self):
        """Sum over frames of links_total (edges per iteration)."""
        cache = {}
        tot = 0
        for c in self.codes:
            if c not in cache:
                cache[c] = ldpcgpu.code_params(int(c))["links_total"]
            tot += cache[c]
        return tot


def single_code(nframes=64, pls=0x10, esn0_db=1.5, seed=2024, max_iter=50):
    specs = [dict(pls=pls, esn0_db=esn0_db, stream=f) for f in range(nframes)]
    return Batch(f"pls{pls:#04x}_x{nframes}_esn0_{esn0_db}dB", specs, seed, max_iter)


def mixed_s2_specs(nframes=8192, seed=7, margin_db=0.3, short_share=0.25):
    """Frame specs of the mixed DVB-S2 workload (no generation)."""
    rng = np.random.default_rng(seed)
    mc = s2_modcods()
    specs = []
    for f in range(nframes):
        short = int(rng.random() < short_share)
        rates = S2_SHORT if short else S2_NORMAL
        rate = rates[rng.integers(len(rates))]
        cands = mc[(short, rate)]
        pls = int(cands[rng.integers(len(cands))])
        specs.append(dict(pls=pls, esn0_db=ldpcgpu.qef_db(pls) + margin_db, stream=f))
    return specs


def mixed_s2(nframes=8192, seed=7, max_iter=50, margin_db=0.3, short_share=0.25):
    specs = mixed_s2_specs(nframes, seed, margin_db, short_share)
    return Batch(f"mixed_s2_x{nframes}", specs, seed, max_iter)
