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


# ---- batches generated on the GPU ----------------------------------------

S2X_NORMAL_HIGH = [0xb8, 0xba, 0xbe, 0xc2, 0xc6, 0xc8, 0xca, 0xcc, 0xce, 0xd0, 0xd2, 0xd4, 0xd6]
S2X_SHORT_LOW = [0xd8, 0xda, 0xdc, 0xde, 0xe0, 0xe2, 0xe4, 0xe6, 0xe8, 0xea]


class GpuBatch:
    """Frames generated on the device: per MODCOD a pool of codewords (host
    encoder) sent through ldpc_gpu_channel with fresh noise.  Frames of a
    MODCOD are contiguous, each in its own N-byte slot."""

    def __init__(self, dec, groups, margin_db, max_iter=50, pool_n=16, seed=11):
        from nms_train import codeword_pool

        self.dec = dec
        self.groups = groups
        self.ptrs = []
        frames, refs, kls = [], [], []
        layout = []
        off = 0
        out_off = 0
        rng = np.random.default_rng(seed)
        for g, (pls, count) in enumerate(groups):
            m = ldpcgpu.pls_lookup(pls)
            fec = ldpcgpu.fec_lookup(m["short_frame"], m["rate"])
            code = ldpcgpu.code_lookup(m["short_frame"], m["rate"])
            pool = codeword_pool(m["short_frame"], m["rate"], pool_n, seed + g)
            rows = rng.integers(0, pool_n, count).astype(np.int32)
            layout.append((pls, code, fec.nldpc, off, count, pool, rows, ldpcgpu.qef_db(pls) + margin_db))
            kb = (fec.kldpc + 7) // 8
            ref = np.packbits(pool[:, :fec.kldpc], axis=1)
            for f in range(count):
                frames.append((code, max_iter, fec.kldpc, -1, off + f * fec.nldpc, out_off))
                refs.append(ref[rows[f]])
                kls.append(fec.kldpc)
                out_off += kb
            off += count * fec.nldpc
            off = (off + 7) // 8 * 8
        self.frames = np.array(frames, dtype=ldpcgpu.FRAME_DTYPE)
        self.nframes = len(frames)
        self.llr_bytes = off
        self.out_bytes = out_off
        self.refs = refs
        self.kldpc = np.array(kls, np.int64)
        self.info_bits = int(self.kldpc.sum())
        self.d_llr = self._alloc(max(off, 8))
        self.d_out = self._alloc(max(out_off, 1))
        self.d_it = self._alloc(4 * self.nframes)
        self.d_ok = self._alloc(self.nframes)
        for pls, code, N, base, count, pool, rows, esn0 in layout:
            d_pool = self._put(pool)
            d_rows = self._put(rows)
            d_e = self._put(np.full(count, esn0, np.float32))
            ldpcgpu.channel(pls, N, d_pool, d_rows, count, d_e, seed * 131 + pls, self.d_llr + base,
                            stream=dec.stream)
        dec.synchronize()
        self.layout = [(pls, code, N, base, count, esn0) for pls, code, N, base, count, _, _, esn0 in layout]

    def _alloc(self, n):
        p = self.dec.malloc(n)
        self.ptrs.append(p)
        return p

    def _put(self, a):
        a = np.ascontiguousarray(a)
        p = self._alloc(a.nbytes)
        self.dec.upload(p, a)
        return p

    def results(self):
        out = np.zeros(self.out_bytes, np.uint8)
        it = np.zeros(self.nframes, np.int32)
        ok = np.zeros(self.nframes, np.uint8)
        self.dec.download(out, self.d_out)
        self.dec.download(it, self.d_it)
        self.dec.download(ok, self.d_ok)
        return out, it, ok

    def frame_errors(self, out, t_of=None):
        """Frames whose decoded kldpc bits differ from the codeword (any bit)."""
        bad = 0
        for f in range(self.nframes):
            o = int(self.frames["out_off"][f])
            ref = self.refs[f]
            bad += int(not np.array_equal(out[o:o + len(ref)], ref))
        return bad

    def free(self):
        for p in self.ptrs:
            self.dec.free(p)
        self.ptrs = []


def s2x_high_order(dec, n_normal=16384, n_short=4096, margin_db=0.2, max_iter=50, seed=11):
    """DVB-S2X high-order batch: normal frames spread evenly over the
    64/128/256APSK MODCODs (and the 4+12+20+28 / 8+16+20+20 APSK ones of
    those rates), short frames over the QPSK / 8PSK ones; AWGN at the
    MODCOD's threshold + margin_db."""
    groups = []
    for lst, n in ((S2X_NORMAL_HIGH, n_normal), (S2X_SHORT_LOW, n_short)):
        base, extra = divmod(n, len(lst))
        for k, pls in enumerate(lst):
            groups.append((pls, base + (1 if k < extra else 0)))
    return GpuBatch(dec, groups, margin_db, max_iter, seed=seed)


T2_RATES = ["1_2", "3_5", "2_3", "3_4", "4_5", "5_6"]
T2_ROT_DEG = [29.0, 16.8, 8.6, 3.576334375]


def t2_blocks(plp, nblocks, snr_db, seed=3, distinct=16):
    """Synthetic DVB-T2 FEC blocks of one PLP (ldpcgpu.T2Plp) as the
    receiver's time deinterleaver hands them over: `distinct` random
    BBFRAMEs through the BCH and LDPC encoders, mapped to Gray QAM
    (codeword bit n to the (cell, plane) of ldpc_gpu_t2_maps' bitsrc;
    bit 0 of an axis its sign, bit j > 0 the sign of |level| - 2^(nb-j)),
    rotated with the Q part sent one cell later, cell-interleaved with the
    block's shift, through AWGN at snr_db (Es/N0), channel state 1; then
    repeated to nblocks (block b is frame b % distinct).  Returns (cells
    (nblocks, ncells, 2) float32, csi (nblocks, ncells) float32, info bytes
    (nblocks, kbch / 8), BCH codewords (distinct, kldpc / 8) packed: what
    the LDPC decoder returns for a decoded block)."""
    N, K, kbch, nc = ldpcgpu.t2_geometry(plp)
    sf = plp.fec_type == 0
    rate = T2_RATES[plp.cod]
    code = "2_3_t2" if (not sf and plp.cod == 2) else "3_5_t2" if (sf and plp.cod == 1) else rate
    m = 2 * (plp.mod + 1)
    nb = m // 2
    perm, src, shift = ldpcgpu.t2_maps(plp, nblocks)
    L = ldpcgpu.sim_lib()
    rng = np.random.default_rng(seed)
    norm = np.sqrt({2: 2.0, 4: 10.0, 6: 42.0, 8: 170.0}[m])
    th = np.deg2rad(T2_ROT_DEG[plp.mod]) if plp.rotation else 0.0
    d = min(distinct, nblocks)
    cells = np.zeros((nblocks, nc, 2), np.float32)
    info = np.zeros((nblocks, kbch // 8), np.uint8)
    bchcw = np.zeros((d, K // 8), np.uint8)
    for b in range(d):
        bits_in = rng.integers(0, 2, kbch, dtype=np.uint8)
        bch = np.zeros(K, np.uint8)
        cw = np.zeros(N, np.uint8)
        if L.sim_bch_encode(sf, rate.encode(), bits_in.ctypes.data, bch.ctypes.data) or \
                L.sim_ldpc_encode(sf, code.encode(), bch.ctypes.data, cw.ctypes.data):
            raise RuntimeError(f"encoder failed for {rate}")
        bchcw[b] = np.packbits(bch)
        planes = np.zeros(nc * m, np.uint8)
        planes[src.astype(np.int64)] = cw
        planes = planes.reshape(nc, m)
        axis = []
        for a in (0, 1):
            bits = planes[:, a::2]
            u = np.where(bits[:, nb - 1] == 0, 1.0, -1.0)
            for j in range(nb - 1, 0, -1):
                u = (u + 2.0 ** (nb - j)) * np.where(bits[:, j - 1] == 0, 1.0, -1.0)
            axis.append(u)
        x = (axis[0] + 1j * axis[1]) / norm * np.exp(1j * th)
        y = x.real + 1j * (np.roll(x.imag, 1) if plp.rotation else x.imag)
        s = 10 ** (-snr_db / 20) / np.sqrt(2)
        y = y + s * (rng.standard_normal(nc) + 1j * rng.standard_normal(nc))
        for bb in range(b, nblocks, d):
            pos = (perm.astype(np.int64) + shift[bb]) % nc
            cells[bb, pos, 0] = y.real
            cells[bb, pos, 1] = y.imag
            info[bb] = np.packbits(bits_in)
    return cells, np.ones((nblocks, nc), np.float32), info, bchcw
