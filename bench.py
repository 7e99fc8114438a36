#!/usr/bin/env python3
"""Decoded info Gbit/s of the MI355X LDPC decoder on one GPU.

Workload (the metric's configuration): 8192 DVB-S2 frames in one call,
75 % normal / 25 % short, rates uniform over the DVB-S2 sets, modulation
uniform among the rate's MODCODs (QPSK/8PSK/16APSK/32APSK), Es/N0 = QEF +
0.3 dB, int8 max-log LLRs, max 50 iterations, per-frame early stop.  The
LLRs are generated on the host, uploaded once, and stay in HBM; a step is
one decode of the whole batch (kernel launches + synchronize).

Printed: one JSON line with the throughput, the decode kernel's roofline
(algorithmic bytes per launch / kernel time from HIP events on the
decoder's stream), and the reference CPU decoder timed on this host's
cores on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dvb-rx_amd"))

import ldpcgpu  # noqa: E402
import workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
LDS_PEAK_GBS = 256 * 128 * 2.4  # 256 CUs x 128 B/clk x 2.4 GHz, GB/s
# VALU issue: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every 4
# cycles (16 lanes) at 2.4 GHz, G wave-instructions/s
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4


def cpu_flags():
    try:
        return open("/proc/cpuinfo").read()
    except OSError:
        return ""


# CPUs of the GPU box's host that one GPU's job may use (its lease pins this
# share whatever the affinity set shows)
CPU_SHARE = 16


def host_cpu():
    """CPU model, logical and physical core counts of this host, the size
    of the process's affinity set, and one CPU of the affinity set per
    physical core (sysfs topology), in order."""
    txt = cpu_flags()
    model, phys = "unknown", set()
    for block in txt.split("\n\n"):
        d = {}
        for line in block.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                d[k.strip()] = v.strip()
        if "processor" in d:
            model = d.get("model name", model)
            phys.add((d.get("physical id", "0"), d.get("core id", d["processor"])))
    try:
        aff = sorted(os.sched_getaffinity(0))
    except AttributeError:
        aff = list(range(os.cpu_count() or 1))
    seen, per_core = set(), []
    for c in aff:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            key = (open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip())
        except OSError:
            key = (str(c),)
        if key not in seen:
            seen.add(key)
            per_core.append(c)
    info = {"model": model, "logical_cpus": os.cpu_count(), "physical_cores": len(phys) or None,
            "affinity_cpus": len(aff), "affinity_physical_cores": len(per_core), "share": CPU_SHARE}
    return info, per_core


def ref_builds():
    """The reference decoder builds (oracle/_ref) this host can run, widest
    first: (name, ctypes library, description)."""
    flags = cpu_flags()
    out = []
    for name, need, desc in (("v4", "avx512bw", "AVX-512BW, 64 lanes"), ("v3", "avx2", "AVX2, 32 lanes")):
        p = os.path.join(ROOT, "oracle", "_ref", f"libdrxref_{name}.so")
        if need in flags and os.path.exists(p):
            lib = C.CDLL(p)
            vp = C.c_void_p
            lib.drx_ldpc_new.restype = vp
            lib.drx_ldpc_new.argtypes = [vp, C.c_char_p]
            lib.drx_ldpc_free.argtypes = [vp]
            lib.drx_ldpc_decode.argtypes = [vp, vp, vp, C.c_int, C.c_int]
            out.append((name, lib, f"libdvbrx ldpc.c built -O3 -march=x86-64-{name} ({desc})"))
    return out


def ref_batches(keys, lanes, seed=0):
    """Frames grouped into the reference's batches: at most `lanes` frames
    of one code (key), shuffled."""
    by_key = {}
    for f, k in enumerate(keys):
        by_key.setdefault(k, []).append(f)
    out = []
    for k, fl in by_key.items():
        for i in range(0, len(fl), lanes):
            out.append(fl[i:i + lanes])
    np.random.default_rng(seed).shuffle(out)
    return out


def ref_decode_rate(lib, keys, llr_of, kl_of, max_iter, cpus, seconds, max_batches=None):
    """The reference decoder on len(cpus) threads, each pinned to one CPU
    with its own handles, decoding batches of LANES frames of one code (the
    reference's batch semantics: a batch iterates until all its frames
    hold their checks) until the time is up, cycling over the frames.  The
    clock runs around the decode calls only (tests/test_ldpc.c:77-95
    times its loop the same way).  keys[f] = (short_frame, rate) of frame
    f, llr_of(f) its int8 LLRs, kl_of(f) its kldpc."""
    lanes = lib.drx_ldpc_lanes()
    batches = ref_batches(keys, lanes)
    cap = max_batches if max_batches is not None else 64 * len(batches)
    lock = threading.Lock()
    st = dict(next=0, frames=0, bits=0, iters=0, nb=0, failed=0, busy=0.0)
    deadline = [0.0]

    def worker(tid):
        try:
            os.sched_setaffinity(0, {cpus[tid]})
        except (AttributeError, OSError):
            pass
        handles = {}
        out = np.zeros((lanes, 8100), np.uint8)
        busy = 0.0
        while True:
            with lock:
                i = st["next"]
                if time.perf_counter() > deadline[0] or i >= cap:
                    break
                st["next"] += 1
            fl = batches[i % len(batches)]
            key = keys[fl[0]]
            if key not in handles:
                p = ldpcgpu.fec_lookup(*key).as_struct()
                handles[key] = (lib.drx_ldpc_new(C.byref(p), key[1].encode()), p)
            n = len(fl)
            ins = (C.c_void_p * n)(*[llr_of(f).ctypes.data for f in fl])
            outs = (C.c_void_p * n)(*[out[k].ctypes.data for k in range(n)])
            t = time.perf_counter()
            it = lib.drx_ldpc_decode(handles[key][0], ins, outs, n, max_iter)
            busy += time.perf_counter() - t
            with lock:
                st["frames"] += n
                st["bits"] += int(sum(kl_of(f) for f in fl))
                st["iters"] += it if it >= 0 else max_iter
                st["failed"] += it < 0
                st["nb"] += 1
        with lock:
            st["busy"] = max(st["busy"], busy)
        for h, _ in handles.values():
            lib.drx_ldpc_free(h)

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(len(cpus))]
    t0 = time.perf_counter()
    deadline[0] = t0 + seconds
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    return {"threads": len(cpus), "gbit_s": round(st["bits"] / wall / 1e9, 4), "frames_s": round(st["frames"] / wall, 1),
            "frames": st["frames"], "batches": st["nb"], "batches_failed": st["failed"],
            "avg_batch_iterations": round(st["iters"] / max(1, st["nb"]), 2), "wall_s": round(wall, 2), "lanes": lanes}


def cpu_baseline(keys, llr_of, kl_of, max_iter, seconds=8.0, what="the workload"):
    """The reference decoder (libdvbrx ldpc.c built from its sources into
    oracle/_ref) on this host, per build (AVX-512BW and AVX2 where the CPU
    has them): all cores of the box's share (one thread per physical core,
    pinned) for `seconds`, and one core for half as long.  value = the
    best all-core rate."""
    info, per_core = host_cpu()
    builds = ref_builds()
    if not builds:
        return None
    allc = per_core[:min(CPU_SHARE, len(per_core))]
    runs = []
    for name, lib, desc in builds:
        for cpus, sec in ((allc, seconds), (allc[:1], seconds / 2)):
            r = ref_decode_rate(lib, keys, llr_of, kl_of, max_iter, cpus, sec)
            r.update(build=name, desc=desc)
            runs.append(r)
            print(f"# cpu baseline {name} x{len(cpus)}: {r['gbit_s']} Gbit/s", file=sys.stderr, flush=True)
    best = max((r for r in runs if r["threads"] == len(allc)), key=lambda r: r["gbit_s"])
    single = max((r for r in runs if r["threads"] == 1), key=lambda r: r["gbit_s"])
    return {
        "value": best["gbit_s"],
        "unit": "Gbit/s",
        "cores": best["threads"],
        "kind": "reference",
        "sample": (f"{best['frames']} frames of {what} decoded in {best['wall_s']} s ({best['batches']} batches of <= "
                   f"{best['lanes']} frames of one code, shuffled, cycling over the workload), {best['threads']} threads "
                   f"pinned one per physical core, {best['desc']}; batch iterations avg {best['avg_batch_iterations']}"),
        "single_core": single["gbit_s"],
        "single_core_build": single["build"],
        "host": info,
        "runs": runs,
    }


def sk_name(di, occ=1):
    """Kernel name of a split kernel instance from its degree code
    (2 * slots per lane - 2 + 100 * wide), plain arithmetic, built for occ
    workgroups per CU."""
    return f"ldpc_sk_kernel<{(di % 100 + 2) // 2}, {di // 100}, 0, {occ}>"


def copy_calibration(mib):
    """A plain device-to-device copy of known size, for scaling the HBM
    traffic counters (FETCH_SIZE / WRITE_SIZE) of the decode kernel."""
    import torch

    n = mib << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    a.fill_(7)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.add(a, 0, out=b)  # one elementwise kernel: reads n bytes, writes n
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)


def single_code_record(dec, reps=20, cpu=True):
    """The reference's own CPU-runnable case: 64 DVB-S2 normal 1/2 QPSK
    frames at Es/N0 1.5 dB, max 50 iterations.  Timed three ways: the
    device-resident batch API per-frame and as one group (the reference's
    batch semantics), and the drop-in drx_ldpc_decode on host buffers
    (PCIe copies included)."""
    b = workloads.single_code()
    d_llr = dec.malloc(b.llr.nbytes)
    d_out = dec.malloc(b.ref.nbytes)
    d_it = dec.malloc(4 * b.nframes)
    d_ok = dec.malloc(b.nframes)
    dec.upload(d_llr, b.llr)
    res = {"workload": b.name, "frames": b.nframes, "info_bits": b.info_bits}
    for mode, group in (("per_frame", -1), ("group", 0)):
        fr = ldpcgpu.frames_for(b.layout, b.max_iter, group)
        dec.decode_batch(fr, d_llr, d_out, d_it, d_ok)
        dec.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            dec.decode_batch(fr, d_llr, d_out, d_it, d_ok)
        dec.synchronize()
        dt = (time.perf_counter() - t) / reps
        it = np.zeros(b.nframes, np.int32)
        dec.download(it, d_it)
        res[mode] = {"ms_per_call": round(dt * 1e3, 3), "gbit_s": round(b.info_bits / dt / 1e9, 3),
                     "iterations": int(it.max())}
    for p in (d_llr, d_out, d_it, d_ok):
        dec.free(p)
    fec = ldpcgpu.fec_lookup(0, "1_2")
    drx = ldpcgpu.DrxLdpc(fec, "1_2")
    llr = b.llr.reshape(b.nframes, -1)
    r, _ = drx.decode(llr, b.max_iter)
    t = time.perf_counter()
    for _ in range(reps):
        drx.decode(llr, b.max_iter)
    dt = (time.perf_counter() - t) / reps
    drx.close()
    res["dropin_host_buffers"] = {"ms_per_call": round(dt * 1e3, 3), "gbit_s": round(b.info_bits / dt / 1e9, 3),
                                  "result": int(r)}
    if cpu:
        res["cpu_baseline"] = single_code_cpu(b)
    return res


def single_code_cpu(b, reps=10):
    """The reference decoder on the 64-frame normal 1/2 QPSK batch as
    tests/test_ldpc.c:77-95 times it: the frames in batches of LANES on one
    thread (a batch cannot
    be split over threads), clock around the decode loop; best of reps
    after a warm-up, per build."""
    _, per_core = host_cpu()
    try:
        orig = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {per_core[0]})
    except (AttributeError, OSError, IndexError):
        orig = None
    out = []
    try:
        for name, lib, desc in ref_builds():
            lanes = lib.drx_ldpc_lanes()
            p = ldpcgpu.fec_lookup(0, "1_2").as_struct()
            h = lib.drx_ldpc_new(C.byref(p), b"1_2")
            o = np.zeros((b.nframes, 8100), np.uint8)
            best, its = None, []
            for r in range(reps + 1):
                its = []
                t = time.perf_counter()
                for f0 in range(0, b.nframes, lanes):
                    n = min(lanes, b.nframes - f0)
                    ins = (C.c_void_p * n)(*[b.frame_llr(f).ctypes.data for f in range(f0, f0 + n)])
                    outs = (C.c_void_p * n)(*[o[f].ctypes.data for f in range(f0, f0 + n)])
                    its.append(lib.drx_ldpc_decode(h, ins, outs, n, b.max_iter))
                dt = time.perf_counter() - t
                if r and (best is None or dt < best):
                    best = dt
            lib.drx_ldpc_free(h)
            out.append({"build": name, "desc": desc, "threads": 1, "ms_per_call": round(best * 1e3, 3),
                        "gbit_s": round(b.info_bits / best / 1e9, 4), "frames_s": round(b.nframes / best, 1),
                        "batch_iterations": its})
    finally:
        if orig is not None:
            os.sched_setaffinity(0, orig)
    return out


def s2x_high_order_record(dec, margin_db=0.2, reps=5, cpu=False, cpu_seconds=6.0):
    """DVB-S2X high-order batch (16384 normal frames on 64/128/256APSK
    MODCODs, 4096 short on QPSK/8PSK, threshold + margin_db, generated on
    the GPU) decoded by a prepared plan: launch shape chosen by the
    autotuner, replayed as one graph; rooflines over the whole decode.
    The S2X threshold is Shannon + 1 dB; at + 0.2 dB most frames do not
    converge in 50 iterations, with the reference decoder as here."""
    b = workloads.s2x_high_order(dec, margin_db=margin_db)
    plan = ldpcgpu.Plan(dec, b.frames, b.d_llr, b.d_out, b.d_it, b.d_ok)
    try:
        plan.tune(2)
        ms = []
        for _ in range(reps):
            plan.launch()
            ms.append(plan.last_ms())
        out, it, ok = b.results()
        info = plan.info()
    finally:
        plan.close()
    errs = b.frame_errors(out)
    sec = float(np.median(ms)) * 1e-3
    hbm = lds = 0
    params = {}
    for f in range(b.nframes):
        c = int(b.frames["code"][f])
        if c not in params:
            params[c] = ldpcgpu.code_params(c)
        E, n = params[c]["links_total"], params[c]["n"]
        hbm += n + (int(b.kldpc[f]) + 7) // 8 + 8
        lds += 4 * E * int(it[f]) + E * (int(it[f]) + 1)
    cpu_rec = None
    if cpu:
        # the reference decoder on the same LLRs (downloaded), all cores of
        # the share, widest build, for `cpu_seconds`
        llr = np.zeros(b.llr_bytes, np.int8)
        dec.download(llr, b.d_llr)
        keys = []
        for pls, count in b.groups:
            m = ldpcgpu.pls_lookup(pls)
            keys += [(m["short_frame"], m["rate"])] * count
        offs = b.frames["llr_off"].astype(np.int64)
        ns = np.array([ldpcgpu.fec_lookup(*k).nldpc for k in keys], np.int64)
        builds = ref_builds()
        if builds:
            info_h, per_core = host_cpu()
            name, lib, desc = builds[0]
            r = ref_decode_rate(lib, keys, lambda f: llr[offs[f]:offs[f] + ns[f]], lambda f: int(b.kldpc[f]),
                                50, per_core[:min(CPU_SHARE, len(per_core))], cpu_seconds)
            cpu_rec = {"value": r["gbit_s"], "unit": "Gbit/s", "cores": r["threads"], "kind": "reference",
                       "sample": f"{r['frames']} frames of this batch in {r['wall_s']} s, batches of <= {r['lanes']} "
                                 f"frames of one MODCOD, {r['threads']} threads pinned one per physical core, {desc}; "
                                 f"batch iterations avg {r['avg_batch_iterations']}, {r['batches_failed']} of "
                                 f"{r['batches']} batches not converged"}
    b.free()
    return {
        "workload": f"S2X high order: {b.nframes} frames ({sum(c for p, c in b.groups if p < 0xd8)} normal on "
                    f"64/128/256APSK MODCODs, the rest short QPSK/8PSK), threshold + {margin_db} dB",
        "gbit_s": round(b.info_bits / sec / 1e9, 3), "ms": round(sec * 1e3, 3),
        "avg_iterations": round(float(it.mean()), 3), "fer_ldpc": errs / b.nframes,
        "frames_unconverged": int((ok == 0).sum()),
        "roofline_hbm": {"achieved_gbs": round(hbm / sec / 1e9, 2), "frac": round(hbm / sec / 1e9 / HBM_PEAK_GBS, 5)},
        "roofline_lds": {"achieved_gbs": round(lds / sec / 1e9, 1), "frac": round(lds / sec / 1e9 / LDS_PEAK_GBS, 5)},
        "autotuner": info,
        "cpu_baseline": cpu_rec,
        "gpu_vs_cpu": round(b.info_bits / sec / 1e9 / cpu_rec["value"], 2) if cpu_rec and cpu_rec["value"] > 0 else None,
    }


def demap_record(dec, nframes=512, reps=5):
    """GPU max-log demapper (ldpc_gpu_demap) on 512 frames of random
    unit-variance symbols per constellation, resident in HBM: Gsym/s and
    the HBM fraction (8 bytes in per symbol, one LLR byte out per bit)."""
    out = []
    for pls in (0x10, 0x30, 0x60, 0xc8, 0xcc, 0xce):
        ns = ldpcgpu.demap_symbols(pls)
        m = ldpcgpu.pls_lookup(pls)
        nldpc = 16200 if m["short_frame"] else 64800
        rng = np.random.default_rng(pls)
        sym = rng.normal(0, 0.7, (nframes, ns, 2)).astype(np.float32)
        n0 = np.full(nframes, 0.1, np.float32)
        d_sym, d_n0, d_llr = dec.malloc(sym.nbytes), dec.malloc(n0.nbytes), dec.malloc(nframes * nldpc)
        try:
            dec.upload(d_sym, sym)
            dec.upload(d_n0, n0)
            ldpcgpu.demap(dec, pls, nframes, d_sym, d_n0, d_llr)
            dec.synchronize()
            dt = []
            for _ in range(reps):
                t = time.perf_counter()
                ldpcgpu.demap(dec, pls, nframes, d_sym, d_n0, d_llr)
                dec.synchronize()
                dt.append(time.perf_counter() - t)
            sec = float(np.median(dt))
            hbm = nframes * (8 * ns + nldpc)
            out.append({"pls": f"{pls:#04x}", "constellation": m["constellation"], "rate": m["rate"],
                        "frames": nframes, "ms": round(sec * 1e3, 4), "gsym_s": round(nframes * ns / sec / 1e9, 3),
                        "hbm_gb_s": round(hbm / sec / 1e9, 1), "hbm_frac": round(hbm / sec / 1e9 / HBM_PEAK_GBS, 4)})
        finally:
            for p_ in (d_sym, d_n0, d_llr):
                dec.free(p_)
    return out


def bch_record(dec, nframes=8192, reps=5):
    """GPU BCH decoder (ldpc_gpu_bch_decode) on 8192 codewords of DVB-S2
    normal 1/2 (t = 12) resident in HBM: all clean (the usual case after
    LDPC decoding), and every codeword with t errors (the Berlekamp-Massey
    and Chien search path on all of them)."""
    fec = ldpcgpu.fec_lookup(0, "1_2")
    # one BCH codeword of the frame source (noiseless frame: its reference bits)
    _, ref, _ = ldpcgpu.generate([dict(short_frame=0, rate="1_2", pls=-1, esn0_db=1000.0, stream=0)], 5)
    cw = np.asarray(ref, np.uint8).reshape(-1)
    nb = len(cw)
    rng = np.random.default_rng(5)
    out = []
    for name, w in (("clean", 0), (f"{fec.t} errors each", fec.t)):
        buf = np.tile(cw, (nframes, 1))
        for f in range(nframes if w else 0):
            for bit in rng.choice(fec.kldpc, w, replace=False):
                buf[f, bit // 8] ^= 0x80 >> (bit % 8)
        d_cw, d_res = dec.malloc(buf.nbytes), dec.malloc(4 * nframes)
        try:
            dt = []
            for _ in range(reps + 1):
                dec.upload(d_cw, buf)
                t = time.perf_counter()
                ldpcgpu.bch_decode(dec, 0, "1_2", nframes, d_cw, nb, d_res)
                dec.synchronize()
                dt.append(time.perf_counter() - t)
            res = np.zeros(nframes, np.int32)
            dec.download(res, d_res)
            sec = float(np.median(dt[1:]))
            out.append({"codewords": name, "frames": nframes, "ms": round(sec * 1e3, 3),
                        "mframes_s": round(nframes / sec / 1e6, 3), "info_gbit_s": round(nframes * fec.kbch / sec / 1e9, 2),
                        "corrected": int((res == w).sum())})
        finally:
            dec.free(d_cw)
            dec.free(d_res)
    return out


T2_PLPS = [
    # (name, cod, mod, fec_type, Es/N0 dB: lowest 0.2 dB step at which 16
    # distinct blocks all decode, plus 0.3 dB, blocks), rotated, with a
    # type-0 time interleaver of 8 TI-blocks (one call is one interleaving
    # frame; 8 keeps each TI-block within the cell interleaver's shifts)
    ("normal 256-QAM 3/5 rotated", 1, 3, 1, 18.7, 4096),
    ("short 64-QAM 2/3 rotated", 2, 2, 0, 14.1, 8192),
]
T2_TI_BLOCKS = 8


def t2_cpu_job(plp, cells, csi, max_iter, seconds=8.0, max_threads=CPU_SHARE):
    """The reference's DVB-T2 FEC job (dvbt2.c decode_job, built from its
    sources into oracle/_ref) on this host: one thread per core, each
    decoding calls of 64 blocks of the same cells until the deadline."""
    flags = cpu_flags()
    lib, what = None, None
    for name, need in (("v4", "avx512bw"), ("v3", "avx2")):
        p = os.path.join(ROOT, "oracle", "_ref", f"libdrxt2ref_{name}.so")
        if need in flags and os.path.exists(p):
            lib, what = C.CDLL(p), f"libdvbrx dvbt2.c decode_job built -O3 -march=x86-64-{name} (clang)"
            break
    if lib is None:
        return None
    lib.t2ref_decode.argtypes = [C.c_int] * 7 + [C.c_void_p] * 2 + [C.c_int] + [C.c_void_p] * 5
    _, per_core = host_cpu()
    cpus = per_core[:min(max_threads, len(per_core))]
    nthreads = len(cpus)
    nb = min(64, cells.shape[0])
    c = np.ascontiguousarray(cells[:nb])
    h = np.ascontiguousarray(csi[:nb])
    N, K, kbch, nc = ldpcgpu.t2_geometry(plp)
    count = [0] * nthreads
    deadline = time.perf_counter() + seconds

    def worker(tid):
        try:
            os.sched_setaffinity(0, {cpus[tid]})
        except (AttributeError, OSError):
            pass
        bb = np.zeros(nb * kbch // 8, np.uint8)
        ok = np.zeros(nb, np.uint8)
        e = np.zeros(nb, np.float32)
        g = np.zeros(nb, np.float32)
        while time.perf_counter() < deadline:
            rc = lib.t2ref_decode(plp.cod, plp.mod, plp.fec_type, plp.rotation, plp.til_length, plp.til_type, nb,
                                  c.ctypes.data, h.ctypes.data, max_iter, bb.ctypes.data, ok.ctypes.data,
                                  e.ctypes.data, g.ctypes.data, None)
            if rc:
                break
            count[tid] += nb

    t0 = time.perf_counter()
    th = [threading.Thread(target=worker, args=(i,)) for i in range(nthreads)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    sec = time.perf_counter() - t0
    blocks = sum(count)
    return {"value": round(blocks * kbch / sec / 1e9, 4), "unit": "Gbit/s", "cores": nthreads, "kind": "reference",
            "blocks_s": round(blocks / sec, 1),
            "sample": f"{what}, {nthreads} threads pinned one per physical core, calls of {nb} blocks, {blocks} blocks in {sec:.1f} s"}


def t2_record(dec, reps=5, cpu=True):
    """DVB-T2 FEC blocks of two PLPs (the T2 PLP codes of the streaming
    mix: normal 3/5, short 2/3), resident in HBM: the front end alone
    (ldpc_gpu_t2_demap: cell and bit deinterleaving, derotation, MER, soft
    values) and the whole job (ldpc_gpu_t2_decode: front end, LDPC up to 30
    iterations, BCH), every block checked against its BBFRAME.  Blocks
    are 16 distinct encoded frames repeated; the reference's job on this
    host's cores beside it."""
    out = []
    for name, cod, mod, ft, snr, nblocks in T2_PLPS:
        plp = ldpcgpu.T2Plp(cod, mod, ft, 1, T2_TI_BLOCKS, 0)
        N, K, kbch, nc = ldpcgpu.t2_geometry(plp)
        cells, csi, info, _ = workloads.t2_blocks(plp, nblocks, snr)
        kb = K // 8
        sizes = dict(cells=cells.nbytes, csi=csi.nbytes, llr=nblocks * N, cw=nblocks * kb, it=4 * nblocks,
                     syn=nblocks, bch=4 * nblocks, err=4 * nblocks, sig=4 * nblocks)
        d = {k: dec.malloc(v) for k, v in sizes.items()}
        try:
            dec.upload(d["cells"], cells)
            dec.upload(d["csi"], csi)

            def front():
                ldpcgpu.t2_demap(dec, plp, nblocks, d["cells"], d["csi"], d["llr"], d["err"], d["sig"])
                dec.synchronize()

            def job(lanes):
                ldpcgpu.t2_decode(dec, plp, nblocks, d["cells"], d["csi"], 30, lanes, d["llr"], d["cw"], d["it"],
                                  d["syn"], d["bch"], d["err"], d["sig"])
                dec.synchronize()

            def timed(fn, *a):
                fn(*a)
                dt = []
                for _ in range(reps):
                    t = time.perf_counter()
                    fn(*a)
                    dt.append(time.perf_counter() - t)
                return float(np.median(dt))

            t_front = timed(front)
            t_job = timed(job, 0)
            res = np.zeros(nblocks, np.int32)
            cw = np.zeros((nblocks, kb), np.uint8)
            it = np.zeros(nblocks, np.int32)
            dec.download(res, d["bch"])
            dec.download(cw, d["cw"])
            dec.download(it, d["it"])
            good = int(((res >= 0) & (cw[:, :kbch // 8] == info).all(1)).sum())
            t_job32 = timed(job, 32)
            # the same job on host buffers (decode_job's interface): PCIe and
            # per-call allocation included
            t0 = time.perf_counter()
            hbb, hok, _, _ = ldpcgpu.t2_decode_host(dec, plp, cells, csi, 30, 0)
            t_host = time.perf_counter() - t0
            host_ok = bool(np.array_equal(hbb, info) and hok.all())
            alg = nblocks * (12 * nc + N)
            rec = {"plp": name, "es_n0_db": snr, "blocks": nblocks, "cells_per_block": nc,
                   "front_end_ms": round(t_front * 1e3, 4), "front_end_gcells_s": round(nblocks * nc / t_front / 1e9, 2),
                   "front_end_hbm_frac": round(alg / t_front / 1e9 / HBM_PEAK_GBS, 4),
                   "job_ms": round(t_job * 1e3, 3), "job_info_gbit_s": round(nblocks * kbch / t_job / 1e9, 3),
                   "job_ms_lanes32": round(t_job32 * 1e3, 3),
                   "job_info_gbit_s_lanes32": round(nblocks * kbch / t_job32 / 1e9, 3),
                   "job_host_buffers_ms": round(t_host * 1e3, 2),
                   "job_host_buffers_info_gbit_s": round(nblocks * kbch / t_host / 1e9, 3),
                   "job_host_buffers_all_decoded": host_ok,
                   "avg_iterations": round(float(it.mean()), 2), "blocks_decoded": good}
            if cpu:
                rec["cpu_baseline"] = t2_cpu_job(plp, cells, csi, 30)
            out.append(rec)
        finally:
            for v in d.values():
                dec.free(v)
    return out


# codes of the per-code microbenchmark: (frame size, rate); the unit
# channel's Es/N0 sits below each code's waterfall, so frames iterate
MICRO_CODES = [(0, "1_4"), (0, "1_3"), (0, "1_2"), (0, "3_5"), (0, "2_3"), (0, "3_4"), (0, "5_6"), (0, "9_10"),
               (1, "1_2"), (1, "3_4"), (1, "8_9"), (0, "154_180"), (0, "2_3_t2")]
MICRO_ESN0 = {"1_4": -3.5, "1_3": -2.6, "1_2": -0.5, "3_5": 0.7, "2_3": 1.5, "3_4": 2.5, "5_6": 3.7, "9_10": 4.9,
              "8_9": 4.7, "154_180": 5.0, "2_3_t2": 1.5, "90_180": -0.5, "135_180": 2.5}


def microbench(dec, nframes=2048, iters=10, reps=3, codes=None):
    """Decoder cost per code: nframes frames of one code on the unit
    channel (+-1 per bit, AWGN) below the code's waterfall, max `iters`
    iterations, per-frame early stop.  Reports microseconds per
    frame-iteration over the whole GPU and picoseconds per edge (edge =
    one of the code's links_total per iteration), plus the level steps and
    the hand-over rounds of the kernel's schedule per iteration."""
    out = []
    for short, rate in (codes or MICRO_CODES):
        fec = ldpcgpu.fec_lookup(short, rate)
        code = ldpcgpu.code_lookup(short, rate)
        prm = ldpcgpu.code_params(code)
        specs = [dict(short_frame=short, rate=rate, esn0_db=MICRO_ESN0[rate], stream=f) for f in range(nframes)]
        llr, _, layout = ldpcgpu.generate(specs, 99, with_ref=False)
        fr = ldpcgpu.frames_for(layout, iters)
        d_llr = dec.malloc(llr.nbytes)
        d_out = dec.malloc(nframes * (fec.kldpc // 8 + 1))
        d_it = dec.malloc(4 * nframes)
        d_ok = dec.malloc(nframes)
        dec.upload(d_llr, llr)
        dec.decode_batch(fr, d_llr, d_out, d_it, d_ok)
        dec.synchronize()
        ms = []
        for _ in range(reps):
            dec.decode_batch(fr, d_llr, d_out, d_it, d_ok)
            dec.synchronize()
            ms.append(dec.last_kernel_ms())
        it = np.zeros(nframes, np.int32)
        dec.download(it, d_it)
        for ptr in (d_llr, d_out, d_it, d_ok):
            dec.free(ptr)
        kms = float(np.median(ms))
        fi = int(it.sum())
        out.append({"code": f"{'short' if short else 'normal'} {rate}", "frames": nframes,
                    "avg_iterations": round(fi / nframes, 3), "kernel_ms": round(kms, 3),
                    "us_per_frame_iteration": round(kms * 1e3 / max(fi, 1), 4),
                    "ps_per_edge": round(kms * 1e9 / max(fi * prm["links_total"], 1), 3),
                    "layers": prm["q"], "level_steps": prm["level_steps"],
                    "gbit_s": round(nframes * fec.kldpc / (kms * 1e-3) / 1e9, 3)})
        print(f"# microbench {out[-1]}", file=sys.stderr, flush=True)
    return out


def learned_record(dec, nf=8192, db=0.2, reps=3):
    """Learned vs plain decoding on the two trained MODCODs (DVB-S2 normal
    3/5 8PSK and 2/3 16APSK) at Es/N0 = QEF + db: frames from pool
    codewords through the GPU channel, decoded by the plain decoder (up to
    50 iterations) and by the learned table (up to 30 learned iterations,
    plain fallback up to 50); frame errors after the BCH code's t."""
    from nms_train import DevBuf, codeword_pool

    out = {}
    for name in ("s2_n3_5_8psk", "s2_n2_3_16apsk"):
        table = json.load(open(os.path.join(ROOT, "dvb-rx_amd", "nms", name + ".json")))
        pls, short, rate = table["pls"], table["short_frame"], table["rate"]
        code = ldpcgpu.code_lookup(short, rate)
        fec = ldpcgpu.fec_lookup(short, rate)
        N, kl = fec.nldpc, fec.kldpc
        kb = (kl + 7) // 8
        pool = codeword_pool(short, rate, 512, 5)
        ref = np.packbits(pool[:, :kl], axis=1)
        rows = np.random.default_rng(6).integers(0, 512, nf).astype(np.int32)
        dec.set_nms(code, np.array(table["S"], np.uint8), np.array(table["B"], np.int16))
        dv = DevBuf(dec)
        rec = {"esn0_db": round(ldpcgpu.qef_db(pls) + db, 3), "frames": nf}
        try:
            d_pool = dv.put(pool)
            d_rows = dv.put(rows)
            d_e = dv.put(np.full(nf, ldpcgpu.qef_db(pls) + db, np.float32))
            d_llr = dv.empty(nf * N)
            ldpcgpu.channel(pls, N, d_pool, d_rows, nf, d_e, 77, d_llr, stream=dec.stream)
            dec.synchronize()
            frames = np.zeros(nf, ldpcgpu.FRAME_DTYPE)
            frames["code"] = code
            frames["nbits"] = kl
            frames["group"] = -1
            frames["llr_off"] = np.arange(nf) * N
            frames["out_off"] = np.arange(nf) * kb
            d_out, d_it, d_ok, d_fb = dv.empty(nf * kb), dv.empty(4 * nf), dv.empty(nf), dv.empty(nf)
            for mode in ("plain", "learned"):
                fr = frames.copy()
                fr["max_iter"] = 50 if mode == "plain" else 30
                dt = []
                for _ in range(reps):
                    dec.synchronize()
                    t = time.perf_counter()
                    if mode == "plain":
                        dec.decode_batch(fr, d_llr, d_out, d_it, d_ok)
                    else:
                        dec.decode_batch_nms(fr, d_llr, d_out, d_it, d_ok, 50, d_fb)
                    dec.synchronize()
                    dt.append(time.perf_counter() - t)
                bits = dv.get(d_out, (nf, kb), np.uint8)
                errs = np.unpackbits(bits ^ ref[rows], axis=1)[:, :kl].sum(axis=1)
                it = dv.get(d_it, nf, np.int32)
                sec = float(np.median(dt))
                rec[mode] = {"gbit_s": round(nf * kl / sec / 1e9, 3), "ms": round(sec * 1e3, 3),
                             "avg_iterations": round(float(it.mean()), 3),
                             "fer_after_bch": float((errs > fec.t).mean()), "fer_ldpc": float((errs > 0).mean())}
                if mode == "learned":
                    rec[mode]["fallback_share"] = float(dv.get(d_fb, nf, np.uint8).mean())
        finally:
            dv.free()
            dec.set_nms(code, None, None)
        out[name] = rec
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=8192)
    ap.add_argument("--cpu-seconds", type=float, default=8.0,
                    help="all-core time of each reference build on the mixed workload (one core: half)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the side measurements (64-frame rate 1/2 batch, learned vs plain)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (HBM counter calibration) and exit")
    ap.add_argument("--serial", action="store_true",
                    help="run the launches of one decode one after the other (per-kernel durations for profiling)")
    ap.add_argument("--codes", default=None, metavar="S:RATE,...",
                    help="microbench only these codes (frame size 0/1 : rate name), e.g. 0:3_4,1:1_2")
    ap.add_argument("--only", choices=["demapper", "bch", "learned", "t2", "microbench"], default=None,
                    help="only run that side record (the demapper's or the BCH decoder's throughput, learned vs "
                         "plain decoding, DVB-T2 FEC blocks) and print it")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: frames are independent, so N GPUs would each decode their "
              "own share of the frame table with no exchange; run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.only:
        dec = ldpcgpu.Decoder(0)
        if args.serial:
            dec.set_concurrent(False)
        if args.only == "microbench" and args.codes:
            codes = [(int(c.split(":")[0]), c.split(":")[1]) for c in args.codes.split(",")]
            rec = microbench(dec, codes=codes)
        else:
            rec = {"demapper": demap_record, "bch": bch_record, "learned": learned_record, "t2": t2_record,
                   "microbench": microbench}[args.only](dec)
        dec.close()
        print(json.dumps({args.only: rec}), flush=True)
        return

    t0 = time.time()
    batch = workloads.mixed_s2(args.frames)
    print(f"# generated {batch.nframes} frames ({batch.llr.nbytes / 1e6:.0f} MB LLRs) in {time.time() - t0:.1f} s",
          file=sys.stderr, flush=True)
    dec = ldpcgpu.Decoder(0)
    if args.serial:
        dec.set_concurrent(False)
    d_llr = dec.malloc(batch.llr.nbytes)
    d_out = dec.malloc(batch.ref.nbytes)
    d_it = dec.malloc(4 * batch.nframes)
    d_ok = dec.malloc(batch.nframes)
    dec.upload(d_llr, batch.llr)

    for _ in range(args.warmup):
        dec.decode_batch(batch.frames, d_llr, d_out, d_it, d_ok)
    dec.synchronize()
    kernel_ms = []
    launches = {}
    t = time.perf_counter()
    for _ in range(args.steps):
        dec.decode_batch(batch.frames, d_llr, d_out, d_it, d_ok)
        dec.synchronize()
        kernel_ms.append(dec.last_kernel_ms())
        for di, nf, ms in dec.last_launches():
            launches.setdefault((di, nf), []).append(ms)
    dt = time.perf_counter() - t
    # per-launch durations with the launches one after another, each kernel
    # alone on the GPU (the rooflines use these), after the timed steps
    serial_ms, serial_occ = {}, {}
    dec.set_concurrent(False)
    for r in range(2):
        dec.decode_batch(batch.frames, d_llr, d_out, d_it, d_ok)
        dec.synchronize()
        if r:
            serial_ms = {(di, nf): ms for di, nf, ms in dec.last_launches()}
            serial_occ = {(di, nf): sh[0] for (di, nf, _), sh in zip(dec.last_launches(), dec.last_launch_shapes())}
            serial_decode_ms = dec.last_kernel_ms()
    dec.set_concurrent(not args.serial)
    out = np.zeros_like(batch.ref)
    it = np.zeros(batch.nframes, np.int32)
    ok = np.zeros(batch.nframes, np.uint8)
    dec.download(out, d_out)
    dec.download(it, d_it)
    dec.download(ok, d_ok)
    frame_err = 0
    bad_f = np.zeros(batch.nframes, bool)
    by_modcod = {}  # pls -> [frames, frames in error]
    for f in range(batch.nframes):
        bad = int(not np.array_equal(batch.frame_ref(f), out[batch.layout[f][4]:batch.layout[f][4] + (batch.kldpc[f] + 7) // 8]))
        bad_f[f] = bool(bad)
        frame_err += bad
        e = by_modcod.setdefault(int(batch.specs[f]["pls"]), [0, 0])
        e[0] += 1
        e[1] += bad
    modcod_err = {}
    for pls, (nf_, ne) in sorted(by_modcod.items()):
        if ne:
            m = ldpcgpu.pls_lookup(pls)
            modcod_err[f"{pls:#04x} {'short' if m['short_frame'] else 'normal'} {m['rate']} {m['constellation']}"] = \
                {"frames": nf_, "in_error": ne, "esn0_db": round(ldpcgpu.qef_db(pls) + 0.3, 2)}
    # MODCODs that fail outright at this margin (more than 10 % of their
    # frames), and the frame error rate of all the others
    failing = [v for v in modcod_err.values() if 10 * v["in_error"] > v["frames"]]
    fer_rest = (frame_err - sum(v["in_error"] for v in failing)) / max(1, batch.nframes - sum(v["frames"] for v in failing))
    ms_step = dt / args.steps * 1e3
    sec_step = dt / args.steps
    # The metric counts decoded information: only frames whose kldpc bits
    # equal the sent ones count (the time of the others still counts).
    # Beside it: all frames' bits, and the bits of the MODCODs with no frame
    # in error in this batch (FER <= 1e-4 cannot be shown on their 60-200
    # frames each; no error is the batch's best estimate of it).
    value = float(batch.kldpc[~bad_f].sum()) / sec_step / 1e9
    value_all = batch.info_bits / sec_step / 1e9
    clean = np.array([by_modcod[int(batch.specs[f]["pls"])][1] == 0 for f in range(batch.nframes)])
    value_clean_modcods = float(batch.kldpc[clean].sum()) / sec_step / 1e9

    # Roofline of the dominant decode kernel.  One decode call launches the
    # split kernel once per (instance, frame size) with frames; an
    # instance is ldpc_sk_kernel<KS, WIDE, NMS> (KS slots per lane: the
    # code's largest layer degree + 2, halved, rounded up; WIDE: the code's
    # layers differ by more than one slot).  The dominant instance is the
    # one with the most launch time; its figures are per launch, averaged
    # over its launches, as a profiler's per-kernel statistics give them.
    # Algorithmic bytes per frame:
    #   HBM: N LLR bytes in + kldpc/8 bytes out + 8 (count, flag);
    #   LDS: per iteration read + write of posterior and message per edge
    #        (4E), one read per edge for each syndrome pass (it + 1).
    kms = float(np.mean(kernel_ms))
    params = {}
    cls_of = np.zeros(batch.nframes, np.int64)
    hbm_f = np.zeros(batch.nframes, np.int64)
    lds_f = np.zeros(batch.nframes, np.int64)
    for f in range(batch.nframes):
        c = int(batch.codes[f])
        if c not in params:
            params[c] = ldpcgpu.code_params(c)
        E = params[c]["links_total"]
        cls_of[f] = params[c]["degree_class"]
        hbm_f[f] = int(batch.n[f]) + (int(batch.kldpc[f]) + 7) // 8 + 8
        lds_f[f] = 4 * E * int(it[f]) + E * (int(it[f]) + 1)
    launch_ms = {k: float(np.mean(v)) for k, v in launches.items()}
    inst = {}
    launch_order = []  # the decode's launches in dispatch order
    for (di, nf), ms in serial_ms.items():
        size = None
        for short in (0, 1):
            m = (cls_of == di) & ((batch.n == 16200) == bool(short))
            if int(m.sum()) == nf:
                size = ("normal", "short")[short]
                sel = m
        if size is None:
            continue
        name = sk_name(di, serial_occ[(di, nf)])
        launch_order.append(f"{name} {size} x{nf}")
        e = inst.setdefault(di, [])
        e.append(dict(kernel=name, size=size, ms=ms, frames=nf, hbm=int(hbm_f[sel].sum()), lds=int(lds_f[sel].sum())))
    dom_di = max(inst, key=lambda d: sum(x["ms"] for x in inst[d]))
    dl = inst[dom_di]
    kname = " / ".join(dict.fromkeys(x["kernel"] for x in dl))
    dom_ms = float(np.mean([x["ms"] for x in dl]))
    hbm_bytes = float(np.mean([x["hbm"] for x in dl]))
    lds_bytes = float(np.mean([x["lds"] for x in dl]))
    edges = batch.edges()
    iters_total = int(it.sum())
    hbm_ach = hbm_bytes / (dom_ms * 1e-3) / 1e9
    lds_ach = lds_bytes / (dom_ms * 1e-3) / 1e9
    # HBM bytes and VALU instructions of the same launches from the
    # committed rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE per
    # dispatch, scaled by what the counters read for a plain copy of known
    # size; SQ_INSTS_VALU per dispatch), for the workload's default size
    traffic, valu = None, None
    tpath = os.path.join(ROOT, "profiles", "counters_mixed_s2_x8192.json")
    if os.path.exists(tpath) and args.frames == 8192:
        try:
            tl = json.load(open(tpath))["launches"]
            traffic = round(float(np.mean([tl[f"{x['kernel']} {x['size']}"]["hbm_bytes"] for x in dl])))
            valu = float(np.mean([tl[f"{x['kernel']} {x['size']}"]["valu_insts"] for x in dl]))
        except (OSError, ValueError, KeyError):
            traffic, valu = None, None
    hbm_rec = {"achieved": round(hbm_ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(hbm_ach / HBM_PEAK_GBS, 5),
               "algorithmic_bytes": round(hbm_bytes), "traffic": traffic}
    roofline = {
        "bound": "hbm", "achieved": hbm_rec["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_rec["frac"],
        "traffic": traffic,
    }
    if valu:
        # the binding resource: VALU issue (instructions from the committed
        # counter profile of these launches, per second of the launch)
        v_ach = valu / (dom_ms * 1e-3) / 1e9
        roofline = {"bound": "valu", "achieved": round(v_ach, 2), "peak": VALU_PEAK_GINST, "unit": "Ginst/s",
                    "frac": round(v_ach / VALU_PEAK_GINST, 4), "traffic": traffic, "hbm": hbm_rec,
                    "valu_insts_per_launch": round(valu)}
    roofline.update({
        "kernel": kname,
        "kernel_ms": round(dom_ms, 4),
        "algorithmic_bytes": round(hbm_bytes),
        "launches": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in x.items() if k != "lds"} for x in dl],
        "serial_launches_ms": {f"{sk_name(k[0])} x{k[1]}": round(v, 4) for k, v in sorted(serial_ms.items())},
        "serial_decode_ms": round(serial_decode_ms, 4),
        "launch_order": launch_order,
        "concurrent_launches_ms": {f"{sk_name(k[0])} x{k[1]}": round(v, 4) for k, v in sorted(launch_ms.items())},
        "concurrent_decode_ms": round(kms, 4),
        "note": "per launch of the kernel alone on the GPU (the decode's launches run one after another in a pass "
                "after the timed steps), averaged over its launches; the timed steps run them concurrently",
    })

    # the same batch as a prepared plan: launch shape chosen by the
    # autotuner, replayed as one graph (after the timed steps; the plan
    # writes the same results into the same buffers)
    plan_rec = None
    if not args.headline_only:
        plan = ldpcgpu.Plan(dec, batch.frames, d_llr, d_out, d_it, d_ok)
        try:
            plan.tune(2)
            pms = []
            for _ in range(3):
                plan.launch()
                pms.append(plan.last_ms())
            pinfo = plan.info()
        finally:
            plan.close()
        pmed = float(np.median(pms))
        plan_rec = {"gbit_s": round(batch.info_bits / (pmed * 1e-3) / 1e9, 3), "ms": round(pmed, 3),
                    "autotuner": pinfo}
    c1 = None if args.headline_only else single_code_record(dec, cpu=not args.no_cpu)
    c4 = None if args.headline_only else learned_record(dec)
    c3 = None if args.headline_only else [s2x_high_order_record(dec, 0.2, cpu=not args.no_cpu), s2x_high_order_record(dec, 1.0, cpu=not args.no_cpu)]
    dm = None if args.headline_only else demap_record(dec)
    bch = None if args.headline_only else bch_record(dec)
    t2 = None if args.headline_only else t2_record(dec, cpu=not args.no_cpu)
    c5 = None
    if not args.headline_only:
        import receiver

        # the 16 receivers on the GPU, two decoder contexts: frames sent as
        # they arrive (dynamic batching), as libdvbrx's receiver batches them
        # (64 frames per receiver, sent when full or after 100 ms), and at
        # 4x and 8x the load
        dec2 = ldpcgpu.Decoder(0)
        try:
            c5 = []
            for dur, load, bmax in ((0.5, 1.0, 1), (0.5, 1.0, 64), (0.25, 4.0, 1), (0.25, 8.0, 1)):
                rec, _, _ = receiver.run([dec, dec2], dur, load, batch_max=bmax)
                c5.append(rec)
                print(f"# receivers load {load} batch {bmax}: p50 {rec['latency_ms']['p50']} ms p99 "
                      f"{rec['latency_ms']['p99']} ms", file=sys.stderr, flush=True)
            c5.append({"decoder": "learned", "applicable": False,
                       "reason": "no learned table for the receivers' MODCODs: the shipped tables are for DVB-S2 "
                                 "3/5 8PSK and 2/3 16APSK; tables trained for 8PSK 3/4 and S2X QPSK 11/20 cut "
                                 "iterations by 0-6 % and lost throughput, and the reference's T2 FEC has none"})
        finally:
            dec2.close()
    cpu = None
    if not args.no_cpu:
        mc = {}
        for sp in batch.specs:
            if sp["pls"] not in mc:
                m = ldpcgpu.pls_lookup(sp["pls"])
                mc[sp["pls"]] = (m["short_frame"], m["rate"])
        keys = [mc[sp["pls"]] for sp in batch.specs]
        cpu = cpu_baseline(keys, batch.frame_llr, lambda f: int(batch.kldpc[f]), batch.max_iter, args.cpu_seconds,
                           "the mixed DVB-S2 workload")
    dec.close()

    res = {
        "metric": "decoded info Gbit/s on 1 MI355X at FER\u22641e-4 (learned vs plain); p99 frame latency",
        "value": round(value, 3),
        "unit": "Gbit/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic: own BCH+LDPC encoder, AWGN, int8 max-log LLRs, resident in HBM",
        "config": {
            "workload": f"mixed DVB-S2 batch, {batch.nframes} frames (75% normal / 25% short), all S2 rates, "
                        "QPSK/8PSK/16APSK/32APSK, Es/N0 = QEF + 0.3 dB, max 50 iterations, per-frame early stop",
            "decoder": "plain layered offset min-sum, bit-exact with the reference",
            "frames": batch.nframes,
            "info_bits_per_step": batch.info_bits,
            "avg_iterations": round(float(it.mean()), 3),
            "fer": frame_err / batch.nframes,
            "gbit_s_all_frames": round(value_all, 3),
            "gbit_s_modcods_without_frame_errors": round(value_clean_modcods, 3),
            "value_counts": "information bits of the frames decoded correctly, per second of the whole decode",
            "fer_modcods_in_error": modcod_err,
            "fer_without_failing_modcods": round(fer_rest, 6),
            "frames_unconverged": int((ok == 0).sum()),
            "autotuned_plan": plan_rec,
        },
        "roofline": roofline,
        "roofline_lds": {
            "achieved": round(lds_ach, 1),
            "peak": LDS_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(lds_ach / LDS_PEAK_GBS, 5),
            "algorithmic_bytes": lds_bytes,
            "edges_per_iteration": edges,
            "iterations_total": iters_total,
        },
        "cpu_baseline": cpu,
        "single_code_64": c1,
        "learned_vs_plain": c4,
        "s2x_high_order_graph": c3,
        "streaming": c5,
        "demapper": dm,
        "bch": bch,
        "t2": t2,
    }
    if cpu:
        res["gpu_vs_cpu"] = round(value / cpu["value"], 2) if cpu["value"] > 0 else None
        res["gpu_vs_cpu_single_core"] = round(value / cpu["single_core"], 1) if cpu["single_core"] > 0 else None
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
