#!/usr/bin/env python3
"""Per-launch counters of the decoder kernel from rocprofv3 runs.

Inputs, each the output directory of one rocprofv3 run (csv format) of
the same command (bench.py on the mixed workload, launches serial):

  --trace DIR        --kernel-trace --stats
  --pmc DIR ...      --pmc runs (any SQ_* counters; one run per pass)
  --fetch DIR        --pmc FETCH_SIZE
  --write DIR        --pmc WRITE_SIZE
  --copy-fetch DIR   bench.py --copy-calibration run with --pmc FETCH_SIZE
  --copy-write DIR   bench.py --copy-calibration run with --pmc WRITE_SIZE

Dispatches of ldpc_sk_kernel<KS, WIDE, NMS, OCC> are keyed by instance and
frame size and averaged over the dispatches of a key.  The traces do not
carry the launch's dynamic LDS, so the frame size comes from the order of
the decode's launches (--order: the JSON line bench.py printed, whose
roofline.launch_order lists them in dispatch order; every decode of the
run has that order): the i-th decoder dispatch of a run is the
(i mod count)-th launch of the list.  FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are
scaled by the factor that maps the plain copy kernel's reading to its
known size.  Durations: bench.py runs its decodes' launches concurrently,
then (the last --serial-decodes decodes) one after another, each kernel
alone on the GPU, as the counter runs do; "ms" is the mean over those
serial dispatches, "ms_concurrent" over the others.  Derived per launch
(serial durations, clock 2.4 GHz):

  valu_issue_frac  SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x cycles): the
                   share of the SIMDs' VALU issue slots used (one wave64
                   VALU instruction per SIMD every 4 cycles)
  waves_per_simd   SQ_WAVE_CYCLES x 4 / (1024 x cycles): waves resident
  valu_per_wave    SQ_INSTS_VALU / SQ_WAVES, likewise salu, lds, branch
  lds_conflict     SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE: share of the
                   LDS's indexed cycles stalled by bank conflicts

Writes the JSON summary (--out) that bench.py reads.
"""
import argparse
import csv
import glob
import json
import os
import re
import sys

CLOCK_GHZ = 2.4
SIMDS = 256 * 4
KRE = re.compile(r"ldpc_sk_kernel<(\d+), *(\d+), *(\d+)(?:, *(\d+))?>")
ORDER = []  # the decode's launches in dispatch order: "name size"


def rows(d, name):
    files = glob.glob(os.path.join(d, "**", f"*{name}.csv"), recursive=True)
    if not files:
        sys.exit(f"no {name}.csv under {d}")
    out = []
    for f in files:
        with open(f) as fh:
            out.extend(csv.DictReader(fh))
    return out


def sk_name(name):
    m = KRE.search(name)
    if not m:
        return None
    return f"ldpc_sk_kernel<{m.group(1)}, {m.group(2)}, {m.group(3)}, {m.group(4) or 1}>"


def label(dispatches):
    """{dispatch id: key} for (dispatch id, kernel name) pairs of one run:
    decoder dispatches in id order, matched to the launch order."""
    sk = sorted((did, sk_name(n)) for did, n in dispatches if sk_name(n))
    out = {}
    for i, (did, nm) in enumerate(sk):
        want = ORDER[i % len(ORDER)]
        if not want.startswith(nm + " "):
            sys.exit(f"dispatch {did} ({nm}) does not match launch {i % len(ORDER)} of the order ({want})")
        out[did] = want
    return out


def counters(d):
    """{dispatch: (key, kernel name, {counter: value summed over its dimensions})}"""
    acc = {}
    for r in rows(d, "counter_collection"):
        did = int(r["Dispatch_Id"])
        e = acc.setdefault(did, (None, r["Kernel_Name"], {}))
        c = r["Counter_Name"]
        e[2][c] = e[2].get(c, 0.0) + float(r["Counter_Value"])
    keys = label([(did, e[1]) for did, e in acc.items()])
    return {did: (keys.get(did), e[1], e[2]) for did, e in acc.items()}


def copy_factor(d, counter, copy_bytes):
    vals = [c[counter] for k, name, c in counters(d).values()
            if counter in c and ("copy" in name.lower() or "elementwise" in name.lower())]
    vals = [v for v in vals if v * 1024 > copy_bytes / 4]  # the big copies, not small fills
    if not vals:
        sys.exit(f"no copy kernel found in {d}")
    vals.sort()
    measured = vals[len(vals) // 2] * 1024.0
    return copy_bytes / measured, measured


def mean(v):
    return sum(v) / len(v) if v else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--order", required=True, help="bench.py's JSON output line (file) of the profiled command")
    ap.add_argument("--pmc", nargs="*", default=[])
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--serial-decodes", type=int, default=2,
                    help="the run's last decodes, whose launches ran one after another")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    with open(a.order) as fh:
        line = [x for x in fh.read().splitlines() if x.startswith("{")][-1]
    ORDER.extend(re.sub(r" x\d+$", "", x) for x in json.loads(line)["roofline"]["launch_order"])

    launches = {}
    trace = rows(a.trace, "kernel_trace")
    keys = label([(int(r["Dispatch_Id"]), r["Kernel_Name"]) for r in trace])
    for r in trace:
        k = keys.get(int(r["Dispatch_Id"]))
        if k is None:
            continue
        e = launches.setdefault(k, {"ms": [], "grid_wg": int(r["Grid_Size_X"]) // int(r["Workgroup_Size_X"]),
                                    "lds": int(r["LDS_Block_Size"]), "vgpr": int(r["VGPR_Count"]),
                                    "sgpr": int(r["SGPR_Count"])})
        e.setdefault("all", []).append((int(r["Dispatch_Id"]), (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6))
    for e in launches.values():
        d = [ms for _, ms in sorted(e.pop("all"))]
        ns = min(a.serial_decodes, len(d))
        e["ms"], e["ms_concurrent"] = d[len(d) - ns:], d[:len(d) - ns]
    vals = {}
    for d in a.pmc:
        for k, name, c in counters(d).values():
            if k is None:
                continue
            for cn, v in c.items():
                if cn != "GRBM_GUI_ACTIVE":
                    vals.setdefault(k, {}).setdefault(cn, []).append(v)
    scale = {}
    if a.fetch and a.copy_fetch:
        scale["fetch"] = copy_factor(a.copy_fetch, "FETCH_SIZE", a.copy_bytes)
        for k, name, c in counters(a.fetch).values():
            if k is not None:
                vals.setdefault(k, {}).setdefault("fetch_bytes", []).append(c["FETCH_SIZE"] * 1024 * scale["fetch"][0])
    if a.write and a.copy_write:
        scale["write"] = copy_factor(a.copy_write, "WRITE_SIZE", a.copy_bytes)
        for k, name, c in counters(a.write).values():
            if k is not None:
                vals.setdefault(k, {}).setdefault("write_bytes", []).append(c["WRITE_SIZE"] * 1024 * scale["write"][0])

    out = {"clock_ghz": CLOCK_GHZ, "simds": SIMDS, "copy_bytes": a.copy_bytes,
           "fetch_scale": scale.get("fetch", (None,))[0], "write_scale": scale.get("write", (None,))[0],
           "launches": {}}
    tot = {"ms": 0.0, "valu": 0.0}
    for k in sorted(launches):
        t = launches[k]
        ms = mean(t["ms"])
        cyc = ms * 1e-3 * CLOCK_GHZ * 1e9
        v = {cn: mean(x) for cn, x in vals.get(k, {}).items()}
        rec = {"ms": round(ms, 4), "ms_concurrent": round(mean(t["ms_concurrent"]), 4) if t["ms_concurrent"] else None,
               "dispatches": len(t["ms"]) + len(t["ms_concurrent"]), "grid_wg": t["grid_wg"], "vgpr": t["vgpr"],
               "sgpr": t["sgpr"]}
        rec.update({cn: round(x) for cn, x in v.items()})
        if "fetch_bytes" in v and "write_bytes" in v:
            rec["hbm_bytes"] = round(v["fetch_bytes"] + v["write_bytes"])
        if "SQ_INSTS_VALU" in v:
            rec["valu_insts"] = round(v["SQ_INSTS_VALU"])
            rec["valu_issue_frac"] = round(v["SQ_INSTS_VALU"] * 4 / (SIMDS * cyc), 4)
            tot["ms"] += ms
            tot["valu"] += v["SQ_INSTS_VALU"]
        if "SQ_WAVE_CYCLES" in v:
            rec["waves_per_simd"] = round(v["SQ_WAVE_CYCLES"] * 4 / (SIMDS * cyc), 3)
        if v.get("SQ_WAVES"):
            for cn, short in (("SQ_INSTS_VALU", "valu"), ("SQ_INSTS_SALU", "salu"), ("SQ_INSTS_LDS", "lds"),
                              ("SQ_INSTS_SMEM", "smem"), ("SQ_INSTS_BRANCH", "branch")):
                if cn in v:
                    rec[f"{short}_per_wave"] = round(v[cn] / v["SQ_WAVES"], 1)
        if v.get("SQ_LDS_IDX_ACTIVE"):
            rec["lds_conflict"] = round(v.get("SQ_LDS_BANK_CONFLICT", 0) / v["SQ_LDS_IDX_ACTIVE"], 4)
        if "SQ_ACTIVE_INST_VALU" in v and "SQ_WAVE_CYCLES" in v and v["SQ_WAVE_CYCLES"]:
            rec["wave_share_valu"] = round(v["SQ_ACTIVE_INST_VALU"] / v["SQ_WAVE_CYCLES"], 4)
            for cn, short in (("SQ_ACTIVE_INST_LDS", "lds"), ("SQ_ACTIVE_INST_SCA", "salu"),
                              ("SQ_WAIT_INST_ANY", "wait_inst"), ("SQ_WAIT_ANY", "wait_any")):
                if cn in v:
                    rec[f"wave_share_{short}"] = round(v[cn] / v["SQ_WAVE_CYCLES"], 4)
        out["launches"][k] = rec
    if tot["ms"]:
        out["all_launches"] = {"ms": round(tot["ms"], 3), "valu_insts": round(tot["valu"]),
                               "valu_issue_frac": round(tot["valu"] * 4 / (SIMDS * tot["ms"] * 1e-3 * CLOCK_GHZ * 1e9), 4)}
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
