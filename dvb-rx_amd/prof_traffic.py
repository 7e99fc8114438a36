#!/usr/bin/env python3
"""HBM traffic per decode launch from rocprofv3 counter runs.

Inputs: four rocprofv3 --pmc output directories (csv format):

  --fetch DIR        bench.py run with --pmc FETCH_SIZE
  --write DIR        bench.py run with --pmc WRITE_SIZE
  --copy-fetch DIR   bench.py --copy-calibration run with --pmc FETCH_SIZE
  --copy-write DIR   bench.py --copy-calibration run with --pmc WRITE_SIZE

FETCH_SIZE and WRITE_SIZE are in KiB per dispatch.  Each is scaled by the
factor that maps the plain copy kernel's reading to its known size (bytes
read = bytes written = the copy size), then the decode launches are
reported per kernel instance and frame size.

Writes a JSON summary (--out) with the per-launch bytes and the scale
factors.
"""
import argparse
import csv
import glob
import json
import os
import re
import sys


def rows(d):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        sys.exit(f"no counter_collection.csv under {d}")
    out = []
    for f in files:
        with open(f) as fh:
            out.extend(csv.DictReader(fh))
    return out


def per_dispatch(d, counter):
    """{dispatch id: (kernel name, lds bytes, grid, value)} for one counter."""
    acc = {}
    for r in rows(d):
        if r.get("Counter_Name") != counter:
            continue
        did = int(r["Dispatch_Id"])
        name = r["Kernel_Name"]
        lds = int(r.get("LDS_Block_Size") or r.get("Lds_Block_Size") or 0)
        grid = int(r.get("Grid_Size") or 0)
        v = float(r["Counter_Value"])
        k = acc.get(did)
        acc[did] = (name, lds, grid, (k[3] if k else 0.0) + v)
    return acc


def copy_factor(d, counter, copy_bytes):
    disp = per_dispatch(d, counter)
    vals = [v for (name, lds, grid, v) in disp.values() if "copy" in name.lower() or "elementwise" in name.lower()]
    vals = [v for v in vals if v * 1024 > copy_bytes / 4]  # the big copies, not the fill of small tensors
    if not vals:
        sys.exit(f"no copy kernel found in {d}")
    vals.sort()
    measured = vals[len(vals) // 2] * 1024.0
    return copy_bytes / measured, measured


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--copy-fetch", required=True)
    ap.add_argument("--copy-write", required=True)
    ap.add_argument("--copy-bytes", type=int, required=True)
    ap.add_argument("--kernel", default="ldpc_pk_kernel")
    ap.add_argument("--labels", default="normal,short", type=lambda x: x.split(","))
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    ff, fm = copy_factor(a.copy_fetch, "FETCH_SIZE", a.copy_bytes)
    wf, wm = copy_factor(a.copy_write, "WRITE_SIZE", a.copy_bytes)
    fetch = per_dispatch(a.fetch, "FETCH_SIZE")
    write = per_dispatch(a.write, "WRITE_SIZE")
    # one decode call launches each kernel instance once per frame size
    # with frames, normal-frame launch first (the launch order of the
    # decoder); label the dispatches of an instance in that order
    launches = {}
    for src, key, fac in ((fetch, "fetch", ff), (write, "write", wf)):
        seen = {}
        for did in sorted(src):
            name, lds, grid, v = src[did]
            # instance <class, learned>: the plain decoder's is named
            # "<kernel><class>", the learned one's "<kernel><class, learned>"
            m = re.search(a.kernel + r"<(\d+)(?:, *(true|false))?>", name)
            if not m:
                continue
            inst = f"{a.kernel}<{m.group(1)}{', learned' if m.group(2) == 'true' else ''}>"
            i = seen.get(inst, 0)
            seen[inst] = i + 1
            label = a.labels[i % len(a.labels)]
            e = launches.setdefault(f"{inst} {label}", {"fetch": [], "write": [], "grid": grid})
            e[key].append(v * 1024.0 * fac)
    summary = {
        "copy_bytes": a.copy_bytes,
        "fetch_scale": ff, "copy_fetch_measured": fm,
        "write_scale": wf, "copy_write_measured": wm,
        "launches": {},
    }
    for k, e in sorted(launches.items()):
        f = sum(e["fetch"]) / max(1, len(e["fetch"]))
        w = sum(e["write"]) / max(1, len(e["write"]))
        summary["launches"][k] = {"fetch_bytes": round(f), "write_bytes": round(w), "bytes": round(f + w),
                                  "dispatches": len(e["fetch"]), "grid_threads": e["grid"]}
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as fh:
        json.dump(summary, fh, indent=1)
    print(json.dumps(summary, indent=1))


if __name__ == "__main__":
    main()
