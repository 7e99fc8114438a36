"""dvb-rx_amd/prof_counters.py on small synthetic rocprofv3 outputs: the
decoder's dispatches are labelled with the decode's launch order, the
per-launch durations come from the serial pass, and the counters become
issue fractions and calibrated HBM bytes."""
import csv
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "dvb-rx_amd", "prof_counters.py")

K2 = "void lgk::(anonymous namespace)::ldpc_sk_kernel<2, 0, 0, 1>(lgk::PkArgs)"
K3 = "void lgk::(anonymous namespace)::ldpc_sk_kernel<3, 0, 0, 2>(lgk::PkArgs)"
ORDER = ["ldpc_sk_kernel<2, 0, 0, 1> normal x10", "ldpc_sk_kernel<2, 0, 0, 1> short x4",
         "ldpc_sk_kernel<3, 0, 0, 2> short x6"]
DECODES = 3  # 1 concurrent, then 2 serial
DUR = {0: 2.0, 1: 0.5, 2: 1.0}  # serial ms of each launch


def _write(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(header)
        w.writerows(rows)


def _dispatches():
    """(dispatch id, kernel, launch index, decode) in dispatch order, with a
    copy kernel between decodes"""
    out, did = [], 1
    for d in range(DECODES):
        for k, name in enumerate([K2, K2, K3]):
            out.append((did, name, k, d))
            did += 1
        out.append((did, "copyBuffer", -1, d))
        did += 1
    return out


@pytest.fixture()
def runs(tmp_path):
    disp = _dispatches()
    tr = []
    t = 0
    for did, name, k, d in disp:
        ms = 0.01 if k < 0 else (DUR[k] * (3.0 if d == 0 else 1.0))
        tr.append(["KERNEL_DISPATCH", did, name, t, t + int(ms * 1e6), 512, 0, 48, 0, 112, 720, 720 * 256])
        t += int(ms * 1e6) + 1000
    _write(str(tmp_path / "trace" / "run_kernel_trace.csv"),
           ["Kind", "Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp", "LDS_Block_Size",
            "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count", "SGPR_Count", "Workgroup_Size_X", "Grid_Size_X"], tr)
    pm = []
    for did, name, k, d in disp:
        if k < 0:
            continue
        # VALU: launch k issues (k + 1) * 1e9 instructions, over two dimensions
        for dim in (0, 1):
            pm.append([did, name, "SQ_INSTS_VALU", (k + 1) * 0.5e9])
            pm.append([did, name, "SQ_WAVES", 1536.0])
    _write(str(tmp_path / "pmc" / "run_counter_collection.csv"),
           ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"], pm)
    for cn, kib in (("FETCH_SIZE", 0.5), ("WRITE_SIZE", 1.0)):
        rows = [[did, name, cn, kib * 1024 * 1024 * (k + 1)] for did, name, k, d in disp if k >= 0]
        _write(str(tmp_path / cn.lower() / "run_counter_collection.csv"),
               ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"], rows)
        # the copy of 1 GiB reads half its bytes under FETCH_SIZE, all under WRITE_SIZE
        _write(str(tmp_path / ("c" + cn.lower()) / "run_counter_collection.csv"),
               ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"],
               [[1, "elementwise_copy", cn, (kib * 1024 * 1024)]])
    with open(tmp_path / "bench.json", "w") as fh:
        fh.write("# log line\n" + json.dumps({"roofline": {"launch_order": ORDER}}) + "\n")
    return tmp_path


def _run(p, *extra):
    out = p / "out.json"
    r = subprocess.run([sys.executable, TOOL, "--trace", str(p / "trace"), "--order", str(p / "bench.json"),
                        "--pmc", str(p / "pmc"), "--fetch", str(p / "fetch_size"), "--write", str(p / "write_size"),
                        "--copy-fetch", str(p / "cfetch_size"), "--copy-write", str(p / "cwrite_size"),
                        "--copy-bytes", str(1 << 30), "--out", str(out), *extra],
                       capture_output=True, text=True)
    return r, out


def test_labels_durations_and_counters(runs):
    r, out = _run(runs)
    assert r.returncode == 0, r.stderr
    d = json.load(open(out))
    L = d["launches"]
    assert sorted(L) == sorted(x.rsplit(" x", 1)[0] for x in ORDER)
    for k, key in enumerate(x.rsplit(" x", 1)[0] for x in ORDER):
        rec = L[key]
        assert rec["ms"] == pytest.approx(DUR[k], abs=1e-3)  # the serial pass only
        assert rec["ms_concurrent"] == pytest.approx(3 * DUR[k], abs=1e-3)
        assert rec["valu_insts"] == pytest.approx((k + 1) * 1e9, rel=1e-6)
        cyc = DUR[k] * 1e-3 * 2.4e9
        assert rec["valu_issue_frac"] == pytest.approx((k + 1) * 1e9 * 4 / (1024 * cyc), rel=1e-3)
        # FETCH scaled x2, WRITE x1 (the copy's readings of its known size)
        assert rec["hbm_bytes"] == pytest.approx((k + 1) * (0.5 * 2 + 1.0) * 1024 ** 3, rel=1e-6)
    assert d["fetch_scale"] == pytest.approx(2.0) and d["write_scale"] == pytest.approx(1.0)


def test_order_that_does_not_match_is_an_error(runs):
    with open(runs / "bench.json", "w") as fh:
        fh.write(json.dumps({"roofline": {"launch_order": list(reversed(ORDER))}}) + "\n")
    r, _ = _run(runs)
    assert r.returncode != 0 and "does not match" in (r.stderr + r.stdout)
