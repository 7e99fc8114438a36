"""The C-ABI libraries: every function include/*.h declares is exported,
struct layouts agree between C and the Python binding, argument checks
that need no device, and graceful failure where no GPU is present.  No
decode runs here."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from helpers import ROOT, ldpcgpu

LIB = os.path.join(ROOT, "dvb-rx_amd", "lib")
HEADERS = {
    "libldpcgpu.so": ["include/ldpc_gpu.h", "include/dvbrx/ldpc.h", "include/ldpc_nms.h"],
    "libdvbs2sim.so": ["include/dvbs2_sim.h"],
}


def declared(header):
    """Function names declared in a C header (comments stripped)."""
    text = open(os.path.join(ROOT, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"^\s*#[^\n]*", "", text, flags=re.M)
    text = re.sub(r"typedef\s+struct\s*\{.*?\}\s*\w+\s*;", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(\w+)\s*\([^;{]*\)\s*;", text)))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    # T: defined in the text section; W: weak (the drop-in's FEC functions)
    return {line.split()[-1] for line in out.splitlines() if " T " in line or " W " in line}


@pytest.mark.parametrize("lib", sorted(HEADERS))
def test_every_declared_symbol_is_exported(lib):
    syms = exported(os.path.join(LIB, lib))
    names = [n for h in HEADERS[lib] for n in declared(h)]
    assert len(names) >= 5
    missing = [n for n in names if n not in syms]
    assert not missing, missing


def test_declared_names():
    names = declared("include/ldpc_gpu.h") + declared("include/dvbrx/ldpc.h")
    for n in ("ldpc_gpu_decode_batch", "ldpc_gpu_create", "drx_ldpc_new", "drx_ldpc_decode", "drx_ldpc_free",
              "drx_ldpc_lanes"):
        assert n in names


def test_frame_struct_layout(tmp_path):
    """ldpc_gpu_frame as the C compiler lays it out == FRAME_DTYPE."""
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "ldpc_gpu.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(ldpc_gpu_frame),"
        " offsetof(ldpc_gpu_frame, code), offsetof(ldpc_gpu_frame, max_iter), offsetof(ldpc_gpu_frame, nbits),"
        " offsetof(ldpc_gpu_frame, group), offsetof(ldpc_gpu_frame, llr_off), offsetof(ldpc_gpu_frame, out_off));"
        "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    d = ldpcgpu.FRAME_DTYPE
    assert vals[0] == d.itemsize
    assert vals[1:] == [d.fields[f][1] for f in ("code", "max_iter", "nbits", "group", "llr_off", "out_off")]


def test_fec_lookup_matches_table():
    L = ldpcgpu.gpu_lib()
    v = [C.c_int() for _ in range(4)]
    assert L.ldpc_gpu_fec_lookup(0, b"1_2", *[C.byref(x) for x in v]) == 0
    assert [x.value for x in v] == [64800, 32400, 32208, 12]
    assert L.ldpc_gpu_fec_lookup(1, b"8_9", *[C.byref(x) for x in v]) == 0
    assert [x.value for x in v] == [16200, 14400, 14232, 12]
    assert L.ldpc_gpu_fec_lookup(0, b"nope", *[C.byref(x) for x in v]) != 0


def test_lanes():
    assert ldpcgpu.gpu_lib().drx_ldpc_lanes() == 64


def _torch_sees_gpu():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_torch_sees_gpu(), reason="checks the no-device path")
def test_no_device_fails_cleanly():
    """Without a usable GPU: create reports a device error, the drop-in
    returns NULL, nothing crashes."""
    L = ldpcgpu.gpu_lib()
    ctx = C.c_void_p()
    rc = L.ldpc_gpu_create(0, C.byref(ctx))
    assert rc in (-3, -2) and not ctx.value
    fec = ldpcgpu.fec_lookup(0, "1_2")
    p = fec.as_struct()
    assert not L.drx_ldpc_new(C.byref(p), b"1_2")
    with pytest.raises(RuntimeError):
        ldpcgpu.Decoder(0)


def test_null_context_is_rejected():
    L = ldpcgpu.gpu_lib()
    frames = np.zeros(1, ldpcgpu.FRAME_DTYPE)
    assert L.ldpc_gpu_decode_batch(None, frames.ctypes.data, 1, None, None, None, None) == -1
    assert L.ldpc_gpu_decode_host(None, frames.ctypes.data, 1, None, 0, None, 0, None, None) == -1


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_workload_composition():
    """The mixed workload: every DVB-S2 MODCOD's rate, about a quarter
    short frames, Es/N0 = QEF + 0.3 dB; deterministic."""
    sys.path.insert(0, os.path.join(ROOT, "dvb-rx_amd"))
    import workloads

    a = workloads.mixed_s2_specs(512, seed=7)
    assert a == workloads.mixed_s2_specs(512, seed=7)
    mods = [ldpcgpu.pls_lookup(s["pls"]) for s in a]
    assert all(s["pls"] < 0x80 for s in a)
    shorts = sum(m["short_frame"] for m in mods)
    assert 0.15 < shorts / len(a) < 0.35
    assert {m["rate"] for m in mods} == set(workloads.S2_NORMAL)
    for s in a[:20]:
        assert abs(s["esn0_db"] - ldpcgpu.qef_db(s["pls"]) - 0.3) < 1e-9


FEC_CASES = [(0, "1_2"), (1, "8_9"), (0, "2_9_VLSNR"), (1, "1_5_VLSNR_SF2"), (0, "2_3_t2"), (0, "7_7")]


def test_dropin_header_fec_lookup(tmp_path):
    """include/dvbrx/ldpc.h on its own declares drx_fec_lookup (and
    drx_bch_*), libldpcgpu.so defines it: a C caller gets the registry's
    parameters, and -1 for an unknown rate (no GPU call)."""
    exe = tmp_path / "fec_header"
    subprocess.run(["gcc", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "fec_header.c"), "-L", os.path.join(ROOT, "dvb-rx_amd", "lib"),
                    "-lldpcgpu", f"-Wl,-rpath,{os.path.join(ROOT, 'dvb-rx_amd', 'lib')}", "-o", str(exe)], check=True)
    args = [x for c in FEC_CASES for x in (str(c[0]), c[1])]
    out = subprocess.run([str(exe)] + args, capture_output=True, text=True, check=True).stdout.split("\n")
    for (short, rate), line in zip(FEC_CASES, out):
        f = line.split()
        if rate == "7_7":
            assert f[2] == "-1"
            continue
        want = ldpcgpu.fec_lookup(short, rate)
        assert [int(x) for x in f[2:]] == [0, want.nldpc, want.kldpc, want.kbch, want.t], line


@pytest.mark.skipif(not os.path.exists("/root/reference/libdvbrx/include/dvbrx/fec.h"),
                    reason="libdvbrx's headers are not mounted here")
def test_dropin_header_with_reference_fec_h(tmp_path):
    """With libdvbrx's include directory on the path the drop-in header
    includes the reference's fec.h, so including fec.h after it (as rx.c
    does) compiles: one drx_fec_params, one set of declarations."""
    subprocess.run(["gcc", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-DWITH_FEC_H", "-I",
                    os.path.join(ROOT, "include"), "-I", "/root/reference/libdvbrx/include",
                    os.path.join(ROOT, "tests", "c", "fec_header.c")], check=True)
