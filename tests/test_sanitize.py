"""The host C code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/c/sanitize_host.c drives the code registry (every code's tables),
the frame source, the oracle and the drop-in shim's host paths; it is built
here from the sources with gcc -fsanitize=address,undefined (every report
fatal) and must run clean.  The shim's GPU calls go to libldpcgpu.so (not
instrumented); on a host without a GPU they fail as they should, which is
one of the paths checked.
"""
import os
import subprocess

import pytest

from helpers import ROOT

SRC = ["tests/c/sanitize_host.c", "dvb-rx_amd/csrc/ldpc_codes.c", "dvb-rx_amd/csrc/dvbs2_sim.c",
       "oracle/ldpc_oracle.c", "dvb-rx_amd/csrc/drx_ldpc_shim.c"]


def test_host_code_under_asan_ubsan(tmp_path):
    exe = tmp_path / "sanitize_host"
    cmd = ["gcc", "-O1", "-g", "-std=gnu11", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-I", "include", "-I", "dvb-rx_amd/csrc", "-I", "oracle"] + SRC + \
          ["-L", "dvb-rx_amd/lib", "-lldpcgpu", f"-Wl,-rpath,{os.path.join(ROOT, 'dvb-rx_amd', 'lib')}", "-lm",
           "-lpthread", "-o", str(exe)]
    b = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if b.returncode and "asan" in b.stderr.lower() and "cannot find" in b.stderr.lower():
        pytest.skip("no sanitizer runtime for this compiler")
    assert b.returncode == 0, b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", ""))
    r = subprocess.run([str(exe)], cwd=ROOT, capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr
    assert "sanitized host run ok" in r.stdout
