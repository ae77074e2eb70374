"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU,
no GPU): tests/sanitize/host_checks.cpp drives

* the oracle (g++ -fsanitize=address,undefined): the per-frame pose
  estimator with its trace for every loss, PROSAC, 3-point and degenerate
  frames, frame-parallel estimation on 3 threads, the samplers, the stop
  rule, the cubic solver and the fundamental-matrix estimator;
* the engine's host side (hipcc, sanitizers on the host pass only): the
  tuning-table parser, boundary validation, PROSAC schedules, batch plans
  of 300 tiny frames, mixed sizes and N = 200000, one-frame plans inside
  reserved capacity, and the packing of their metadata.

A sanitizer report ends the program with a non-zero status.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}"
    return r


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_checks")
    _run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
          "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-ffp-contract=off",
          "-DTR_CHECK_ORACLE", "-I", os.path.join(ROOT, "include"),
          "-I", os.path.join(ROOT, "oracle"), SRC, os.path.join(ROOT, "oracle", "oracle.cpp"),
          "-lpthread", "-o", exe], ROOT)
    r = subprocess.run([exe, "oracle"], capture_output=True, text=True, env=ENV, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="no hipcc")
def test_engine_host_under_asan_ubsan(tmp_path):
    csrc = os.path.join(ROOT, "turboransac_amd", "csrc")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", csrc]
    objs = []
    procs = []
    for src in ("kernels.hip", "engine.hip"):
        obj = str(tmp_path / (src + ".o"))
        objs.append(obj)
        procs.append(subprocess.Popen(
            [HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             *SAN, *inc, "-c", os.path.join(csrc, src), "-o", obj],
            stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
    hc = str(tmp_path / "host_checks.o")
    rocm_inc = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(HIPCC))), "include")
    _run([HIPCC, "-O1", "-g", "-std=c++17", *SAN, "-D__HIP_PLATFORM_AMD__", "-DTR_CHECK_ENGINE",
          *inc, "-I", rocm_inc, "-x", "c++", "-c", SRC, "-o", hc], ROOT)
    for p in procs:
        out, _ = p.communicate()
        assert p.returncode == 0, out[-4000:]
    exe = str(tmp_path / "engine_checks")
    _run([HIPCC, "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
          hc, *objs, "-o", exe], ROOT)
    r = subprocess.run([exe, "engine"], capture_output=True, text=True, env=ENV, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
