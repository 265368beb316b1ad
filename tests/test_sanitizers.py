"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only, gcc/g++ builds):

- the oracle's entry points on small cases (oracle/xc_oracle.c with its own FFT);
- the serving loop's scheduler pieces (xc_sched.hpp: first-fit arena, request queue with deadlines and
  cancellation), including 20000 random allocations and releases;
- the ThmPair table writer and readers (xc_thmpair.hip compiled as C++), with over-long lines;
- the C++ drop-in's host-only functions (FFTSize, FFTSizeSQR, ParabPeakFFT at the image borders).

Each driver (tests/sanitize/*) aborts on the first sanitizer report or failed check."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
LIB = os.path.join(ROOT, "alignment_projects_amd", "lib")


def _have_asan():
    if not shutil.which("gcc") or not shutil.which("g++"):
        return False
    out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return os.path.isabs(out) and os.path.exists(out)


pytestmark = pytest.mark.skipif(not _have_asan(), reason="gcc with libasan not available")


def _run(cmd, exe, args=(), env_extra=None):
    subprocess.run(cmd, check=True, cwd=ROOT)
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    env.update(env_extra or {})
    r = subprocess.run([exe, *args], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr
    return r.stdout


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_check")
    out = _run(["gcc", *SAN, "-std=gnu11", "-Iinclude", "oracle/xc_oracle.c", "tests/sanitize/oracle_check.c", "-o",
                exe, "-lm", "-ldl", "-lpthread"], exe)
    assert "oracle ok" in out


def test_scheduler_under_sanitizers(tmp_path):
    exe = str(tmp_path / "sched_check")
    out = _run(["g++", *SAN, "-std=c++17", "tests/sanitize/sched_check.cpp", "-o", exe], exe)
    assert "sched ok" in out


def test_thmpair_under_sanitizers(tmp_path):
    exe = str(tmp_path / "thmpair_check")
    d = tmp_path / "tables"
    d.mkdir()
    out = _run(["g++", *SAN, "-std=c++17", "-x", "c++", "alignment_projects_amd/csrc/xc_thmpair.hip", "-x", "none",
                "tests/sanitize/thmpair_check.cpp", "-o", exe], exe, [str(d)])
    assert "thmpair ok" in out


def test_drop_in_host_functions_under_sanitizers(tmp_path):
    if not os.path.exists(os.path.join(LIB, "libxcorr.so")):
        pytest.skip("libxcorr.so not built")
    exe = str(tmp_path / "shim_check")
    # the HIP runtime's own allocations in libxcorr.so are not ours to leak-check
    out = _run(["g++", *SAN, "-std=c++11", "-Iinclude", "alignment_projects_amd/compat/aln_corr_compat.cpp",
                "tests/sanitize/shim_check.cpp", "-L" + LIB, "-lxcorr", "-Wl,-rpath," + LIB, "-o", exe], exe,
               env_extra={"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1"})
    assert "shim ok" in out
