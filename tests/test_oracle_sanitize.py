"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the
driver tests/oracle_sanitize.cpp, compiled with oracle/oracle.cpp into one
instrumented executable (no preloading), takes every oracle entry point the
tests use through small inputs (each disk and error mode, supersampled RGB,
the Hittable scene, the HDRI sky, a tensor metric, the spin clamp, invalid and
exhausted caller states, the dual-number gradient).  Any sanitizer finding
aborts the run; leaks are reported at exit."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_clean_under_asan_and_ubsan(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not found")
    probe = subprocess.run([cxx, "-print-file-name=libasan.so"], capture_output=True, text=True)
    if probe.returncode != 0 or not os.path.isabs(probe.stdout.strip()):
        pytest.skip("g++ has no AddressSanitizer runtime")
    exe = str(tmp_path / "oracle_sanitize")
    subprocess.run([cxx, "-std=c++17", "-O1", "-g", "-ffp-contract=off", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-pthread",
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "oracle"),
                    os.path.join(ROOT, "tests", "oracle_sanitize.cpp"), os.path.join(ROOT, "oracle", "oracle.cpp"),
                    "-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout)
    print(r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "oracle sanitize: ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
