"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer.

atlas-engine_amd/Makefile.sanitize compiles the C ABI's host sources and the
oracle's sources with -fsanitize=address,undefined (host compiler only; the
device objects are the ordinary ones) into tests/cpp/sanitize_test.cpp's
program, which builds BLASes and AABB trees on the host, traces them with
the engine's CPU traversal and compares every result with the oracle (soups
of 1 to 20,000 triangles, a heightfield with equal-t ties, degenerate
triangles, tmin > 0, argument errors).  Any sanitizer report fails the run.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "atlas-engine_amd")
BIN = os.path.join(PKG, "_build_san", "sanitize_test")


def test_host_code_is_clean_under_asan_and_ubsan():
    b = subprocess.run(["make", "-f", "Makefile.sanitize", "-j8"], cwd=PKG, capture_output=True, text=True,
                       timeout=900)
    assert b.returncode == 0, b.stdout[-3000:] + b.stderr[-3000:]
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1",
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    p = subprocess.run([BIN], capture_output=True, text=True, env=env, timeout=600)
    out = p.stdout + p.stderr
    assert p.returncode == 0, out[-4000:]
    assert "address sanitizer on" in p.stdout
    assert "sanitize ok" in p.stdout
    assert "runtime error" not in out and "ERROR: AddressSanitizer" not in out
