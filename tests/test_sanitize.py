"""AddressSanitizer + UndefinedBehaviorSanitizer build of the CPU oracle.

The oracle's C restatement (oracle/fft_oracle.c) is compiled with
``-fsanitize=address,undefined -fno-sanitize-recover=all`` together with a
host-only driver (tests/sanitize/oracle_sanitize_main.c) that calls every entry
point on exact-size buffers: prime, smooth and Bluestein lengths, n-D views
with negative and padded byte strides, in-place calls, halfcomplex, r2r and
upfirdn.  Any out-of-bounds access, use after free, leak, signed overflow or
misaligned access ends the run with a report, and the test fails.  Host code
only: no GPU code is built with a sanitizer.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_asan_ubsan(tmp_path):
    exe = tmp_path / "oracle_sanitize"
    cmd = ["gcc", "-std=c11", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "oracle"),
           os.path.join(HERE, "sanitize", "oracle_sanitize_main.c"), os.path.join(ROOT, "oracle", "fft_oracle.c"),
           "-lm", "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr
    # verify_asan_link_order=0: the environment may preload other libraries ahead of the runtime
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "oracle sanitizer run clean" in r.stdout
