"""Runs the reference's FFT tests restated in C++ against the C++ host mirror
(tests/cpp/test_fft_cpp.cpp over impulse_amd/host/impulse_fft.hpp)."""
import os
import subprocess

import pytest

from tests.conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_mirror_reference_tests():
    exe = os.path.join(ROOT, "tests", "cpp", "test_fft_cpp")
    assert os.path.exists(exe), "build the C++ test program first (__graft_entry__.build())"
    r = subprocess.run([exe, "8192", "2e-15"], capture_output=True, text=True, timeout=600)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
