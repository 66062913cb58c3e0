"""Runs tests/cpp/test_api: the reference's own gtest cases (test_fft.cpp,
test_dispatch.cpp, test_real_fft.cpp, test_dit.cpp, with DummyData and FFT_Eps
as in test/test_util.h) compiled against the drop-in <genFFT/fft.h> header,
which resolves to the gfx950 engine through the C ABI."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_api")


def test_reference_cases_through_cpp_header(gpu):
    assert os.path.exists(BIN), "tests/cpp/test_api not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=600)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "0 failed" in r.stdout
