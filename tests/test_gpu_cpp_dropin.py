"""Runs tests/cpp/st_ntt_test: the reference's st_ntt_test.cu call sequence
through the C++ drop-in header (include/stntt/ntt_kernel.hpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(__file__), "cpp", "st_ntt_test")


def test_cpp_dropin_replays_reference_test():
    assert os.path.exists(BIN), "build() compiles tests/cpp/st_ntt_test"
    print("running", BIN, flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=110)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "0 failures" in r.stdout
