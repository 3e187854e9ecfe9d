"""The device field arithmetic (stntt_amd/csrc/field_*.hpp) compiled for the
host and checked against 128-bit integer arithmetic on edge and random
values (tests/cpp/field_test.cpp).  CPU only."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "field_test")


def _stale() -> bool:
    """the binary is missing or older than the test or a field header"""
    if not os.path.exists(BIN):
        return True
    csrc = os.path.join(ROOT, "stntt_amd", "csrc")
    srcs = [os.path.join(ROOT, "tests", "cpp", "field_test.cpp")] + [
        os.path.join(csrc, f) for f in os.listdir(csrc) if f.startswith("field_")]
    return any(os.path.getmtime(s) > os.path.getmtime(BIN) for s in srcs)


def test_field_arithmetic_exact():
    if _stale():
        subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++20", "-o", BIN,
                        os.path.join(ROOT, "tests", "cpp", "field_test.cpp")], check=True, capture_output=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
