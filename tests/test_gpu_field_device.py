"""The field arithmetic of stntt_amd/csrc/field_*.hpp executed on the GPU on
edge and random values (tests/cpp/field_device_test.hip), checked against
128-bit integer arithmetic on the host: the Goldilocks product, reduction,
add, sub and every power-of-two shift, the 64-bit Montgomery field for
primes up to 2^64 - 59, and the 32-bit Montgomery field.  The borrow and
carry paths that random transform data almost never reaches are fed
directly."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "field_device_test")


def test_device_field_arithmetic_exact():
    assert os.path.exists(BIN), "tests/cpp/field_device_test is not built (__graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert " 0 failures" in r.stdout
