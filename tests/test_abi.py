"""The C ABI library loads, exports every declared symbol, and its host-only
logic works without a GPU (no compute calls here)."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

from oracle import oracle as O
from stntt_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "stntt.h")


def declared_symbols() -> set[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(stntt_\w+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "build() first"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (stntt_\w+)", out))
    missing = declared_symbols() - exported
    assert not missing, missing
    # and the Python binding table covers exactly the header
    assert {name for name, _, _ in _lib.SIGNATURES} == declared_symbols()


def test_library_loads():
    L = _lib.lib()
    assert L.stntt_version().startswith(b"stntt")


@pytest.mark.parametrize("L", [1, 5, 12, 16, 20, 24, 27])
def test_find_params_matches_oracle(L):
    from stntt_amd import find_params
    assert find_params(1 << L, 11) == O.find_params(1 << L, 11)


def test_find_params_min_mod():
    from stntt_amd import find_params
    for n, mm in [(16, 1000), (64, 81 * 64 + 1), (1024, 2**30)]:
        assert find_params(n, mm) == O.find_params(n, mm)


def test_errors_without_device():
    L = _lib.lib()
    r, m = ctypes.c_uint64(), ctypes.c_uint64()
    assert L.stntt_find_params(0, 11, ctypes.byref(r), ctypes.byref(m)) == _lib.STNTT_EINVAL
    assert b"positive" in L.stntt_last_error()
    # null context is rejected before any device call
    assert L.stntt_ntt_goldilocks(None, None, 4, 1, 0, None) == _lib.STNTT_EINVAL
    assert L.stntt_ntt_u32(None, None, 4, 1, 3, 17, 0, 0, None) == _lib.STNTT_EINVAL
