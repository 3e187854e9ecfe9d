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


CSV_BIN = os.path.join(os.path.dirname(__file__), "..", "stntt_amd", "bench_csv")


def test_bench_csv_has_reference_columns():
    """bench_csv prints src/bench.cu's CSV; metrics/stats.py's grouping
    (median per log2_N and Block size) applies to it unchanged."""
    import io

    import pandas as pd

    assert os.path.exists(CSV_BIN), "build() compiles stntt_amd/bench_csv"
    print("running", CSV_BIN, "2", flush=True)
    r = subprocess.run([CSV_BIN, "2"], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr[-2000:]
    df = pd.read_csv(io.StringIO(r.stdout))
    assert list(df.columns) == ["log2_N", "Block size", "Kernel time", "findParams time", "initArgs time",
                                "Kernel+transfers time", "Total time"]
    assert len(df) == 5 * 19 * 2  # block sizes x sizes 2^2..2^20 x timed repetitions
    med = df.groupby(["log2_N", "Block size"]).median()
    assert len(med) >= 19
    assert (df["Kernel time"] > 0).all() and (df["Total time"] >= df["Kernel+transfers time"]).all()
    # the reference's geometry: min(n/2, max block) threads for batch 1
    row = df[(df["log2_N"] == 12)]
    assert set(row["Block size"]) == {64, 128, 256, 512, 1024}
