"""Compile the tuning table into the library.

Reads gfx950.json (written by ``python -m stntt_amd.autotune``) and writes
``csrc/tuned_gfx950.hpp``, which every new context installs, so callers of
the C ABI get the tuned splits without the Python package.  ``build()``
runs this before compiling; a CPU test checks the committed header is in
sync with the table.

    python -m stntt_amd.tuning.gen_tuned
"""
from __future__ import annotations

import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "gfx950.json")
HEADER = os.path.join(os.path.dirname(HERE), "csrc", "tuned_gfx950.hpp")
FIELD_IDS = {"u32": 0, "goldilocks": 1, "u64": 2}  # stntt_field
MIN_LOGR_TILE, MAX_LOGR_TILE = 4, 9  # csrc/plan.hpp
MIN_LOGR_BIG, MAX_LOGR_BIG = 10, 12   # csrc/plan.hpp: the large-radix kernel
BIG_TILE_LOG_WORDS = 14               # csrc/ntt_big.hpp: a tile holds 2^14 words


def check_pass_split(field: str, L: int, split: list[int], logc: list[int] | None = None) -> None:
    """The rules stntt_set_pass_shape enforces (a rejected entry would make
    every context creation fail, so they are checked at build time)."""
    tile_logc = 5 if field == "u32" else 4  # 128-byte rows of 4- or 8-byte words
    logc = logc or [-1] * len(split)
    assert 2 <= len(split) <= 4 and sum(split) == L and len(logc) == len(split), (field, L, split, logc)
    for i, (r, c) in enumerate(zip(split, logc)):
        j = sum(split[:i])
        if r <= MAX_LOGR_TILE:
            assert r >= MIN_LOGR_TILE and c in (-1, tile_logc), (field, L, split, logc)
            assert L - r >= tile_logc and (i == 0 or j >= tile_logc), (field, L, split)  # n/R >= C, J >= C
        else:
            assert r <= MAX_LOGR_BIG, (field, L, split)
            full = BIG_TILE_LOG_WORDS - r
            widths = [c] if c >= 0 else range(full, full - 3, -1)
            assert any(full - 2 <= w <= full and L - r >= w and (i == 0 or j >= w) for w in widths), \
                (field, L, split, logc)


def check_batch_split(field: str, L: int, logs: list[int]) -> None:
    """The rules of batch_validate_split (LDS-resident path)."""
    assert field in ("u32", "goldilocks"), field
    assert 1 <= len(logs) <= 8 and sum(logs) == L and all(1 <= x <= 4 for x in logs), (field, L, logs)


def render(table: dict) -> str:
    lines = [
        "// Generated by stntt_amd/tuning/gen_tuned.py from stntt_amd/tuning/gfx950.json",
        "// (python -m stntt_amd.autotune on an MI355X).  Do not edit by hand.",
        "#pragma once",
        "",
        "#include <cstdint>",
        "",
        "namespace stntt {",
        "",
        "// stages per pass of the multi-pass path for one transform, and the tile",
        "// columns (log2) of large-radix passes (-1: the library's choice)",
        "struct TunedPassSplit {",
        "    int field;",
        "    uint32_t log2n;",
        "    uint32_t npasses;",
        "    uint32_t logr[4];",
        "    int32_t logc[4];",
        "};",
        "",
        "// sub-step radices (log2, 4 bits each from the lowest) of the batched kernel",
        "struct TunedBatchSplit {",
        "    int field;",
        "    uint32_t log2n;",
        "    uint32_t nsub;",
        "    uint32_t radix_logs;",
        "};",
        "",
        "inline constexpr TunedPassSplit TUNED_PASS_SPLITS[] = {",
    ]
    for field, entries in sorted(table.get("pass_split", {}).items(), key=lambda kv: FIELD_IDS[kv[0]]):
        for L, split in sorted(entries.items(), key=lambda kv: int(kv[0])):
            logc = table.get("pass_logc", {}).get(field, {}).get(L) or [-1] * len(split)
            check_pass_split(field, int(L), split, logc)
            padded = list(split) + [0] * (4 - len(split))
            cpad = list(logc) + [-1] * (4 - len(logc))
            lines.append(f"    {{{FIELD_IDS[field]}, {int(L)}, {len(split)}, {{{', '.join(map(str, padded))}}}, "
                         f"{{{', '.join(map(str, cpad))}}}}},  // {field}")
    lines += ["};", "", "inline constexpr TunedBatchSplit TUNED_BATCH_SPLITS[] = {"]
    for field, entries in sorted(table.get("batch_split", {}).items(), key=lambda kv: FIELD_IDS[kv[0]]):
        for L, logs in sorted(entries.items(), key=lambda kv: int(kv[0])):
            check_batch_split(field, int(L), logs)
            packed = 0
            for i, lr in enumerate(logs):
                packed |= lr << (4 * i)
            lines.append(f"    {{{FIELD_IDS[field]}, {int(L)}, {len(logs)}, 0x{packed:x}u}},  // {field} {logs}")
    lines += ["};", "", "}  // namespace stntt", ""]
    return "\n".join(lines)


def generate(table_path: str = TABLE, header_path: str = HEADER) -> str:
    with open(table_path) as f:
        text = render(json.load(f))
    old = open(header_path).read() if os.path.exists(header_path) else None
    if old != text:
        with open(header_path, "w") as f:
            f.write(text)
    return header_path


if __name__ == "__main__":
    print(generate())
