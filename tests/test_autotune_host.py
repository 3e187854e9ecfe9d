"""Host logic of the autotuner: candidate splits and table merging (no GPU)."""
import json
import os

from stntt_amd import autotune as A
from stntt_amd.tuning import gen_tuned as G


def test_pass_candidates_cover_every_stage():
    for L, logc in ((14, 5), (20, 5), (24, 4), (26, 5)):
        cands = A.pass_candidates(L, logc)
        assert cands, L
        for c in cands:
            assert sum(c) == L and all(4 <= p <= 9 for p in c)
            # every pass after the first starts at J >= the tile's columns
            assert all(sum(c[:i]) >= logc for i in range(1, len(c)))
        assert len(cands[0]) <= len(cands[-1])  # fewest passes first
    # below two tile passes there is nothing to tune (the tuner skips these)
    assert A.pass_candidates(9, 5) == [] and A.pass_candidates(7, 4) == []


def test_merge_keeps_entries_not_retuned():
    old = {"device": "gfx950", "pass_split": {"u32": {"20": [6, 7, 7], "24": [6, 9, 9]}, "goldilocks": {"24": [9, 8, 7]}},
           "batch_split": {"u32": {"12": [4, 4, 4]}}, "measured_us": {"pass": {"u32": {"20": {"[6, 7, 7]": 18.0}}}}}
    new = {"device": "gfx950", "pass_split": {"u32": {"24": [8, 8, 8]}, "goldilocks": {}},
           "batch_split": {"u32": {}, "goldilocks": {}}, "measured_us": {"pass": {"u32": {"24": {"[8, 8, 8]": 110.0}}}}}
    m = A.merge(old, new)
    assert m["pass_split"]["u32"] == {"20": [6, 7, 7], "24": [8, 8, 8]}
    assert m["pass_split"]["goldilocks"] == {"24": [9, 8, 7]}
    assert m["batch_split"]["u32"] == {"12": [4, 4, 4]}
    assert m["measured_us"]["pass"]["u32"]["24"] == {"[8, 8, 8]": 110.0}
    assert old["pass_split"]["u32"]["24"] == [6, 9, 9]  # inputs untouched


def test_shipped_table_is_valid():
    with open(A.TABLE) as f:
        table = json.load(f)
    for field, entries in table["pass_split"].items():
        logc = A.word_log2c(field)
        for L, split in entries.items():
            assert sum(split) == int(L) and all(sum(split[:i]) >= logc for i in range(1, len(split))), (field, L)
    assert set(table["pass_split"]) <= set(A.PASS_FIELDS)
    assert set(table["batch_split"]) <= set(A.FIELDS)
    assert os.path.basename(A.TABLE) == "gfx950.json"


def test_compiled_table_matches_json():
    """csrc/tuned_gfx950.hpp (installed by every new context) is the
    rendering of tuning/gfx950.json."""
    with open(A.TABLE) as f:
        table = json.load(f)
    with open(G.HEADER) as f:
        assert f.read() == G.render(table)
    # every tuned split appears as a row of the header
    text = G.render(table)
    for field, entries in table["pass_split"].items():
        for L, split in entries.items():
            row = ", ".join(map(str, list(split) + [0] * (4 - len(split))))
            logc = table.get("pass_logc", {}).get(field, {}).get(L) or [-1] * len(split)
            crow = ", ".join(map(str, list(logc) + [-1] * (4 - len(logc))))
            assert f"{{{G.FIELD_IDS[field]}, {L}, {len(split)}, {{{row}}}, {{{crow}}}}}" in text, (field, L)


def test_table_checks_reject_what_the_library_rejects():
    import pytest
    for field, L, split in (("goldilocks", 24, [8, 8, 8]), ("u32", 24, [9, 7, 8]), ("goldilocks", 13, [9, 4])):
        G.check_pass_split(field, L, split)
    # large-radix passes (2^10..2^12) with the tile widths the kernel takes
    G.check_pass_split("goldilocks", 24, [12, 12], [2, 1])
    G.check_pass_split("u32", 22, [11, 11], [-1, -1])
    G.check_pass_split("goldilocks", 24, [10, 7, 7])
    for field, L, split in (("goldilocks", 26, [13, 13]),    # radix above 2^12
                            ("u32", 12, [4, 8]),             # second pass with J = 2^4 < 32 columns
                            ("u32", 13, [9, 4]),             # n / R = 2^4 < 32 columns
                            ("u32", 24, [8, 8, 7])):         # covers 2^23
        with pytest.raises(AssertionError):
            G.check_pass_split(field, L, split)
    with pytest.raises(AssertionError):
        G.check_pass_split("goldilocks", 24, [12, 12], [3, -1])   # wider than a 2^14-word tile
    with pytest.raises(AssertionError):
        G.check_pass_split("goldilocks", 24, [12, 12], [-1, 4])   # a tile pass width on a large-radix pass
    with pytest.raises(AssertionError):
        G.check_batch_split("u32", 12, [5, 4, 3])  # sub-step radix above 2^4
    with pytest.raises(AssertionError):
        G.check_batch_split("u64", 12, [4, 4, 4])  # no batched 64-bit prime field


def test_big_candidates_fit_the_kernel():
    for L in range(20, 25):
        for split, logc in A.big_candidates(L):
            G.check_pass_split("goldilocks", L, split, logc)
    assert ([12, 12], [2, 2]) in A.big_candidates(24) and not A.big_candidates(19)


def test_measure_alternates_order_and_takes_medians():
    """Every candidate runs once untimed, then rounds alternate the order, so
    no candidate is always timed first (right after a pause, while the GPU
    clock ramps up); the result is the per-candidate median."""
    order = []
    cur = {}
    cands = [[8, 8, 8], [8, 9, 7], [9, 6, 9]]
    med = A.measure(cands, lambda c: cur.update(c=str(c)), lambda: order.append(cur["c"]), lambda: None,
                    reps=2, rounds=3, warm_s=0.0)
    assert set(med) == {str(c) for c in cands} and all(v >= 0 for v in med.values())
    names = [str(c) for c in cands]
    assert len(order) == 3 + 3 * 3 * 2  # one untimed run each, then rounds x candidates x reps
    assert order[:3] == names
    timed = order[3:]
    assert [timed[i] for i in range(0, len(timed), 2)] == names + names[::-1] + names
    assert all(timed[i] == timed[i + 1] for i in range(0, len(timed), 2))


def test_single_pass_sizes_match_the_library():
    """The tuner skips the sizes the library always runs in one pass; the
    limits it assumes are the ones csrc/plan.hpp compiles in."""
    import re
    src = open(os.path.join(os.path.dirname(A.__file__), "csrc", "plan.hpp")).read()
    m = re.search(r"constexpr int narrow_max_logr\(\) \{\s*return sizeof\(typename FP::T\) == 4 \? (\d+) : (\d+);", src)
    assert m, "narrow_max_logr not found"
    assert A.single_pass_max_log("u32") == int(m.group(1))
    assert A.single_pass_max_log("goldilocks") == int(m.group(2))
    assert A.single_pass_max_log("u64") == int(m.group(2))
