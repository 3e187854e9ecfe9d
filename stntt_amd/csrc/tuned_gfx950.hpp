// Generated by stntt_amd/tuning/gen_tuned.py from stntt_amd/tuning/gfx950.json
// (python -m stntt_amd.autotune on an MI355X).  Do not edit by hand.
#pragma once

#include <cstdint>

namespace stntt {

// stages per tile pass of the multi-pass path for one transform
struct TunedPassSplit {
    int field;
    uint32_t log2n;
    uint32_t npasses;
    uint32_t logr[4];
};

// sub-step radices (log2, 4 bits each from the lowest) of the batched kernel
struct TunedBatchSplit {
    int field;
    uint32_t log2n;
    uint32_t nsub;
    uint32_t radix_logs;
};

inline constexpr TunedPassSplit TUNED_PASS_SPLITS[] = {
    {0, 10, 2, {5, 5, 0, 0}},  // u32
    {0, 11, 2, {6, 5, 0, 0}},  // u32
    {0, 12, 2, {6, 6, 0, 0}},  // u32
    {0, 13, 2, {6, 7, 0, 0}},  // u32
    {0, 14, 2, {7, 7, 0, 0}},  // u32
    {0, 15, 3, {6, 4, 5, 0}},  // u32
    {0, 16, 3, {6, 6, 4, 0}},  // u32
    {0, 17, 3, {6, 6, 5, 0}},  // u32
    {0, 18, 3, {6, 6, 6, 0}},  // u32
    {0, 19, 3, {6, 6, 7, 0}},  // u32
    {0, 20, 3, {7, 6, 7, 0}},  // u32
    {0, 21, 3, {8, 5, 8, 0}},  // u32
    {0, 22, 3, {9, 7, 6, 0}},  // u32
    {0, 23, 3, {9, 8, 6, 0}},  // u32
    {0, 24, 3, {9, 7, 8, 0}},  // u32
    {0, 25, 3, {9, 9, 7, 0}},  // u32
    {0, 26, 3, {9, 9, 8, 0}},  // u32
    {1, 8, 2, {4, 4, 0, 0}},  // goldilocks
    {1, 9, 2, {5, 4, 0, 0}},  // goldilocks
    {1, 10, 2, {6, 4, 0, 0}},  // goldilocks
    {1, 11, 2, {5, 6, 0, 0}},  // goldilocks
    {1, 12, 2, {6, 6, 0, 0}},  // goldilocks
    {1, 13, 3, {5, 4, 4, 0}},  // goldilocks
    {1, 14, 3, {6, 4, 4, 0}},  // goldilocks
    {1, 15, 3, {6, 4, 5, 0}},  // goldilocks
    {1, 16, 3, {6, 4, 6, 0}},  // goldilocks
    {1, 17, 3, {6, 5, 6, 0}},  // goldilocks
    {1, 18, 3, {6, 6, 6, 0}},  // goldilocks
    {1, 19, 3, {8, 6, 5, 0}},  // goldilocks
    {1, 20, 3, {6, 6, 8, 0}},  // goldilocks
    {1, 21, 3, {8, 6, 7, 0}},  // goldilocks
    {1, 22, 3, {8, 6, 8, 0}},  // goldilocks
    {1, 23, 3, {8, 8, 7, 0}},  // goldilocks
    {1, 24, 3, {8, 8, 8, 0}},  // goldilocks
    {1, 25, 3, {8, 8, 9, 0}},  // goldilocks
    {1, 26, 3, {8, 9, 9, 0}},  // goldilocks
    {2, 8, 2, {4, 4, 0, 0}},  // u64
    {2, 9, 2, {4, 5, 0, 0}},  // u64
    {2, 10, 2, {4, 6, 0, 0}},  // u64
    {2, 11, 2, {6, 5, 0, 0}},  // u64
    {2, 12, 2, {6, 6, 0, 0}},  // u64
    {2, 13, 3, {4, 4, 5, 0}},  // u64
    {2, 14, 3, {4, 6, 4, 0}},  // u64
    {2, 15, 3, {4, 6, 5, 0}},  // u64
    {2, 16, 3, {6, 4, 6, 0}},  // u64
    {2, 17, 3, {6, 6, 5, 0}},  // u64
    {2, 18, 3, {6, 6, 6, 0}},  // u64
    {2, 19, 3, {8, 6, 5, 0}},  // u64
    {2, 20, 3, {6, 6, 8, 0}},  // u64
    {2, 21, 3, {5, 8, 8, 0}},  // u64
    {2, 22, 3, {6, 8, 8, 0}},  // u64
    {2, 23, 3, {8, 9, 6, 0}},  // u64
    {2, 24, 3, {8, 8, 8, 0}},  // u64
    {2, 25, 3, {9, 8, 8, 0}},  // u64
    {2, 26, 3, {9, 8, 9, 0}},  // u64
};

inline constexpr TunedBatchSplit TUNED_BATCH_SPLITS[] = {
    {0, 10, 4, 0x4222u},  // u32 [2, 2, 2, 4]
    {0, 11, 4, 0x3332u},  // u32 [2, 3, 3, 3]
    {0, 12, 3, 0x444u},  // u32 [4, 4, 4]
    {0, 13, 4, 0x4333u},  // u32 [3, 3, 3, 4]
    {0, 14, 4, 0x3344u},  // u32 [4, 4, 3, 3]
    {0, 15, 4, 0x3444u},  // u32 [4, 4, 4, 3]
    {0, 16, 6, 0x433222u},  // u32 [2, 2, 2, 3, 3, 4]
    {0, 17, 6, 0x333332u},  // u32 [2, 3, 3, 3, 3, 3]
    {0, 18, 6, 0x222444u},  // u32 [4, 4, 4, 2, 2, 2]
    {1, 10, 4, 0x4222u},  // goldilocks [2, 2, 2, 4]
    {1, 11, 4, 0x2333u},  // goldilocks [3, 3, 3, 2]
    {1, 12, 3, 0x444u},  // goldilocks [4, 4, 4]
    {1, 13, 4, 0x4432u},  // goldilocks [2, 3, 4, 4]
    {1, 14, 6, 0x222224u},  // goldilocks [4, 2, 2, 2, 2, 2]
    {1, 15, 6, 0x432222u},  // goldilocks [2, 2, 2, 2, 3, 4]
    {1, 16, 6, 0x223333u},  // goldilocks [3, 3, 3, 3, 2, 2]
    {1, 17, 6, 0x433322u},  // goldilocks [2, 2, 3, 3, 3, 4]
    {1, 18, 6, 0x222444u},  // goldilocks [4, 4, 4, 2, 2, 2]
};

}  // namespace stntt
