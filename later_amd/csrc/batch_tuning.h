// Variant of the one-workgroup batched kernel per size bucket (see
// bucket_of): 0 = 16-column panels / 64-column blocks of earlier Q columns, 1 = 32 / 64,
// 2 = 16 / 128, 3 = 32 / 128.  Written by later_amd/autotune.py from
// timings on AMD Radeon Graphics (gfx950); rows are n classes (n <= 32, <= 64, ..., > 224),
// columns m classes (m < 128, < 256, ..., >= 4096).
#pragma once
constexpr int kDefaultTuning[kBuckets] = {
    3, 1, 3, 1, 1, 1, 3,
    1, 1, 1, 1, 1, 1, 1,
    1, 1, 1, 1, 1, 1, 1,
    1, 1, 1, 1, 1, 1, 1,
    3, 1, 1, 1, 1, 1, 1,
    3, 1, 1, 1, 1, 1, 1,
    3, 1, 1, 1, 1, 1, 1,
    3, 1, 1, 1, 1, 1, 1,
};
