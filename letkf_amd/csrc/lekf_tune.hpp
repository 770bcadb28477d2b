// lekf_tune.hpp -- launch parameters chosen at run time (set by the autotuner
// through lekf_tune_set; process-wide, read when a launch is issued).
#pragma once
#include <cstddef>

struct LekfTune {
    int gram_waves = 2;                 // waves per point of the k > 32 Gram kernel: 2, 4, 8
    size_t gram_ws_bytes = 16ull << 30; // Gram workspace per launch (points per sub-batch)
    int sel_per_cu = 3;                 // selection workgroups (scratch slots) per CU: 1..4
    int eig_split = 1;                  // 32 < k <= 64: three-kernel eigensolver (1) or fused (0)
    long split_rcap = 0;                // rotations a split stream holds before its point is
                                        // solved in full by the fallback (0: 2k^2 + 8k)
    long mixed_kappa = 500;             // mixed forms, split eigensolver: condition number of
                                        // C above which a point is redone in FP64 (0: never)
};

extern LekfTune g_lekf_tune;
