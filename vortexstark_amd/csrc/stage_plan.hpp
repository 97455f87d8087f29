// Chunk depth of the Merkle stage kernels: a wave hashes 2^c chunks of 128
// items, merging them as it goes, and ends with one 64 -> 1 reduction on
// narrow levels, so its output is one node 7 + c levels above its items.
// Pure host logic (no device code), shared by the tree driver and its
// autotuner.
#pragma once
#include <stdint.h>

#include <atomic>

namespace vx {

constexpr uint32_t MERKLE_LOGC_MAX = 5;       // chunks per wave in a stage: up to 32
constexpr uint32_t STAGE_MIN_WAVES = 3072;    // default rule: three waves per SIMD
constexpr uint32_t STAGE_PLAN_LOGS = 40;      // table index: log2 of the items of one launch

// Tuned depth per log2 of the items one stage launch covers (every proof of a
// batch), stored as c + 1 (0: not tuned, use the default rule).
inline std::atomic<uint8_t>* stage_plan_table() {
    static std::atomic<uint8_t> t[STAGE_PLAN_LOGS];
    return t;
}

inline uint32_t floor_log2(uint64_t x) { return x ? 63u - (uint32_t)__builtin_clzll(x) : 0u; }

// Depth of a stage over n items per proof for `batch` proofs: the tuned depth
// for n * batch items when there is one and use_plan is set, else the largest
// c that still leaves STAGE_MIN_WAVES waves in the launch; never above
// logc_max, and always at least two waves per proof.
inline uint32_t pick_logc(uint64_t n, uint32_t logc_max = MERKLE_LOGC_MAX, uint32_t batch = 1, bool use_plan = true) {
    const uint32_t lt = floor_log2(n * batch);
    int c = use_plan && lt < STAGE_PLAN_LOGS ? (int)stage_plan_table()[lt].load(std::memory_order_relaxed) - 1 : -1;
    if (c < 0) {
        c = 0;
        for (int k = (int)logc_max; k > 0; k--) {
            const uint64_t w = n >> (7 + k);
            if (w >= 2 && w * batch >= STAGE_MIN_WAVES) {
                c = k;
                break;
            }
        }
    }
    if (c > (int)logc_max) c = (int)logc_max;
    while (c > 0 && (n >> (7 + c)) < 2) c--;
    return (uint32_t)c;
}

// Most nodes any stage of a tree of n_leaves items can output (per proof),
// whatever depth is picked: a wave covers at least 128 items.
inline uint64_t max_stage_nodes(uint64_t n_leaves, uint64_t tail_items) {
    return n_leaves <= tail_items ? 0 : n_leaves >> 7;
}

}  // namespace vx
