// Host-side driver of the Merkle engine: commits a leaf source with one stage
// kernel over the leaves, stage kernels over node layers while more than
// PROVER_TAIL_ITEMS nodes remain, and the tail kernel (root + transcript).
#pragma once
#include "common.hpp"
#include "merkle_engine.hpp"

namespace vx {

// Trees of at most this many items go straight to the tail kernel; larger
// ones are reduced by stage kernels until this many nodes remain (one more
// tail level is cheaper than one more stage launch).
constexpr uint32_t PROVER_TAIL_ITEMS = MERKLE_TAIL_MAX;

struct Scratch {
    uint32_t* a;
    uint32_t* b;
};

// Commits n_leaves (power of two >= 2) items of a leaf source, for `batch`
// proofs at once whose buffers are bs words apart (bs = 0 for one proof).
// Trees of at most direct_max leaves go straight to the tail kernel.
template <class Src, class Mark>
void commit_tree(const Src& src, uint64_t n_leaves, LevelStore st, TailOut out, Scratch sc, hipStream_t s,
                 Mark&& mark, uint32_t batch = 1, uint64_t bs = 0, uint64_t direct_max = PROVER_TAIL_ITEMS) {
    if (n_leaves <= direct_max) {
        merkle_tail_kernel<Src><<<dim3(1, batch), TAIL_THREADS, 0, s>>>(src, 0, (uint32_t)n_leaves, st, out, bs);
        check_launch("merkle_tail_kernel");
        mark("tail");
        return;
    }
    uint32_t logc = pick_logc(n_leaves, StageLogcMax<Src>::value, batch, stage_uses_plan<Src>());
    uint64_t waves = n_leaves >> (7 + logc);
    // deep leaf stages hand on MERKLE_STAGE_KEEP nodes a wave (at most n / 128
    // in all, the scratch bound of the shallowest stage); the others one
    const uint32_t keep = logc >= MERKLE_KEEP_LOGC ? MERKLE_STAGE_KEEP : 1u;
    merkle_stage_kernel<Src><<<dim3(ceil_div(waves, MERKLE_STAGE_WAVES), batch), 64 * MERKLE_STAGE_WAVES, 0, s>>>(
        src, 0, n_leaves, logc, sc.a, st, bs, keep);
    check_launch("merkle_stage_kernel");
    mark("stage_leaf");
    uint64_t n = waves * keep;
    uint32_t lvl = 7 + logc - (keep == 1 ? 0u : 4u);  // 16 = 64 >> 2 nodes: two of the six levels done
    uint32_t* cur = sc.a;
    uint32_t* nxt = sc.b;
    while (n > PROVER_TAIL_ITEMS) {
        logc = pick_logc(n, MERKLE_LOGC_MAX, batch);
        waves = n >> (7 + logc);
        merkle_stage_kernel<SrcNodes><<<dim3(ceil_div(waves, MERKLE_STAGE_WAVES), batch), 64 * MERKLE_STAGE_WAVES,
                                        0, s>>>(SrcNodes{cur}, lvl, n, logc, nxt, st, bs, 1u);
        check_launch("merkle_stage_kernel<nodes>");
        mark("stage_node");
        n = waves;
        lvl += 7 + logc;
        uint32_t* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    merkle_tail_kernel<SrcNodes><<<dim3(1, batch), TAIL_THREADS, 0, s>>>(SrcNodes{cur}, lvl, (uint32_t)n, st, out, bs);
    check_launch("merkle_tail_kernel<nodes>");
    mark("tail");
}

// Line-fold layers of at most LINE_QUAD_MAX values (and more than the FRI tail
// kernel takes) of a single proof commit through line_fold_quad_stage_kernel:
// 16 leaves a wave up to 2^13 values, 32 at 2^14, so the tail starts from at
// most 512 nodes. Such a layer's chain is latency-bound for one proof; a batch
// of proofs fills the device with its waves, where the lane-per-pair stage
// does the same hashing in fewer instructions (about 980 a compression
// against 4 x 410), so batches keep it.
constexpr uint64_t LINE_QUAD_MAX = 1ull << 14;
inline uint32_t line_quad_leaves(uint64_t n) { return n <= (1ull << 13) ? 16u : 32u; }
// Stage output (nodes per proof) of the quad line-fold stages of the layers
// E/4, E/8, ... that take it: 0 if none does.
inline uint64_t max_line_quad_nodes(uint64_t e) {
    uint64_t mx = 0;
    for (uint64_t n = e / 4; n >= 2; n /= 2)
        if (n <= LINE_QUAD_MAX && n >= 2048) mx = std::max<uint64_t>(mx, n / line_quad_leaves(n));
    return mx;
}

// Commits a line-fold layer of n values (2048 <= n <= LINE_QUAD_MAX, a power
// of two) with the quad leaf stage and the tail kernel.
template <class Mark>
void commit_line_fold_quad(const SrcLineFold& src, uint64_t n, LevelStore st, TailOut out, Scratch sc, hipStream_t s,
                           Mark&& mark, uint32_t batch = 1, uint64_t bs = 0) {
    const uint32_t lpw = line_quad_leaves(n);
    const uint64_t waves = n / lpw;
    const dim3 grid((uint32_t)ceil_div(waves, LINE_QUAD_WAVES), batch);
    if (lpw == 16)
        line_fold_quad_stage_kernel<1><<<grid, 64 * LINE_QUAD_WAVES, 0, s>>>(src, n, sc.a, st, bs);
    else
        line_fold_quad_stage_kernel<2><<<grid, 64 * LINE_QUAD_WAVES, 0, s>>>(src, n, sc.a, st, bs);
    check_launch("line_fold_quad_stage_kernel");
    mark("stage_leaf");
    const uint32_t lvl = lpw == 16 ? 4u : 5u;
    merkle_tail_kernel<SrcNodes><<<dim3(1, batch), TAIL_THREADS, 0, s>>>(SrcNodes{sc.a}, lvl, (uint32_t)waves, st, out,
                                                                         bs);
    check_launch("merkle_tail_kernel<nodes>");
    mark("tail");
}

// Largest stage output of a tree of n_leaves items (per proof), for sizing
// the stage scratch buffers.
inline uint64_t max_stage_nodes(uint64_t n_leaves) { return max_stage_nodes(n_leaves, PROVER_TAIL_ITEMS); }

}  // namespace vx
