// Pass plans of the circle NTT: how the log_n layers of a transform are split
// into LDS-tiled passes. Pure host logic (no device code), shared by the NTT
// driver and its autotuner.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <vector>

namespace vx {

// Pass shapes. The contiguous pass (c = 0) holds up to 2^14 words in LDS;
// strided passes move 2^c contiguous words per row (c = 4: 64-byte, c = 5:
// 128-byte, c = 6: 256-byte segments) of up to 2^14-word tiles. Which split a
// size uses comes from the plan table: the default below, or what the
// autotuner measured.
constexpr uint32_t NTT_LOW_MAX = 14;   // most layers of the contiguous pass
constexpr uint32_t NTT_HIGH_MAX6 = 8;  // most layers of a strided pass, c = 6
constexpr uint32_t NTT_HIGH_MAX5 = 9;  // most layers of a strided pass, c = 5
constexpr uint32_t NTT_HIGH_MAX4 = 10; // most layers of a strided pass, c = 4
constexpr uint32_t NTT_LOW = 12;       // default split: contiguous pass
constexpr uint32_t NTT_HIGH = 7;       // default split: most layers per strided pass
constexpr uint32_t NTT_COLS = 5;       // default split: strided row segment

struct Pass {
    uint32_t s, m, c;
    // 1: the top pass of an evaluation whose input fills only the first quarter
    // (a blowup-4 LDE). Its two top layers only copy that quarter into all four,
    // so the tile holds 2^(m-2) rows of the quarter and is transformed once per
    // quarter: m layers for the LDS of m - 2. One workgroup does the four
    // quarters in turn (one load from HBM, three from its L2), so the input can
    // be overwritten in place: its own quarter 0 is stored last.
    uint32_t r = 0;
};
inline uint32_t pass_rows(const Pass& p) { return p.r ? p.m - 2 : p.m; }

inline std::vector<Pass> default_plan(uint32_t k) {
    std::vector<Pass> ps;
    if (k == 0) return ps;
    uint32_t m0 = k < NTT_LOW ? k : NTT_LOW;
    ps.push_back({0, m0, 0});
    uint32_t R = k - m0;
    if (R) {
        uint32_t np = (R + NTT_HIGH - 1) / NTT_HIGH, s = m0;
        for (uint32_t i = 0; i < np; i++) {
            uint32_t mi = R / np + (i < R % np ? 1 : 0);
            ps.push_back({s, mi, NTT_COLS});
            s += mi;
        }
    }
    return ps;
}

// A plan is valid when its passes cover layers 0 .. k-1 in order, the first is
// contiguous, every strided pass starts at a stride of at least its segment
// and every shape has a compiled kernel. Replicating top passes are valid only
// in plans for quarter-filled evaluations (lde).
inline bool plan_valid(uint32_t k, const std::vector<Pass>& ps, bool lde = false) {
    if (k == 0) return ps.empty();
    if (ps.empty() || ps.size() > 8 || ps[0].c != 0 || ps[0].s != 0) return false;
    uint32_t s = 0;
    for (const Pass& p : ps) {
        if (p.s != s || p.m == 0 || p.r > 1) return false;
        if (p.r && (!lde || p.m < 3 || p.s + p.m != k)) return false;
        const uint32_t rows = pass_rows(p);
        if (p.c == 0 && (p.s != 0 || rows > NTT_LOW_MAX)) return false;
        if (p.c == 6 && (rows > NTT_HIGH_MAX6 || p.s < 6)) return false;
        if (p.c == 5 && (rows > NTT_HIGH_MAX5 || p.s < 5)) return false;
        if (p.c == 4 && (rows > NTT_HIGH_MAX4 || p.s < 4)) return false;
        if (p.c != 0 && p.c != 4 && p.c != 5 && p.c != 6) return false;
        s += p.m;
    }
    return s == k;
}

// Candidate splits of a size-2^k transform for the autotuner: every contiguous
// pass depth that leaves at most three strided passes, with 64-, 128- or
// 256-byte row segments and the fewest (or one more) balanced strided passes.
inline std::vector<std::vector<Pass>> plan_candidates(uint32_t k) {
    std::vector<std::vector<Pass>> out;
    auto add = [&](const std::vector<Pass>& ps) {
        if (!plan_valid(k, ps)) return;
        for (const auto& q : out)
            if (q.size() == ps.size() &&
                std::equal(q.begin(), q.end(), ps.begin(),
                           [](const Pass& x, const Pass& y) { return x.m == y.m && x.c == y.c; }))
                return;
        out.push_back(ps);
    };
    const uint32_t lo = std::max<uint32_t>(k >= 16 ? 8 : 1, k > 3 * NTT_HIGH_MAX4 ? k - 3 * NTT_HIGH_MAX4 : 1);
    for (uint32_t m0 = lo; m0 <= std::min(k, NTT_LOW_MAX); m0++) {
        const uint32_t R = k - m0;
        if (R == 0) {
            add({{0, m0, 0}});
            continue;
        }
        for (uint32_t c : {4u, 5u, 6u}) {
            const uint32_t hmax = c == 4 ? NTT_HIGH_MAX4 : c == 5 ? NTT_HIGH_MAX5 : NTT_HIGH_MAX6;
            const uint32_t npmin = (R + hmax - 1) / hmax;
            for (uint32_t np = npmin; np <= npmin + 1 && np <= 7 && np <= R; np++) {
                std::vector<Pass> ps = {{0, m0, 0}};
                uint32_t s0 = m0;
                for (uint32_t i = 0; i < np; i++) {
                    const uint32_t mi = R / np + (i < R % np ? 1 : 0);
                    ps.push_back({s0, mi, c});
                    s0 += mi;
                }
                add(ps);
            }
        }
    }
    add(default_plan(k));
    return out;
}

// Candidates for a blowup-4 evaluation (input in the first quarter): the
// general ones, whose top pass skips its two copy-only layers inside a full
// tile, and the same splits with a replicating top pass, which carries two
// layers more for its LDS.
inline std::vector<std::vector<Pass>> lde_plan_candidates(uint32_t k) {
    std::vector<std::vector<Pass>> out = plan_candidates(k);
    auto add = [&](const std::vector<Pass>& ps) {
        if (plan_valid(k, ps, true)) out.push_back(ps);
    };
    if (k < 3) return out;
    if (k - 2 <= NTT_LOW_MAX) add({{0, k, 0, 1}});
    const uint32_t lo = std::max<uint32_t>(k >= 16 ? 8 : 1, k > 3 * NTT_HIGH_MAX4 + 2 ? k - 3 * NTT_HIGH_MAX4 - 2 : 1);
    for (uint32_t m0 = lo; m0 <= std::min(k, NTT_LOW_MAX); m0++) {
        if (k - m0 < 3) continue;
        const uint32_t R = k - m0 - 2;  // layers to transform above the contiguous pass
        for (uint32_t c : {4u, 5u, 6u}) {
            const uint32_t hmax = c == 4 ? NTT_HIGH_MAX4 : c == 5 ? NTT_HIGH_MAX5 : NTT_HIGH_MAX6;
            const uint32_t npmin = std::max<uint32_t>(1, (R + hmax - 1) / hmax);
            for (uint32_t np = npmin; np <= npmin + 1 && np <= 7 && np <= R; np++) {
                std::vector<Pass> ps = {{0, m0, 0}};
                uint32_t s0 = m0;
                for (uint32_t i = 0; i < np; i++) {
                    const uint32_t mi = R / np + (i < R % np ? 1 : 0) + (i + 1 == np ? 2 : 0);
                    ps.push_back({s0, mi, c, i + 1 == np ? 1u : 0u});
                    s0 += mi;
                }
                add(ps);
            }
        }
    }
    return out;
}

}  // namespace vx
