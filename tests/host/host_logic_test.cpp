// Host-logic unit test of libvortex_hip (built with g++, no GPU): the NTT
// pass plans (validity, default split, autotuner candidates) and the result
// ring allocator of the proof server. Exits non-zero on the first failure.
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <random>
#include <set>
#include <vector>

#include "../../vortexstark_amd/csrc/ntt_plan.hpp"
#include "../../vortexstark_amd/csrc/result_ring.hpp"
#include "../../vortexstark_amd/csrc/stage_plan.hpp"

#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                \
        }                                                                \
    } while (0)

using namespace vx;

static std::vector<Pass> make(std::initializer_list<std::pair<uint32_t, uint32_t>> mc) {
    std::vector<Pass> ps;
    uint32_t s = 0;
    for (auto [m, c] : mc) {
        ps.push_back({s, m, c});
        s += m;
    }
    return ps;
}

static void test_plans() {
    // every default split and every autotuner candidate is valid and covers k layers
    for (uint32_t k = 1; k <= 30; k++) {
        CHECK(plan_valid(k, default_plan(k)));
        const auto cands = plan_candidates(k);
        CHECK(!cands.empty());
        bool has_default = false;
        for (const auto& ps : cands) {
            CHECK(plan_valid(k, ps));
            uint32_t sum = 0;
            for (const auto& p : ps) sum += p.m;
            CHECK(sum == k);
            if (ps.size() == default_plan(k).size()) {
                bool same = true;
                for (size_t i = 0; i < ps.size(); i++)
                    same &= ps[i].m == default_plan(k)[i].m && ps[i].c == default_plan(k)[i].c;
                has_default |= same;
            }
        }
        CHECK(has_default);
    }
    CHECK(plan_valid(0, {}));
    // the default split of the plain workload's sizes
    CHECK(default_plan(26).size() == 3 && default_plan(26)[0].m == 12 && default_plan(26)[1].m == 7);
    // shapes without a compiled kernel, gaps, wrong first pass
    CHECK(!plan_valid(16, make({{15, 0}, {1, 5}})));   // contiguous pass too deep
    CHECK(!plan_valid(16, make({{10, 0}, {5, 4}})));   // does not cover 16 layers
    CHECK(!plan_valid(16, make({{3, 0}, {13, 5}})));   // strided pass too deep
    CHECK(!plan_valid(16, make({{4, 0}, {12, 4}})));   // strided pass too deep (c = 4)
    CHECK(!plan_valid(16, make({{10, 4}, {6, 5}})));   // first pass not contiguous
    CHECK(!plan_valid(16, make({{3, 0}, {7, 4}, {6, 5}})));  // c = 4 pass starting at stride 2^3
    CHECK(plan_valid(16, make({{4, 0}, {6, 4}, {6, 5}})));
    CHECK(!plan_valid(16, make({{2, 0}, {7, 4}, {7, 5}})));  // c = 4 pass starting at stride 2^2
    CHECK(!plan_valid(16, make({{8, 0}, {8, 3}})));          // no kernel for 8-word segments
    CHECK(plan_valid(16, make({{8, 0}, {8, 6}})));           // 256-byte segments, 8 layers
    CHECK(!plan_valid(16, make({{7, 0}, {9, 6}})));          // strided pass too deep (c = 6)
    CHECK(!plan_valid(16, make({{5, 0}, {5, 6}, {6, 5}})));  // c = 6 pass starting at stride 2^5
    CHECK(!plan_valid(16, make({{8, 0}, {0, 5}, {8, 5}})));  // empty pass

    // blowup-4 evaluation plans: a replicating top pass holds m - 2 rows
    auto repl_top = [](std::vector<Pass> ps) {
        ps.back().r = 1;
        return ps;
    };
    const auto p22 = repl_top(make({{12, 0}, {10, 5}}));  // 8 tile rows, c = 5: two passes for 2^22
    CHECK(!plan_valid(22, p22));                        // only in an LDE plan
    CHECK(plan_valid(22, p22, true));
    CHECK(!plan_valid(22, make({{12, 0}, {10, 5}}), true));  // 10 rows at c = 5 without replication
    CHECK(plan_valid(26, repl_top(make({{14, 0}, {12, 4}})), true));
    CHECK(!plan_valid(26, repl_top(make({{14, 0}, {12, 5}})), true));  // 10 tile rows at c = 5
    CHECK(plan_valid(16, repl_top(make({{16, 0}})), true));            // one contiguous pass of 2^14 rows
    CHECK(!plan_valid(17, repl_top(make({{17, 0}})), true));           // 2^15 rows
    CHECK(!plan_valid(16, repl_top(make({{14, 0}, {2, 5}})), true));   // no tile rows left
    {
        auto ps = make({{8, 0}, {4, 5}, {4, 5}});
        ps[1].r = 1;  // replicating pass below the top
        CHECK(!plan_valid(16, ps, true));
    }
    for (uint32_t k = 3; k <= 30; k++) {
        const auto cands = lde_plan_candidates(k);
        bool has_repl = false;
        for (const auto& ps : cands) {
            CHECK(plan_valid(k, ps, true));
            uint32_t sum = 0;
            for (size_t i = 0; i < ps.size(); i++) {
                sum += ps[i].m;
                CHECK(ps[i].r == 0 || i + 1 == ps.size());
            }
            CHECK(sum == k);
            has_repl |= ps.back().r == 1;
        }
        CHECK(has_repl);
        CHECK(cands.size() > plan_candidates(k).size());
    }
    // the two-pass LDE plans of the provers' largest sizes are among the candidates
    for (auto [k, want] : {std::pair<uint32_t, std::vector<Pass>>{22, p22},
                           {26, repl_top(make({{14, 0}, {12, 4}}))}}) {
        bool found = false;
        for (const auto& ps : lde_plan_candidates(k))
            found |= ps.size() == want.size() &&
                     std::equal(ps.begin(), ps.end(), want.begin(), [](const Pass& x, const Pass& y) {
                         return x.s == y.s && x.m == y.m && x.c == y.c && x.r == y.r;
                     });
        CHECK(found);
    }
}

static void test_ring() {
    ResultRing r;
    std::vector<uint8_t> mem(100);
    r.base = mem.data();
    r.size = 100;
    CHECK(r.alloc(0, 1) == -1);
    CHECK(r.alloc(101, 1) == -1);
    const int64_t a = r.alloc(60, 2);  // [0, 60)
    CHECK(a == 0 && r.at(a) == mem.data());
    const int64_t b = r.alloc(30, 1);  // [60, 90)
    CHECK(b == 1 && r.at(b) == mem.data() + 60);
    CHECK(r.alloc(20, 1) == -1);       // 10 left at the end, nothing at the front
    r.release(a);                       // one of two proofs released: still live
    CHECK(r.alloc(20, 1) == -1);
    r.release(a);                       // [0, 60) free
    const int64_t c = r.alloc(50, 1);  // wraps to [0, 50)
    CHECK(c == 2 && r.at(c) == mem.data());
    const int64_t d = r.alloc(10, 1);  // [50, 60)
    CHECK(d == 3 && r.at(d) == mem.data() + 50);
    CHECK(r.alloc(1, 1) == -1);        // full up to the oldest live region
    r.release(c);                       // not the oldest: space stays held
    CHECK(r.alloc(1, 1) == -1);
    r.release(b);                       // oldest released: b and c pop, d remains
    CHECK(r.regions.size() == 1);
    const int64_t e = r.alloc(40, 1);  // after d: [60, 100)
    CHECK(e == 4 && r.at(e) == mem.data() + 60);
    r.release(7);                       // unknown ids are ignored
    r.release(d);
    r.release(e);
    CHECK(r.regions.empty());

    // random traffic: live regions never overlap and stay inside the buffer
    std::mt19937_64 rng(7);
    ResultRing q;
    std::vector<uint8_t> big(1 << 16);
    q.base = big.data();
    q.size = big.size();
    std::vector<std::pair<int64_t, uint32_t>> live;  // id, remaining releases
    for (int step = 0; step < 20000; step++) {
        if (live.empty() || rng() % 3) {
            const uint64_t bytes = 1 + rng() % 9000;
            const uint32_t n = 1 + rng() % 3;
            const int64_t id = q.alloc(bytes, n);
            if (id >= 0) live.push_back({id, n});
        } else {
            const size_t i = rng() % live.size();
            q.release(live[i].first);
            if (--live[i].second == 0) live.erase(live.begin() + i);
        }
        // check the regions still held
        std::vector<std::pair<uint64_t, uint64_t>> spans;
        for (const auto& reg : q.regions) {
            CHECK(reg.off + reg.bytes <= q.size);
            spans.push_back({reg.off, reg.off + reg.bytes});
        }
        std::sort(spans.begin(), spans.end());
        for (size_t i = 1; i < spans.size(); i++) CHECK(spans[i - 1].second <= spans[i].first);
    }
}

static void test_stage_plan() {
    for (uint32_t i = 0; i < STAGE_PLAN_LOGS; i++) stage_plan_table()[i].store(0);
    // default rule: the largest depth leaving STAGE_MIN_WAVES waves
    CHECK(pick_logc(1ull << 26) == 5);  // 16384 waves
    CHECK(pick_logc(1ull << 23) == 4);  // 4096 waves (depth 5 leaves 2048)
    CHECK(pick_logc(1ull << 19) == 0);
    CHECK(pick_logc(1ull << 26, 4) == 4);
    CHECK(pick_logc(1ull << 16, 5, 256) == 5);  // 2^24 items in the launch, 16 waves per proof
    CHECK(pick_logc(1ull << 12, 5, 1024) == 3);  // two waves per proof at most depth 3
    for (uint32_t l = 8; l <= 30; l++)
        for (uint32_t batch : {1u, 3u, 64u, 1024u})
            for (uint32_t mx = 0; mx <= MERKLE_LOGC_MAX; mx++) {
                const uint64_t n = 1ull << l;
                const uint32_t c = pick_logc(n, mx, batch);
                CHECK(c <= mx);
                CHECK(c == 0 || (n >> (7 + c)) >= 2);
                CHECK((n >> (7 + c)) <= max_stage_nodes(n, 2048) || n <= 2048);
            }
    // tuned depths are used, clipped to the source's maximum and to two waves per proof
    stage_plan_table()[20].store(2 + 1);
    CHECK(pick_logc(1ull << 20) == 2);
    CHECK(pick_logc(1ull << 14, 5, 64) == 2);  // 2^20 items in the launch
    CHECK(pick_logc(1ull << 20, 1) == 1);
    // memory-heavy sources ignore the table: the default rule's depth 2 (4096 waves)
    stage_plan_table()[21].store(5 + 1);
    CHECK(pick_logc(1ull << 21) == 5);
    CHECK(pick_logc(1ull << 21, 4, 1, false) == 2);
    stage_plan_table()[9].store(5 + 1);
    CHECK(pick_logc(1ull << 9) == 1);  // 2^9 items: depth 1 is the deepest with two waves
    for (uint32_t i = 0; i < STAGE_PLAN_LOGS; i++) stage_plan_table()[i].store(0);
    CHECK(max_stage_nodes(2048, 2048) == 0 && max_stage_nodes(4096, 2048) == 32);
}

int main() {
    test_stage_plan();
    test_plans();
    test_ring();
    std::printf("host logic: ok\n");
    return 0;
}
