// Runs the oracle's main paths under AddressSanitizer and UndefinedBehavior-
// Sanitizer (built with g++, no GPU): proofs at small sizes through prove,
// serialize, deserialize and verify, a tampered proof, the memory-lean
// summary against the full proof, the batch inverse with zeros, the stwo
// pieces and the wide commitment, the DEEP quotient pieces and a PoW search.
// Any invalid access or undefined operation aborts the run; the result checks
// exit non-zero on failure.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../oracle/oracle.h"

#define CHECK(c)                                                                        \
    do {                                                                                \
        if (!(c)) {                                                                     \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                               \
        }                                                                               \
    } while (0)

static const uint32_t P = 0x7FFFFFFFu;

static std::vector<uint32_t> words(uint64_t n, uint64_t seed) {
    std::vector<uint32_t> v(n);
    uint64_t st = seed;
    for (auto& x : v) x = (uint32_t)(orc_splitmix64(&st) % P);
    return v;
}

static void proofs() {
    for (uint32_t log_n : {2u, 4u, 7u, 10u}) {
        const uint32_t nq = log_n == 10 ? 100 : 80, last = log_n == 10 ? 4 : 3;
        if (last > log_n + 1) continue;
        const uint64_t len = orc_prove_fibonacci(3, 5, log_n, nq, last, nullptr, 0, nullptr);
        CHECK(len > 0);
        std::vector<uint8_t> pf(len);
        uint32_t stats[17];
        CHECK(orc_prove_fibonacci(3, 5, log_n, nq, last, pf.data(), len, stats) == len);
        CHECK(orc_verify_krak(pf.data(), len, nq, last) == 0);
        // a flipped byte in the last decommitment and a truncated proof are rejected
        std::vector<uint8_t> bad = pf;
        bad[len - 5] ^= 0x40;
        CHECK(orc_verify_krak(bad.data(), len, nq, last) != 0);
        CHECK(orc_verify_krak(pf.data(), len - 4, nq, last) != 0);
        // the lean summary gives the same commitments
        const uint32_t J = log_n + 2 - last;
        std::vector<uint32_t> roots(8 * (2 + J)), lastw(4u << last);
        std::vector<uint64_t> q(nq);
        CHECK(orc_prove_summary(3, 5, log_n, nq, last, roots.data(), lastw.data(), q.data()) == 0);
        CHECK(std::memcmp(roots.data(), stats, 64) == 0);
    }
}

static void batch_inverse_zeros() {
    std::vector<uint32_t> x = words(64 * 5 + 9, 11), out(x.size());
    for (uint64_t i : {0ull, 5ull, 63ull, 64ull, 130ull, 200ull, 64ull * 5 + 8}) x[i] = 0;
    orc_batch_inverse(x.data(), out.data(), x.size());
    for (uint64_t i = 0; i < x.size(); i++) {
        const bool cleared = x[i] == 0 || x[i - i % 64] == 0;
        CHECK(cleared ? out[i] == 0 : (uint64_t)out[i] * x[i] % P == 1);
    }
}

static void stwo_and_wide() {
    for (uint32_t k : {2u, 5u, 9u}) {
        const uint64_t n = 1ull << k;
        std::vector<uint32_t> tw(n / 2), itw(n / 2);
        orc_stwo_twiddles(k, tw.data(), itw.data());
        std::vector<uint32_t> d = words(n, k), d0 = d;
        orc_stwo_ntt_interpolate(d.data(), itw.data(), k);
        orc_stwo_ntt_evaluate(d.data(), tw.data(), k);
        CHECK(d == d0);
    }
    const uint32_t logs[5] = {4, 4, 6, 7, 7};
    std::vector<std::vector<uint32_t>> cols;
    std::vector<const uint32_t*> ptrs;
    for (int c = 0; c < 5; c++) cols.push_back(words(1ull << logs[c], 100 + c));
    for (auto& c : cols) ptrs.push_back(c.data());
    uint32_t root[8], beta[4];
    std::vector<uint32_t> fri(8 * 16), lastw(4 << 2);
    CHECK(orc_wide_commit(ptrs.data(), logs, 5, 2, 2, root, fri.data(), lastw.data(), beta) > 0);
    std::vector<uint32_t> leaves(8ull << 9);
    orc_lifted_leaves(ptrs.data(), logs, 5, 9, leaves.data());
}

static void deep_and_pow() {
    const uint32_t L = 6, n = 1u << L;
    std::vector<uint32_t> c0 = words(n, 1), c1 = words(n / 4, 2);
    const uint32_t* cs[2] = {c0.data(), c1.data()};
    const uint32_t idx[1] = {0};
    std::vector<uint32_t> b = words(4, 3), c = words(4, 4);
    std::vector<uint32_t> o[4];
    for (auto& v : o) v.assign(n, 0);
    orc_accumulate_numerators(cs, idx, b.data(), c.data(), 1, n, o[0].data(), o[1].data(), o[2].data(), o[3].data());
    std::vector<uint32_t> xs(n), ys(n);
    orc_circle_domain_brev(L, xs.data(), ys.data());
    std::vector<uint32_t> col = words(4 * n, 5), cur = words(n, 6);
    orc_accumulate_lift(col.data(), cur.data(), n, 2);
    const uint32_t state[8] = {1, 2, 3, 4, 5, 6, 7, 8};
    uint32_t pre[8];
    orc_pow_prefix(state, 8, 0, pre);
    const uint64_t nonce = orc_pow_search(pre, 8, 0, 0, 1ull << 16);
    CHECK(nonce != UINT64_MAX && orc_pow_zeros(pre, nonce, 0) >= 8);
}

int main() {
    orc_set_threads(2);
    proofs();
    batch_inverse_zeros();
    stwo_and_wide();
    deep_and_pow();
    std::printf("oracle under sanitizers: ok\n");
    return 0;
}
