// Oracle — CPU restatement of VortexSTARK's Fibonacci commit path.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// CPU-baseline leg of bench.py as the checker. Product code never uses it.
//
// Each function cites the reference file:line it restates (paths relative to
// the VortexSTARK tree). The code favours literal fidelity over speed; the only
// liberty taken is running independent loop iterations on several threads
// (parallel_for), which cannot change any result.
//
// Parity pins (where each function is checked against something other than
// itself) are listed in tests/test_oracle_*.py.

#include "oracle.h"

#include <array>
#include <atomic>
#include <chrono>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t P = 0x7FFFFFFFu;

// ---------------------------------------------------------------- threads
std::atomic<uint32_t> g_threads{0};

uint32_t n_threads() {
    uint32_t t = g_threads.load();
    if (t == 0) {
        t = std::thread::hardware_concurrency();
        if (t == 0) t = 1;
    }
    return t;
}

// Runs fn(begin, end) over [0, n) split into contiguous ranges.
void parallel_for(uint64_t n, const std::function<void(uint64_t, uint64_t)>& fn) {
    uint32_t t = n_threads();
    if (t <= 1 || n < 8192) {
        fn(0, n);
        return;
    }
    if (t > n) t = (uint32_t)n;
    std::vector<std::thread> ths;
    uint64_t chunk = (n + t - 1) / t;
    for (uint32_t i = 0; i < t; i++) {
        uint64_t b = i * chunk, e = b + chunk < n ? b + chunk : n;
        if (b >= e) break;
        ths.emplace_back([&fn, b, e] { fn(b, e); });
    }
    for (auto& th : ths) th.join();
}

// ---------------------------------------------------------------- M31
// src/field/m31.rs:24-29 (reduce), :86-121 (add/sub/mul/neg), :32-50 (inverse via pow)
inline uint32_t reduce(uint64_t v) {
    uint32_t lo = (uint32_t)(v & P);
    uint32_t hi = (uint32_t)(v >> 31);
    uint32_t r = lo + hi;
    return r >= P ? r - P : r;
}
inline uint32_t add(uint32_t a, uint32_t b) {
    uint32_t r = a + b;
    return r >= P ? r - P : r;
}
inline uint32_t sub(uint32_t a, uint32_t b) { return a >= b ? a - b : a + P - b; }
inline uint32_t mul(uint32_t a, uint32_t b) { return reduce((uint64_t)a * b); }
inline uint32_t neg(uint32_t a) { return a == 0 ? 0 : P - a; }
inline uint32_t dbl(uint32_t a) {  // m31.rs:52-56
    uint32_t r = a << 1;
    return r >= P ? r - P : r;
}
uint32_t m31_pow(uint32_t base, uint32_t exp) {
    uint32_t r = 1;
    while (exp > 0) {
        if (exp & 1) r = mul(r, base);
        base = mul(base, base);
        exp >>= 1;
    }
    return r;
}
inline uint32_t inv(uint32_t a) { return m31_pow(a, P - 2); }

// ---------------------------------------------------------------- CM31 / QM31
// src/field/cm31.rs:78-87 (mul), :37-51 (norm/inverse)
struct CM31 {
    uint32_t a, b;
};
inline CM31 cadd(CM31 x, CM31 y) { return {add(x.a, y.a), add(x.b, y.b)}; }
inline CM31 csub(CM31 x, CM31 y) { return {sub(x.a, y.a), sub(x.b, y.b)}; }
inline CM31 cmul(CM31 x, CM31 y) {
    return {sub(mul(x.a, y.a), mul(x.b, y.b)), add(mul(x.a, y.b), mul(x.b, y.a))};
}
inline CM31 cmul_m31(CM31 x, uint32_t s) { return {mul(x.a, s), mul(x.b, s)}; }
inline CM31 cneg(CM31 x) { return {neg(x.a), neg(x.b)}; }
CM31 cinv(CM31 x) {
    uint32_t n = add(mul(x.a, x.a), mul(x.b, x.b));
    uint32_t in = inv(n);
    return {mul(x.a, in), mul(neg(x.b), in)};
}

// src/field/qm31.rs:11-14 (layout a + b*u), :114-132 (mul), :62-88 (norm/inverse)
struct QM31 {
    CM31 a, b;
};
inline QM31 qfrom(const uint32_t v[4]) { return {{v[0], v[1]}, {v[2], v[3]}}; }
inline void qto(QM31 q, uint32_t v[4]) {
    v[0] = q.a.a;
    v[1] = q.a.b;
    v[2] = q.b.a;
    v[3] = q.b.b;
}
inline QM31 qadd(QM31 x, QM31 y) { return {cadd(x.a, y.a), cadd(x.b, y.b)}; }
inline QM31 qsub(QM31 x, QM31 y) { return {csub(x.a, y.a), csub(x.b, y.b)}; }
inline QM31 qmul(QM31 x, QM31 y) {
    CM31 ac = cmul(x.a, y.a);
    CM31 bd = cmul(x.b, y.b);
    CM31 ad = cmul(x.a, y.b);
    CM31 bc = cmul(x.b, y.a);
    // bd * (2 + i) = (2p - q) + (p + 2q) i   (qm31.rs:122-125)
    CM31 bdu = {sub(dbl(bd.a), bd.b), add(bd.a, dbl(bd.b))};
    return {cadd(ac, bdu), cadd(ad, bc)};
}
inline QM31 qmul_m31(QM31 x, uint32_t s) { return {cmul_m31(x.a, s), cmul_m31(x.b, s)}; }
QM31 qinv(QM31 x) {
    CM31 a2 = cmul(x.a, x.a), b2 = cmul(x.b, x.b);
    CM31 b2u = {sub(dbl(b2.a), b2.b), add(b2.a, dbl(b2.b))};
    CM31 n = csub(a2, b2u);
    CM31 in = cinv(n);
    QM31 conj = {x.a, cneg(x.b)};
    return {cmul(conj.a, in), cmul(conj.b, in)};
}

// ---------------------------------------------------------------- circle
// src/circle.rs:17-91
struct Pt {
    uint32_t x, y;
};
const Pt IDENTITY = {1, 0};
const Pt GENERATOR = {2, 1268011823u};  // circle.rs:26-29
inline Pt pmul(Pt p, Pt q) { return {sub(mul(p.x, q.x), mul(p.y, q.y)), add(mul(p.x, q.y), mul(p.y, q.x))}; }
inline Pt pdouble(Pt p) { return pmul(p, p); }
Pt repeated_double(Pt p, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) p = pdouble(p);
    return p;
}
Pt mul_scalar(Pt p, uint32_t k) {  // circle.rs:66-77
    Pt r = IDENTITY, base = p;
    while (k > 0) {
        if (k & 1) r = pmul(r, base);
        base = pdouble(base);
        k >>= 1;
    }
    return r;
}

struct Coset {
    Pt initial, step;
    uint32_t log_size;
    // circle.rs:183-185: initial * step^(i as u32)
    Pt at(uint64_t i) const { return pmul(initial, mul_scalar(step, (uint32_t)i)); }
};
Coset subgroup(uint32_t k) { return {IDENTITY, repeated_double(GENERATOR, 31 - k), k}; }  // :100-108
Coset half_coset(uint32_t k) {                                                          // :112-121
    return {repeated_double(GENERATOR, 30 - k), repeated_double(GENERATOR, 31 - k), k};
}
Coset half_odds(uint32_t k) {  // :132-143 (k = 30 uses G^1)
    return {repeated_double(GENERATOR, k >= 29 ? 0 : 29 - k), repeated_double(GENERATOR, 31 - k), k};
}
Coset coset_of(int kind, uint32_t k) {
    if (kind == 0) return subgroup(k);
    if (kind == 1) return half_coset(k);
    return half_odds(k);
}
// All points in order, by sequential multiplication (circle.rs:262-272). The
// reference GPU kernel (twiddles.cu:80-105) evaluates initial*step^i by
// square-and-multiply; both give the same group element, hence the same coordinates.
void all_points(const Coset& c, std::vector<uint32_t>& xs, std::vector<uint32_t>& ys) {
    uint64_t n = 1ull << c.log_size;
    xs.resize(n);
    ys.resize(n);
    uint32_t nt = n_threads();
    uint64_t chunk = (n + nt - 1) / nt;
    parallel_for(nt, [&](uint64_t tb, uint64_t te) {
        for (uint64_t t = tb; t < te; t++) {
            uint64_t b = t * chunk, e = b + chunk < n ? b + chunk : n;
            if (b >= e) continue;
            Pt cur = c.at(b);
            for (uint64_t i = b; i < e; i++) {
                xs[i] = cur.x;
                ys[i] = cur.y;
                cur = pmul(cur, c.step);
            }
        }
    });
}

inline uint32_t bitrev(uint32_t x, uint32_t bits) {
    // __brev(x) >> (32 - bits)  (twiddles.cu:37)
    uint32_t r = 0;
    for (uint32_t i = 0; i < 32; i++) r |= ((x >> i) & 1u) << (31 - i);
    return bits == 0 ? 0 : r >> (32 - bits);
}

// ---------------------------------------------------------------- batch inverse
// cuda/batch_inverse.cu:8-41 — chunks of 64 with the kernel's zero handling.
void batch_inverse(const uint32_t* in, uint32_t* out, uint64_t n) {
    const uint64_t CH = 64;
    uint64_t chunks = (n + CH - 1) / CH;
    parallel_for(chunks, [&](uint64_t cb, uint64_t ce) {
        uint32_t prefix[64];
        for (uint64_t c = cb; c < ce; c++) {
            uint64_t start = c * CH, end = start + CH < n ? start + CH : n;
            uint64_t count = end - start;
            prefix[0] = in[start];
            for (uint64_t i = 1; i < count; i++) {
                uint32_t v = in[start + i];
                prefix[i] = v == 0 ? prefix[i - 1] : mul(prefix[i - 1], v);
            }
            uint32_t iv = inv(prefix[count - 1]);
            for (uint64_t i = count - 1; i > 0; i--) {
                uint32_t v = in[start + i];
                if (v == 0) {
                    out[start + i] = 0;
                } else {
                    out[start + i] = mul(iv, prefix[i - 1]);
                    iv = mul(iv, v);
                }
            }
            out[start] = in[start] == 0 ? 0 : iv;
        }
    });
}

// ---------------------------------------------------------------- NTT twiddles
// circle.rs:357-451 (compute_both_twiddles_gpu) with twiddles.cu:127-142 (extract_and_squash).
struct NttTw {
    uint32_t log_n;
    std::vector<uint32_t> line, circle, iline, icircle, offsets, sizes;
};
void build_ntt_twiddles(const Coset& c, NttTw& tw, bool inverses) {
    uint32_t log_n = c.log_size;
    uint64_t n = 1ull << log_n;
    uint32_t n_line = log_n > 0 ? log_n - 1 : 0;
    tw.log_n = log_n;
    std::vector<uint32_t> xs, ys;
    all_points(c, xs, ys);
    tw.circle.assign(ys.begin(), ys.begin() + n / 2);
    tw.line.assign(n > 0 ? n - 1 : 0, 0);
    tw.offsets.clear();
    tw.sizes.clear();
    std::vector<uint32_t> cur = std::move(xs), nxt;
    uint64_t cur_n = n, off = 0;
    for (uint32_t l = 0; l < n_line; l++) {
        uint64_t ls = cur_n / 2;
        tw.offsets.push_back((uint32_t)off);
        tw.sizes.push_back((uint32_t)ls);
        nxt.assign(ls, 0);
        for (uint64_t i = 0; i < ls; i++) {
            uint32_t x = cur[2 * i];
            tw.line[off + i] = x;
            uint32_t x2 = mul(x, x);
            nxt[i] = sub(add(x2, x2), 1);
        }
        off += ls;
        cur.swap(nxt);
        cur_n = ls;
    }
    if (inverses) {
        tw.iline.assign(tw.line.size(), 0);
        tw.icircle.assign(tw.circle.size(), 0);
        if (!tw.line.empty()) batch_inverse(tw.line.data(), tw.iline.data(), tw.line.size());
        if (!tw.circle.empty()) batch_inverse(tw.circle.data(), tw.icircle.data(), tw.circle.size());
    }
}

// ---------------------------------------------------------------- NTT
// cuda/circle_ntt.cu:8-19 (butterflies), :22-50 (layer kernel)
void ntt_layer(uint32_t* d, const uint32_t* tw, uint32_t li, uint64_t n, bool fwd) {
    uint64_t half = n / 2;
    parallel_for(half, [&](uint64_t b, uint64_t e) {
        uint64_t stride = 1ull << li;
        for (uint64_t t = b; t < e; t++) {
            uint64_t h = t >> li, l = t & (stride - 1);
            uint64_t i0 = (h << (li + 1)) + l, i1 = i0 + stride;
            uint32_t v0 = d[i0], v1 = d[i1], w = tw[h];
            if (fwd) {
                uint32_t tmp = mul(v1, w);
                d[i1] = sub(v0, tmp);
                d[i0] = add(v0, tmp);
            } else {
                d[i0] = add(v0, v1);
                d[i1] = mul(sub(v0, v1), w);
            }
        }
    });
}
// circle_ntt.cu:181-210
void ntt_evaluate(uint32_t* d, const uint32_t* line, const uint32_t* circle, const uint32_t* offsets,
                  uint32_t log_n) {
    uint64_t n = 1ull << log_n;
    int n_line = log_n > 0 ? (int)log_n - 1 : 0;
    for (int l = n_line - 1; l >= 0; l--) ntt_layer(d, line + offsets[l], (uint32_t)l + 1, n, true);
    ntt_layer(d, circle, 0, n, true);
}
// circle_ntt.cu:212-249
void ntt_interpolate(uint32_t* d, const uint32_t* iline, const uint32_t* icircle, const uint32_t* offsets,
                     uint32_t log_n) {
    uint64_t n = 1ull << log_n;
    uint32_t n_line = log_n > 0 ? log_n - 1 : 0;
    ntt_layer(d, icircle, 0, n, false);
    for (uint32_t l = 0; l < n_line; l++) ntt_layer(d, iline + offsets[l], l + 1, n, false);
    uint32_t ex = (30u * log_n) % 31u;
    uint32_t inv_n = ex == 0 ? 1u : (1u << ex);
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) d[i] = mul(d[i], inv_n);
    });
}
std::vector<uint32_t> offsets_for(uint32_t log_n) {
    std::vector<uint32_t> o;
    uint64_t n = 1ull << log_n, off = 0, cur = n;
    for (uint32_t l = 0; l + 1 < log_n; l++) {
        o.push_back((uint32_t)off);
        off += cur / 2;
        cur /= 2;
    }
    return o;
}

// ---------------------------------------------------------------- Blake2s
// cuda/include/blake2s.cuh (compress, IV, G), channel.rs:222-298 (single block helper)
const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                        0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
const uint8_t SIGMA[10][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0},
};
inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
inline void G(uint32_t* v, int a, int b, int c, int d, uint32_t x, uint32_t y) {
    v[a] = v[a] + v[b] + x;
    v[d] = rotr(v[d] ^ v[a], 16);
    v[c] = v[c] + v[d];
    v[b] = rotr(v[b] ^ v[c], 12);
    v[a] = v[a] + v[b] + y;
    v[d] = rotr(v[d] ^ v[a], 8);
    v[c] = v[c] + v[d];
    v[b] = rotr(v[b] ^ v[c], 7);
}
void compress(uint32_t h[8], const uint32_t m[16], uint32_t t, uint32_t f) {
    uint32_t v[16];
    for (int i = 0; i < 8; i++) {
        v[i] = h[i];
        v[i + 8] = IV[i];
    }
    v[12] ^= t;
    v[14] ^= f;
    for (int r = 0; r < 10; r++) {
        const uint8_t* s = SIGMA[r];
        G(v, 0, 4, 8, 12, m[s[0]], m[s[1]]);
        G(v, 1, 5, 9, 13, m[s[2]], m[s[3]]);
        G(v, 2, 6, 10, 14, m[s[4]], m[s[5]]);
        G(v, 3, 7, 11, 15, m[s[6]], m[s[7]]);
        G(v, 0, 5, 10, 15, m[s[8]], m[s[9]]);
        G(v, 1, 6, 11, 12, m[s[10]], m[s[11]]);
        G(v, 2, 7, 8, 13, m[s[12]], m[s[13]]);
        G(v, 3, 4, 9, 14, m[s[14]], m[s[15]]);
    }
    for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[i + 8];
}
inline void init_h(uint32_t h[8]) {
    for (int i = 0; i < 8; i++) h[i] = IV[i];
    h[0] ^= 0x01010020u;  // digest 32, fanout 1, depth 1; h[6] untouched (domain 0)
}
inline uint32_t le32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
inline void put_le32(uint8_t* p, uint32_t w) {
    p[0] = (uint8_t)w;
    p[1] = (uint8_t)(w >> 8);
    p[2] = (uint8_t)(w >> 16);
    p[3] = (uint8_t)(w >> 24);
}
// channel.rs:222-298: one block, t = min(len, 64), final flag set.
void blake2s_single(const uint8_t* in, uint64_t len, uint8_t out[32]) {
    uint8_t mb[64] = {0};
    uint64_t l = len < 64 ? len : 64;
    memcpy(mb, in, l);
    uint32_t m[16], h[8];
    for (int i = 0; i < 16; i++) m[i] = le32(mb + 4 * i);
    init_h(h);
    compress(h, m, (uint32_t)l, 0xFFFFFFFFu);
    for (int i = 0; i < 8; i++) put_le32(out + 4 * i, h[i]);
}
// Standard Blake2s-256 (RFC 7693), as used through the `blake2` crate by
// Channel::mix_felts (channel.rs:48-57).
void blake2s_full(const uint8_t* in, uint64_t len, uint8_t out[32]) {
    uint32_t h[8], m[16];
    init_h(h);
    uint64_t off = 0;
    while (len - off > 64) {
        for (int i = 0; i < 16; i++) m[i] = le32(in + off + 4 * i);
        off += 64;
        compress(h, m, (uint32_t)off, 0);
    }
    uint8_t mb[64] = {0};
    memcpy(mb, in + off, len - off);
    for (int i = 0; i < 16; i++) m[i] = le32(mb + 4 * i);
    compress(h, m, (uint32_t)len, 0xFFFFFFFFu);
    for (int i = 0; i < 8; i++) put_le32(out + 4 * i, h[i]);
}

// ---------------------------------------------------------------- Merkle
// Leaf hash: cuda/blake2s.cu:13-62 / merkle.rs:538-556 — words of the first
// min(n_cols,16) columns, t = 4*min(n_cols,16).
inline void leaf_hash_words(const uint32_t* vals, uint32_t n, uint32_t out[8]) {
    uint32_t eff = n > 16 ? 16 : n;
    uint32_t m[16] = {0};
    for (uint32_t i = 0; i < eff; i++) m[i] = vals[i];
    init_h(out);
    compress(out, m, eff * 4, 0xFFFFFFFFu);
}
// Node hash: blake2s.cu:65-89 / merkle.rs:827-842 — H(left || right), t = 64.
inline void node_hash(const uint32_t* l, const uint32_t* r, uint32_t out[8]) {
    uint32_t m[16];
    for (int i = 0; i < 8; i++) {
        m[i] = l[i];
        m[8 + i] = r[i];
    }
    init_h(out);
    compress(out, m, 64, 0xFFFFFFFFu);
}
void leaf_layer(const uint32_t* const* cols, uint32_t n_cols, uint64_t n_leaves, uint32_t* out) {
    parallel_for(n_leaves, [&](uint64_t b, uint64_t e) {
        uint32_t vals[16];
        uint32_t eff = n_cols > 16 ? 16 : n_cols;
        for (uint64_t i = b; i < e; i++) {
            for (uint32_t c = 0; c < eff; c++) vals[c] = cols[c][i];
            leaf_hash_words(vals, n_cols, out + 8 * i);
        }
    });
}
void node_layer(const uint32_t* ch, uint64_t np, uint32_t* par) {
    parallel_for(np, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) node_hash(ch + 16 * i, ch + 16 * i + 8, par + 8 * i);
    });
}
// Full tree as layers[0] = leaves ... layers[depth] = root (merkle.rs:772-782).
struct Tree {
    std::vector<std::vector<uint32_t>> layers;
    const uint32_t* root() const { return layers.back().data(); }
    uint32_t depth() const { return (uint32_t)layers.size() - 1; }
    // merkle.rs:700-711 / prover.rs:2221-2232: siblings from the leaf level up, root excluded.
    void path(uint64_t idx, std::vector<uint32_t>& out) const {
        out.clear();
        for (uint32_t l = 0; l < depth(); l++) {
            const uint32_t* s = layers[l].data() + 8 * (idx ^ 1);
            out.insert(out.end(), s, s + 8);
            idx >>= 1;
        }
    }
};
void build_tree(const uint32_t* const* cols, uint32_t n_cols, uint64_t n, Tree& t) {
    t.layers.clear();
    t.layers.emplace_back(8 * n);
    leaf_layer(cols, n_cols, n, t.layers[0].data());
    while (t.layers.back().size() > 8) {
        uint64_t np = t.layers.back().size() / 16;
        std::vector<uint32_t> par(8 * np);
        node_layer(t.layers.back().data(), np, par.data());
        t.layers.push_back(std::move(par));
    }
}

// ---------------------------------------------------------------- AIR
// air.rs:30-45 (sequential trace)
void fib_trace(uint32_t a, uint32_t b, uint32_t log_n, uint32_t* out) {
    uint64_t n = 1ull << log_n;
    out[0] = a;
    if (n > 1) out[1] = b;
    for (uint64_t i = 2; i < n; i++) {
        uint32_t s = out[i - 1] + out[i - 2];
        out[i] = s >= P ? s - P : s;
    }
}
// constraint.cu:10-36: q[i] = alpha * (t[i+2] - (t[i+1] + t[i])), indices mod n.
void quotient(const uint32_t* t, uint64_t n, QM31 alpha, uint32_t* q0, uint32_t* q1, uint32_t* q2, uint32_t* q3) {
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            uint32_t c = sub(t[(i + 2) % n], add(t[(i + 1) % n], t[i]));
            QM31 r = qmul_m31(alpha, c);
            q0[i] = r.a.a;
            q1[i] = r.a.b;
            q2[i] = r.b.a;
            q3[i] = r.b.b;
        }
    });
}

// ---------------------------------------------------------------- FRI
// twiddles.cu:23-52 + batch_inverse.cu: tw[i] = 1 / coord(domain.at(bitrev_k(2i))).
void fold_twiddles(const Coset& d, bool y, std::vector<uint32_t>& out) {
    uint64_t n = 1ull << d.log_size, half = n / 2;
    std::vector<uint32_t> src(half);
    parallel_for(half, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            Pt p = d.at(bitrev((uint32_t)(i << 1), d.log_size));
            src[i] = y ? p.y : p.x;
        }
    });
    out.assign(half, 0);
    if (half) batch_inverse(src.data(), out.data(), half);
}
// fri.cu:10-39 / fri.rs:411-427: f' = (f0 + f1) + alpha * ((f0 - f1) * tw)
inline QM31 fold_pair(QM31 f0, QM31 f1, QM31 alpha, uint32_t tw) {
    return qadd(qadd(f0, f1), qmul(alpha, qmul_m31(qsub(f0, f1), tw)));
}
struct Soa4 {
    std::vector<uint32_t> c[4];
    void resize(uint64_t n) {
        for (auto& v : c) v.assign(n, 0);
    }
    uint64_t size() const { return c[0].size(); }
    QM31 get(uint64_t i) const { return {{c[0][i], c[1][i]}, {c[2][i], c[3][i]}}; }
    void set(uint64_t i, QM31 q) {
        c[0][i] = q.a.a;
        c[1][i] = q.a.b;
        c[2][i] = q.b.a;
        c[3][i] = q.b.b;
    }
};
void fold_line(const Soa4& in, const std::vector<uint32_t>& tw, QM31 alpha, Soa4& out) {
    uint64_t half = in.size() / 2;
    out.resize(half);
    parallel_for(half, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) out.set(i, fold_pair(in.get(2 * i), in.get(2 * i + 1), alpha, tw[i]));
    });
}
// fri.cu:44-78: dst = dst * alpha^2 + f' (dst is zero-initialised by the prover, fri.rs:30-37).
void fold_circle(Soa4& dst, const Soa4& src, const std::vector<uint32_t>& tw, QM31 alpha) {
    uint64_t half = src.size() / 2;
    QM31 a2 = qmul(alpha, alpha);
    parallel_for(half, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            QM31 fp = fold_pair(src.get(2 * i), src.get(2 * i + 1), alpha, tw[i]);
            dst.set(i, qadd(qmul(dst.get(i), a2), fp));
        }
    });
}

// ---------------------------------------------------------------- channel
// src/channel.rs:15-206
struct Channel {
    uint8_t state[32] = {0};
    uint32_t n_draws = 0;
    void mix_digest(const uint32_t d[8]) {  // :30-39
        uint8_t in[64];
        memcpy(in, state, 32);
        for (int i = 0; i < 8; i++) put_le32(in + 32 + 4 * i, d[i]);
        blake2s_single(in, 64, state);
        n_draws = 0;
    }
    void mix_u64(uint64_t v) {  // :117-128, the PoW nonce: H(state || lo LE || hi LE)
        uint8_t in[40];
        memcpy(in, state, 32);
        put_le32(in + 32, (uint32_t)v);
        put_le32(in + 36, (uint32_t)(v >> 32));
        blake2s_single(in, 40, state);
        n_draws = 0;
    }
    void mix_felts(const uint32_t* w, uint64_t n_felts) {  // :48-61
        if (n_felts == 0) return;
        std::vector<uint8_t> in(32 + 16 * n_felts);
        memcpy(in.data(), state, 32);
        for (uint64_t i = 0; i < 4 * n_felts; i++) put_le32(in.data() + 32 + 4 * i, w[i]);
        blake2s_full(in.data(), in.size(), state);
        n_draws = 0;
    }
    void squeeze(uint8_t out[32]) {  // :192-199
        uint8_t in[37];
        memcpy(in, state, 32);
        put_le32(in + 32, n_draws);
        in[36] = 0;
        n_draws++;
        blake2s_single(in, 37, out);
    }
    QM31 draw_felt() {  // :67-86
        for (;;) {
            uint8_t b[32];
            squeeze(b);
            uint32_t w[8];
            bool ok = true;
            for (int i = 0; i < 8; i++) {
                w[i] = le32(b + 4 * i);
                if (w[i] >= 2 * P) ok = false;
            }
            if (ok) return {{w[0] % P, w[1] % P}, {w[2] % P, w[3] % P}};
        }
    }
    uint64_t draw_number(uint64_t bound) {  // :94-98
        uint8_t b[32];
        squeeze(b);
        uint64_t raw = (uint64_t)le32(b) | ((uint64_t)le32(b + 4) << 32);
        return raw % bound;
    }
};

// ---------------------------------------------------------------- proof
// prover.rs:101-137
struct Decom {
    uint32_t width;  // 1 (trace) or 4 (SoA4)
    std::vector<uint32_t> values, sibling_values;           // width words per query
    std::vector<std::vector<uint32_t>> paths, sibling_paths;  // 8 words per level
};
struct Proof {
    uint32_t trace_root[8], quotient_root[8];
    std::vector<std::array<uint32_t, 8>> fri_roots;
    std::vector<uint32_t> last_layer;  // 4 words per QM31
    uint32_t log_n, a, b;
    std::vector<uint64_t> queries;
    Decom trace, quotient;
    std::vector<Decom> fri;
};

void decommit(const Tree& t, const uint32_t* const* cols, uint32_t width, const std::vector<uint64_t>& idx, Decom& d) {
    // prover.rs:1184-1198 and 1230-1247: values at idx, siblings at idx^1, both auth paths.
    d.width = width;
    d.values.clear();
    d.sibling_values.clear();
    d.paths.clear();
    d.sibling_paths.clear();
    std::vector<uint32_t> p;
    for (uint64_t i : idx) {
        for (uint32_t c = 0; c < width; c++) d.values.push_back(cols[c][i]);
    }
    for (uint64_t i : idx) {
        for (uint32_t c = 0; c < width; c++) d.sibling_values.push_back(cols[c][i ^ 1]);
    }
    for (uint64_t i : idx) {
        t.path(i, p);
        d.paths.push_back(p);
    }
    for (uint64_t i : idx) {
        t.path(i ^ 1, p);
        d.sibling_paths.push_back(p);
    }
}

// Twiddles and FRI twiddles for one log size (the ProverCache, prover.rs:143-189).
struct Cache {
    uint32_t log_n, log_e;
    NttTw trace_tw, eval_tw;
    std::vector<std::vector<uint32_t>> fri_tw;  // [0] circle on half_coset(log_e); [j] line on half_odds(log_e-j)
    void build(uint32_t ln, uint32_t log_last) {
        log_n = ln;
        log_e = ln + 2;  // BLOWUP_BITS = 2 (prover.rs:86)
        build_ntt_twiddles(half_coset(ln), trace_tw, true);
        // prover.rs:626-630: the whole group when log_e = 31 (no half coset exists)
        build_ntt_twiddles(log_e <= 30 ? half_coset(log_e) : subgroup(log_e), eval_tw, false);
        // fri.rs:143-160 (precompute_fri_twiddles(log_e, stop))
        fri_tw.clear();
        fri_tw.emplace_back();
        fold_twiddles(log_e <= 30 ? half_coset(log_e) : subgroup(log_e), true, fri_tw.back());
        for (uint32_t ls = log_e - 1; ls > log_last; ls--) {
            fri_tw.emplace_back();
            fold_twiddles(half_odds(ls), false, fri_tw.back());
        }
    }
};

// prover.rs:1721-1978 (prove_with_cache), parameterised by n_queries and the last layer size.
void prove(uint32_t a, uint32_t b, const Cache& C, uint32_t n_queries, uint32_t log_last, Proof& pf,
           const uint32_t* trace_in) {
    uint32_t log_n = C.log_n, log_e = C.log_e;
    uint64_t n = 1ull << log_n, E = 1ull << log_e;
    pf.log_n = log_n;
    pf.a = a;
    pf.b = b;
    // Step 1-3: trace -> interpolate -> zero pad -> evaluate
    std::vector<uint32_t> eval(E, 0);
    memcpy(eval.data(), trace_in, n * 4);
    ntt_interpolate(eval.data(), C.trace_tw.iline.data(), C.trace_tw.icircle.data(), C.trace_tw.offsets.data(),
                    log_n);
    // zero pad: eval[n..E) already zero (constraint.cu:81-90)
    ntt_evaluate(eval.data(), C.eval_tw.line.data(), C.eval_tw.circle.data(), C.eval_tw.offsets.data(), log_e);
    // Step 4: trace commitment (1 column)
    Tree ttree;
    const uint32_t* tcols[1] = {eval.data()};
    build_tree(tcols, 1, E, ttree);
    memcpy(pf.trace_root, ttree.root(), 32);
    // Step 5: quotient
    Channel ch;
    ch.mix_digest(pf.trace_root);
    QM31 alpha = ch.draw_felt();
    Soa4 q;
    q.resize(E);
    quotient(eval.data(), E, alpha, q.c[0].data(), q.c[1].data(), q.c[2].data(), q.c[3].data());
    // Step 6: quotient commitment
    Tree qtree;
    const uint32_t* qcols[4] = {q.c[0].data(), q.c[1].data(), q.c[2].data(), q.c[3].data()};
    build_tree(qcols, 4, E, qtree);
    memcpy(pf.quotient_root, qtree.root(), 32);
    ch.mix_digest(pf.quotient_root);
    // Step 7: FRI
    std::vector<Soa4> layers;
    std::vector<Tree> trees;
    QM31 fa = ch.draw_felt();
    layers.emplace_back();
    layers.back().resize(E / 2);
    fold_circle(layers.back(), q, C.fri_tw[0], fa);
    uint32_t cur = log_e - 1;
    auto commit_layer = [&]() {
        trees.emplace_back();
        const Soa4& L = layers.back();
        const uint32_t* lc[4] = {L.c[0].data(), L.c[1].data(), L.c[2].data(), L.c[3].data()};
        build_tree(lc, 4, L.size(), trees.back());
        std::array<uint32_t, 8> r;
        memcpy(r.data(), trees.back().root(), 32);
        pf.fri_roots.push_back(r);
        ch.mix_digest(r.data());
    };
    pf.fri_roots.clear();
    commit_layer();
    uint32_t twi = 1;
    while (cur > log_last) {  // prover.rs:1858-1906 (GPU loop and CPU tail share the arithmetic)
        QM31 al = ch.draw_felt();
        Soa4 nx;
        fold_line(layers.back(), C.fri_tw[twi], al, nx);
        twi++;
        cur--;
        layers.push_back(std::move(nx));
        commit_layer();
    }
    const Soa4& last = layers.back();
    pf.last_layer.clear();
    for (uint64_t i = 0; i < last.size(); i++)
        for (int c = 0; c < 4; c++) pf.last_layer.push_back(last.c[c][i]);
    // Step 8: queries (prover.rs:1922-1927)
    ch.mix_felts(pf.last_layer.data(), last.size());
    pf.queries.clear();
    for (uint32_t i = 0; i < n_queries; i++) pf.queries.push_back(ch.draw_number(E));
    decommit(ttree, tcols, 1, pf.queries, pf.trace);
    decommit(qtree, qcols, 4, pf.queries, pf.quotient);
    std::vector<uint64_t> fi;
    for (uint64_t qi : pf.queries) fi.push_back(qi / 2);
    pf.fri.clear();
    for (size_t j = 0; j < layers.size(); j++) {
        const Soa4& L = layers[j];
        const uint32_t* lc[4] = {L.c[0].data(), L.c[1].data(), L.c[2].data(), L.c[3].data()};
        pf.fri.emplace_back();
        decommit(trees[j], lc, 4, fi, pf.fri.back());
        for (auto& v : fi) v /= 2;
    }
}

// ---------------------------------------------------------------- proof summary, memory-lean
// The commitments of prove() without the trees: the same trace/quotient roots,
// FRI roots, last layer and query indices (prover.rs:1721-1927), computed with
// root-only Merkle commitments over leaves produced on the fly (the quotient
// from the LDE, as the reference's fused path does, prover.rs:696-733) and each
// FRI layer freed once the next exists. Twiddles are built lean as well. It
// serves traces whose full trees do not fit host memory (2^29 rows: a 2^31-leaf
// tree alone is 128 GiB); at sizes where prove() runs the two must agree.

// Twiddles of build_ntt_twiddles without the all-points arrays: circle[i] =
// y(pts[i]) for i < n/2; line layer 0 holds x(pts[2i]); every later layer is
// computed from the one before (circle.rs:394-425: X_{l+1}[i] = 2 X_l[2i]^2 - 1,
// and layer l keeps the even entries X_l[2i]).
void build_ntt_twiddles_lean(const Coset& c, NttTw& tw, bool inverses) {
    uint32_t log_n = c.log_size;
    uint64_t n = 1ull << log_n;
    uint32_t n_line = log_n > 0 ? log_n - 1 : 0;
    tw.log_n = log_n;
    tw.circle.assign(n / 2, 0);
    tw.line.assign(n > 0 ? n - 1 : 0, 0);
    tw.offsets = offsets_for(log_n);
    tw.sizes.clear();
    for (uint32_t l = 0; l < n_line; l++) tw.sizes.push_back((uint32_t)(n >> (l + 1)));
    const uint64_t half = n / 2;
    const uint32_t nt = n_threads();
    const uint64_t chunk = (half + nt - 1) / nt;
    const Pt step2 = pmul(c.step, c.step);
    parallel_for(nt, [&](uint64_t tb, uint64_t te) {
        for (uint64_t t = tb; t < te; t++) {
            uint64_t b = t * chunk, e = b + chunk < half ? b + chunk : half;
            if (b >= e) continue;
            Pt p = c.at(b), q = c.at(2 * b);
            for (uint64_t i = b; i < e; i++) {
                tw.circle[i] = p.y;
                if (n_line) tw.line[i] = q.x;
                p = pmul(p, c.step);
                q = pmul(q, step2);
            }
        }
    });
    for (uint32_t l = 1; l < n_line; l++) {
        const uint32_t* prev = tw.line.data() + tw.offsets[l - 1];
        uint32_t* cur = tw.line.data() + tw.offsets[l];
        parallel_for(n >> (l + 1), [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; i++) {
                uint32_t x = prev[2 * i];
                uint32_t x2 = mul(x, x);
                cur[i] = sub(add(x2, x2), 1);
            }
        });
    }
    if (inverses) {
        tw.iline.assign(tw.line.size(), 0);
        tw.icircle.assign(tw.circle.size(), 0);
        if (!tw.line.empty()) batch_inverse(tw.line.data(), tw.iline.data(), tw.line.size());
        if (!tw.circle.empty()) batch_inverse(tw.circle.data(), tw.icircle.data(), tw.circle.size());
    }
}

// Root of the complete binary tree over n leaves (n a power of two), leaf(i, h)
// writing leaf i's hash: subtrees of up to 2^14 leaves are reduced on their own
// threads, then the tree over their roots.
template <class F>
void stream_root(uint64_t n, F&& leaf, uint32_t root[8]) {
    const uint64_t S = n < (1ull << 14) ? n : (1ull << 14);
    const uint64_t ns = n / S;
    std::vector<uint32_t> tops(8 * ns);
    parallel_for(ns, [&](uint64_t b, uint64_t e) {
        std::vector<uint32_t> buf(8 * S);
        for (uint64_t s = b; s < e; s++) {
            for (uint64_t i = 0; i < S; i++) leaf(s * S + i, buf.data() + 8 * i);
            for (uint64_t m = S; m > 1; m /= 2)
                for (uint64_t i = 0; i < m / 2; i++) node_hash(&buf[16 * i], &buf[16 * i + 8], &buf[8 * i]);
            memcpy(&tops[8 * s], buf.data(), 32);
        }
    });
    std::vector<uint32_t> nx;
    while (tops.size() > 8) {
        nx.assign(tops.size() / 2, 0);
        node_layer(tops.data(), tops.size() / 16, nx.data());
        tops.swap(nx);
    }
    memcpy(root, tops.data(), 32);
}
inline void soa4_leaf(const Soa4& L, uint64_t i, uint32_t h[8]) {
    uint32_t v[4] = {L.c[0][i], L.c[1][i], L.c[2][i], L.c[3][i]};
    leaf_hash_words(v, 4, h);
}

void prove_summary(uint32_t a, uint32_t b, uint32_t log_n, uint32_t n_queries, uint32_t log_last, Proof& pf) {
    const uint32_t log_e = log_n + 2;
    const uint64_t n = 1ull << log_n, E = 1ull << log_e;
    pf.log_n = log_n;
    pf.a = a;
    pf.b = b;
    std::vector<uint32_t> eval(E, 0);
    fib_trace(a, b, log_n, eval.data());
    {
        NttTw ttw;
        build_ntt_twiddles_lean(half_coset(log_n), ttw, true);
        ntt_interpolate(eval.data(), ttw.iline.data(), ttw.icircle.data(), ttw.offsets.data(), log_n);
    }
    {
        NttTw etw;
        build_ntt_twiddles_lean(log_e <= 30 ? half_coset(log_e) : subgroup(log_e), etw, false);
        ntt_evaluate(eval.data(), etw.line.data(), etw.circle.data(), etw.offsets.data(), log_e);
    }
    const uint32_t* ev = eval.data();
    stream_root(E, [&](uint64_t i, uint32_t* h) { leaf_hash_words(ev + i, 1, h); }, pf.trace_root);
    Channel ch;
    ch.mix_digest(pf.trace_root);
    const QM31 alpha = ch.draw_felt();
    auto q_at = [&](uint64_t i) {
        return qmul_m31(alpha, sub(ev[(i + 2) % E], add(ev[(i + 1) % E], ev[i])));
    };
    stream_root(E, [&](uint64_t i, uint32_t* h) {
        uint32_t v[4];
        qto(q_at(i), v);
        leaf_hash_words(v, 4, h);
    }, pf.quotient_root);
    ch.mix_digest(pf.quotient_root);
    const QM31 fa = ch.draw_felt();
    Soa4 cur;
    {
        std::vector<uint32_t> tw;
        fold_twiddles(log_e <= 30 ? half_coset(log_e) : subgroup(log_e), true, tw);
        cur.resize(E / 2);
        parallel_for(E / 2, [&](uint64_t b0, uint64_t e0) {
            for (uint64_t i = b0; i < e0; i++) cur.set(i, fold_pair(q_at(2 * i), q_at(2 * i + 1), fa, tw[i]));
        });
    }
    std::vector<uint32_t>().swap(eval);
    pf.fri_roots.clear();
    auto commit = [&]() {
        std::array<uint32_t, 8> r;
        stream_root(cur.size(), [&](uint64_t i, uint32_t* h) { soa4_leaf(cur, i, h); }, r.data());
        pf.fri_roots.push_back(r);
        ch.mix_digest(r.data());
    };
    commit();
    for (uint32_t ls = log_e - 1; ls > log_last; ls--) {
        const QM31 al = ch.draw_felt();
        std::vector<uint32_t> tw;
        fold_twiddles(half_odds(ls), false, tw);
        Soa4 nx;
        fold_line(cur, tw, al, nx);
        cur = std::move(nx);
        commit();
    }
    pf.last_layer.clear();
    for (uint64_t i = 0; i < cur.size(); i++)
        for (int c = 0; c < 4; c++) pf.last_layer.push_back(cur.c[c][i]);
    ch.mix_felts(pf.last_layer.data(), cur.size());
    pf.queries.clear();
    for (uint32_t i = 0; i < n_queries; i++) pf.queries.push_back(ch.draw_number(E));
    (void)n;
}

// stark_cli.rs:907-975 (KRAK v1)
struct W {
    std::vector<uint8_t> b;
    void u32(uint32_t v) {
        uint8_t t[4];
        put_le32(t, v);
        b.insert(b.end(), t, t + 4);
    }
    void words(const uint32_t* p, size_t n) {
        for (size_t i = 0; i < n; i++) u32(p[i]);
    }
};
void write_decom(W& w, const Decom& d) {  // :918-946
    uint32_t n = (uint32_t)(d.values.size() / d.width);
    w.u32(n);
    w.words(d.values.data(), d.values.size());
    w.words(d.sibling_values.data(), d.sibling_values.size());
    for (auto& p : d.paths) {
        w.u32((uint32_t)(p.size() / 8));
        w.words(p.data(), p.size());
    }
    for (auto& p : d.sibling_paths) {
        w.u32((uint32_t)(p.size() / 8));
        w.words(p.data(), p.size());
    }
}
std::vector<uint8_t> serialize(const Proof& pf) {  // :948-975
    W w;
    w.u32(0x4B52414Bu);
    w.u32(1);
    w.u32(pf.log_n);
    w.u32(pf.a);
    w.u32(pf.b);
    w.words(pf.trace_root, 8);
    w.words(pf.quotient_root, 8);
    w.u32((uint32_t)pf.fri_roots.size());
    for (auto& r : pf.fri_roots) w.words(r.data(), 8);
    w.u32((uint32_t)(pf.last_layer.size() / 4));
    w.words(pf.last_layer.data(), pf.last_layer.size());
    w.u32((uint32_t)pf.queries.size());
    for (uint64_t q : pf.queries) w.u32((uint32_t)q);
    write_decom(w, pf.trace);
    write_decom(w, pf.quotient);
    w.u32((uint32_t)pf.fri.size());
    for (auto& d : pf.fri) write_decom(w, d);
    return std::move(w.b);
}

// stark_cli.rs:1003-1089 (deserialize), with bounds checks instead of panics.
struct R {
    const uint8_t* p;
    uint64_t n, off = 0;
    bool bad = false;
    uint32_t u32() {
        if (off + 4 > n) {
            bad = true;
            return 0;
        }
        uint32_t v = le32(p + off);
        off += 4;
        return v;
    }
};
bool read_decom(R& r, uint32_t width, Decom& d, uint32_t max_n) {
    d.width = width;
    uint32_t n = r.u32();
    if (r.bad || n > max_n) return false;
    d.values.resize((uint64_t)n * width);
    d.sibling_values.resize((uint64_t)n * width);
    for (auto& v : d.values) v = r.u32();
    for (auto& v : d.sibling_values) v = r.u32();
    d.paths.assign(n, {});
    d.sibling_paths.assign(n, {});
    for (int s = 0; s < 2; s++) {
        auto& ps = s == 0 ? d.paths : d.sibling_paths;
        for (auto& p : ps) {
            uint32_t depth = r.u32();
            if (r.bad || depth > 64) return false;
            p.resize(8ull * depth);
            for (auto& v : p) v = r.u32();
        }
    }
    return !r.bad;
}
bool deserialize(const uint8_t* data, uint64_t len, Proof& pf) {
    R r{data, len};
    if (r.u32() != 0x4B52414Bu) return false;
    if (r.u32() != 1) return false;
    pf.log_n = r.u32();
    pf.a = r.u32();
    pf.b = r.u32();
    for (auto& v : pf.trace_root) v = r.u32();
    for (auto& v : pf.quotient_root) v = r.u32();
    uint32_t nf = r.u32();
    if (r.bad || nf > 64) return false;
    pf.fri_roots.assign(nf, {});
    for (auto& rt : pf.fri_roots)
        for (auto& v : rt) v = r.u32();
    uint32_t nl = r.u32();
    if (r.bad || nl > (1u << 20)) return false;
    pf.last_layer.resize(4ull * nl);
    for (auto& v : pf.last_layer) v = r.u32();
    uint32_t nq = r.u32();
    if (r.bad || nq > 4096) return false;
    pf.queries.resize(nq);
    for (auto& q : pf.queries) q = r.u32();
    if (!read_decom(r, 1, pf.trace, 4096)) return false;
    if (!read_decom(r, 4, pf.quotient, 4096)) return false;
    uint32_t nd = r.u32();
    if (r.bad || nd > 64) return false;
    pf.fri.assign(nd, {});
    for (auto& d : pf.fri)
        if (!read_decom(r, 4, d, 4096)) return false;
    return !r.bad && r.off == len;
}

// ---------------------------------------------------------------- verifier
// verifier.rs:30-36
uint32_t fold_twiddle_at(const Coset& d, uint64_t folded, bool circle) {
    Pt p = d.at(bitrev((uint32_t)(folded << 1), d.log_size));
    return inv(circle ? p.y : p.x);
}
bool verify_path(const uint32_t root[8], const uint32_t leaf[8], uint64_t idx, const std::vector<uint32_t>& path) {
    // merkle.rs:494-534
    uint32_t cur[8], nx[8];
    memcpy(cur, leaf, 32);
    for (size_t l = 0; l < path.size() / 8; l++) {
        const uint32_t* s = path.data() + 8 * l;
        if (idx % 2 == 0)
            node_hash(cur, s, nx);
        else
            node_hash(s, cur, nx);
        memcpy(cur, nx, 32);
        idx /= 2;
    }
    return memcmp(cur, root, 32) == 0;
}
// verifier.rs:256-310
int check_decom(const uint32_t root[8], const Decom& d, const std::vector<uint64_t>& idx, uint32_t nq) {
    uint32_t w = d.width;
    if (d.values.size() != (uint64_t)nq * w || d.paths.size() != nq) return 1;
    if (d.sibling_values.size() != (uint64_t)nq * w || d.sibling_paths.size() != nq) return 1;
    for (uint32_t q = 0; q < nq; q++) {
        uint32_t lh[8];
        leaf_hash_words(d.values.data() + (uint64_t)q * w, w, lh);
        if (!verify_path(root, lh, idx[q], d.paths[q])) return 2;
        leaf_hash_words(d.sibling_values.data() + (uint64_t)q * w, w, lh);
        if (!verify_path(root, lh, idx[q] ^ 1, d.sibling_paths[q])) return 3;
    }
    return 0;
}
QM31 q4(const uint32_t* v) { return qfrom(v); }
// verifier.rs:39-236, with N_QUERIES and the last-layer size as parameters.
int verify(const Proof& pf, uint32_t nq, uint32_t log_last) {
    uint32_t log_e = pf.log_n + 2;
    uint64_t E = 1ull << log_e;
    if (pf.log_n < 1 || log_e > 31) return 10;
    if (pf.a >= P || pf.b >= P) return 11;
    if (pf.queries.size() != nq) return 12;
    Channel ch;
    ch.mix_digest(pf.trace_root);
    (void)ch.draw_felt();
    ch.mix_digest(pf.quotient_root);
    std::vector<QM31> alphas;
    alphas.push_back(ch.draw_felt());
    for (auto& r : pf.fri_roots) {
        ch.mix_digest(r.data());
        alphas.push_back(ch.draw_felt());
    }
    if (log_e < log_last || pf.fri_roots.size() != log_e - log_last) return 13;
    if (pf.last_layer.size() != 4ull << log_last) return 14;
    ch.mix_felts(pf.last_layer.data(), pf.last_layer.size() / 4);
    for (uint32_t i = 0; i < nq; i++)
        if (pf.queries[i] != ch.draw_number(E)) return 15;
    uint32_t zero[8] = {0};
    if (!memcmp(pf.trace_root, zero, 32)) return 16;
    if (!memcmp(pf.quotient_root, zero, 32)) return 17;
    if (check_decom(pf.trace_root, pf.trace, pf.queries, nq)) return 20;
    if (check_decom(pf.quotient_root, pf.quotient, pf.queries, nq)) return 21;
    if (pf.fri.size() != pf.fri_roots.size()) return 22;
    std::vector<uint64_t> fi;
    for (auto q : pf.queries) fi.push_back(q / 2);
    for (size_t j = 0; j < pf.fri.size(); j++) {
        if (check_decom(pf.fri_roots[j].data(), pf.fri[j], fi, nq)) return 23;
        for (auto& v : fi) v /= 2;
    }
    size_t nfl = pf.fri.size();
    for (uint32_t q = 0; q < nq; q++) {
        uint64_t cur = pf.queries[q];
        uint32_t cl = log_e;
        {
            Coset d = cl <= 30 ? half_coset(cl) : subgroup(cl);
            uint64_t f = cur / 2;
            const uint32_t* v = pf.quotient.values.data() + 4ull * q;
            const uint32_t* s = pf.quotient.sibling_values.data() + 4ull * q;
            QM31 f0 = cur % 2 == 0 ? q4(v) : q4(s), f1 = cur % 2 == 0 ? q4(s) : q4(v);
            QM31 exp = fold_pair(f0, f1, alphas[0], fold_twiddle_at(d, f, true));
            QM31 act = q4(pf.fri[0].values.data() + 4ull * q);
            if (memcmp(&exp, &act, 16)) return 30;
            cur = f;
            cl--;
        }
        for (size_t l = 0; l + 1 < nfl; l++) {
            Coset d = half_odds(cl);
            uint64_t f = cur / 2;
            const uint32_t* v = pf.fri[l].values.data() + 4ull * q;
            const uint32_t* s = pf.fri[l].sibling_values.data() + 4ull * q;
            QM31 f0 = cur % 2 == 0 ? q4(v) : q4(s), f1 = cur % 2 == 0 ? q4(s) : q4(v);
            QM31 exp = fold_pair(f0, f1, alphas[l + 1], fold_twiddle_at(d, f, false));
            QM31 act = q4(pf.fri[l + 1].values.data() + 4ull * q);
            if (memcmp(&exp, &act, 16)) return 31;
            cur = f;
            cl--;
        }
        if (nfl > 0) {
            if (cur >= pf.last_layer.size() / 4) return 32;
            if (memcmp(pf.fri[nfl - 1].values.data() + 4ull * q, pf.last_layer.data() + 4 * cur, 16)) return 33;
        }
    }
    return 0;
}


// ---------------------------------------------------------------- stwo-basis NTT
// Twiddles of src/ntt.rs:262-324 (StwoNttCache::new, restating stwo's
// slow_precompute_twiddles): for a domain of size 2^log_n, start from
// half_odds(log_n - 1); per level store the x-coordinates of the first half of
// the coset in bit-reversed order, then double the coset; finally append 1.
// Total 2^(log_n-1) words.
void stwo_twiddles(uint32_t log_n, std::vector<uint32_t>& fwd, std::vector<uint32_t>& inv_out) {
    uint64_t tc = 1ull << (log_n - 1);
    Coset cur = half_odds(log_n - 1);
    fwd.clear();
    for (uint32_t lv = 0, L = cur.log_size; lv < L; lv++) {
        uint64_t half = (1ull << cur.log_size) / 2;
        uint32_t layer_log = 0;
        while ((1ull << layer_log) < half) layer_log++;
        std::vector<uint32_t> xs(half), brt(half);
        Pt pt = cur.initial;
        for (uint64_t j = 0; j < half; j++) {
            xs[j] = pt.x;
            pt = pmul(pt, cur.step);
        }
        for (uint64_t i = 0; i < half; i++) brt[layer_log == 0 ? 0 : bitrev((uint32_t)i, layer_log)] = xs[i];
        fwd.insert(fwd.end(), brt.begin(), brt.end());
        cur = {pmul(cur.initial, cur.initial), pmul(cur.step, cur.step), cur.log_size - 1};
    }
    fwd.push_back(1);
    while (fwd.size() < tc) fwd.push_back(0);
    fwd.resize(tc);
    // ntt.rs:307-320: Montgomery batch inverse (no zero entries occur)
    inv_out.assign(tc, 0);
    std::vector<uint32_t> prefix(tc);
    prefix[0] = fwd[0];
    for (uint64_t i = 1; i < tc; i++) prefix[i] = mul(prefix[i - 1], fwd[i]);
    uint32_t ip = inv(prefix[tc - 1]);
    for (uint64_t i = tc - 1; i > 0; i--) {
        inv_out[i] = mul(ip, prefix[i - 1]);
        ip = mul(ip, fwd[i]);
    }
    inv_out[0] = ip;
}
// cuda/circle_ntt_stwo.cu: line layer k (stride 2^(k+1)) reads the flat buffer at
// half_n - 2^(n_line - k) with index h = i0 >> (k+2); the circle layer pairs
// (2t, 2t+1) with [y, -y, -x, x][t % 4], x = T[2(t/4)], y = T[2(t/4)+1].
void stwo_ntt(uint32_t* d, const uint32_t* T, uint32_t log_n, bool fwd) {
    uint64_t n = 1ull << log_n, half = n / 2;
    uint32_t n_line = log_n - 1;
    auto circle = [&]() {
        parallel_for(half, [&](uint64_t b, uint64_t e) {
            for (uint64_t t = b; t < e; t++) {
                uint32_t x = T[2 * (t / 4)], y = T[2 * (t / 4) + 1];
                uint32_t w = (t % 4 == 0) ? y : (t % 4 == 1) ? neg(y) : (t % 4 == 2) ? neg(x) : x;
                uint32_t v0 = d[2 * t], v1 = d[2 * t + 1];
                if (fwd) {
                    uint32_t tmp = mul(v1, w);
                    d[2 * t + 1] = sub(v0, tmp);
                    d[2 * t] = add(v0, tmp);
                } else {
                    d[2 * t] = add(v0, v1);
                    d[2 * t + 1] = mul(sub(v0, v1), w);
                }
            }
        });
    };
    auto layer = [&](uint32_t k) {
        const uint32_t* tw = T + (half - (1ull << (n_line - k)));
        ntt_layer(d, tw, k + 1, n, fwd);
    };
    if (fwd) {
        for (int k = (int)n_line - 1; k >= 0; k--) layer((uint32_t)k);
        if (log_n >= 1) circle();
    } else {
        if (log_n >= 1) circle();
        for (uint32_t k = 0; k < n_line; k++) layer(k);
        uint32_t ex = (30u * log_n) % 31u;
        uint32_t inv_n = ex == 0 ? 1u : (1u << ex);
        for (uint64_t i = 0; i < n; i++) d[i] = mul(d[i], inv_n);
    }
}

// ---------------------------------------------------------------- lifted leaves
// cuda/merkle_leaves_lifted.cu:26-98 and crates/cuda-backend/src/merkle_ops.rs:24-148:
// leaf L hashes, with standard multi-block Blake2s, the LE32 words of every
// column (in the given order, ascending log size) at its lifted row.
inline uint32_t lifted_row(uint32_t leaf, uint32_t S, uint32_t L) {
    if (L == S) return leaf;
    uint32_t sh = S - L + 1;
    return ((leaf >> sh) << 1) | (leaf & 1);
}
void lifted_leaves(const uint32_t* const* cols, const uint32_t* logs, uint32_t n_cols, uint32_t S, uint32_t* out) {
    uint64_t n = 1ull << S;
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        std::vector<uint8_t> msg(4 * n_cols);
        for (uint64_t i = b; i < e; i++) {
            for (uint32_t c = 0; c < n_cols; c++) put_le32(msg.data() + 4 * c, cols[c][lifted_row((uint32_t)i, S, logs[c])]);
            uint8_t h[32];
            blake2s_full(msg.data(), msg.size(), h);
            for (int k = 0; k < 8; k++) out[8 * i + k] = le32(h + 4 * k);
        }
    });
}

// ---------------------------------------------------------------- wide commit
// The batched wide-commitment workload: stwo-basis LDE of every column (blowup
// 2^blowup), one lifted Blake2s tree over all of them, then FRI on
// f = sum_c beta^c * lifted_c (beta drawn after the root), the circle fold on the
// canonic domain of size 2^S (fri_ops.rs:40-78 twiddles: 1/y of
// domain.at(bitrev(2i))) and line folds on half_odds (fri.rs:143-160), each
// layer committed as an SoA4 tree, down to 2^log_last.
struct WideOut {
    uint32_t root[8];
    std::vector<std::array<uint32_t, 8>> fri_roots;
    std::vector<uint32_t> last;  // 4 words per element
    uint32_t beta[4];
};
void wide_commit(const std::vector<std::vector<uint32_t>>& trace, const std::vector<uint32_t>& logs, uint32_t blowup,
                 uint32_t log_last, WideOut& out, std::vector<std::vector<uint32_t>>* lde_out) {
    uint32_t C = (uint32_t)trace.size();
    std::vector<std::vector<uint32_t>> ev(C);
    uint32_t S = 0;
    for (uint32_t c = 0; c < C; c++) {
        uint32_t l = logs[c], le = l + blowup;
        std::vector<uint32_t> tw, itw, etw, eitw;
        stwo_twiddles(l, tw, itw);
        stwo_twiddles(le, etw, eitw);
        std::vector<uint32_t> coef(trace[c]);
        stwo_ntt(coef.data(), itw.data(), l, false);
        ev[c].assign(1ull << le, 0);
        memcpy(ev[c].data(), coef.data(), coef.size() * 4);  // extend (poly_ops.rs:496-514)
        stwo_ntt(ev[c].data(), etw.data(), le, true);
        if (le > S) S = le;
    }
    std::vector<uint32_t> elogs(C);
    std::vector<const uint32_t*> cp(C);
    for (uint32_t c = 0; c < C; c++) {
        elogs[c] = logs[c] + blowup;
        cp[c] = ev[c].data();
    }
    uint64_t n = 1ull << S;
    std::vector<uint32_t> lay(8 * n), nx;
    lifted_leaves(cp.data(), elogs.data(), C, S, lay.data());
    while (lay.size() > 8) {
        nx.assign(lay.size() / 2, 0);
        node_layer(lay.data(), lay.size() / 16, nx.data());
        lay.swap(nx);
    }
    memcpy(out.root, lay.data(), 32);
    Channel ch;
    ch.mix_digest(out.root);
    QM31 beta = ch.draw_felt();
    qto(beta, out.beta);
    // f = sum_c beta^c * lifted_c
    Soa4 f;
    f.resize(n);
    std::vector<QM31> pw(C);
    QM31 acc = {{1, 0}, {0, 0}};
    for (uint32_t c = 0; c < C; c++) {
        pw[c] = acc;
        acc = qmul(acc, beta);
    }
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            QM31 s = {{0, 0}, {0, 0}};
            for (uint32_t c = 0; c < C; c++) s = qadd(s, qmul_m31(pw[c], cp[c][lifted_row((uint32_t)i, S, elogs[c])]));
            f.set(i, s);
        }
    });
    // circle fold on the canonic domain: twiddle i = 1 / y(half_odds(S-1).at(bitrev_{S-1}(i)))
    std::vector<uint32_t> src(n / 2), tw(n / 2);
    Coset ho = half_odds(S - 1);
    for (uint64_t i = 0; i < n / 2; i++) src[i] = ho.at(bitrev((uint32_t)i, S - 1)).y;
    batch_inverse(src.data(), tw.data(), n / 2);
    QM31 a0 = ch.draw_felt();
    std::vector<Soa4> layers(1);
    layers[0].resize(n / 2);
    fold_circle(layers[0], f, tw, a0);
    out.fri_roots.clear();
    auto commit = [&](const Soa4& L) {
        const uint32_t* lc[4] = {L.c[0].data(), L.c[1].data(), L.c[2].data(), L.c[3].data()};
        uint32_t r[8];
        std::vector<uint32_t> cur(8 * L.size()), nn;
        leaf_layer(lc, 4, L.size(), cur.data());
        while (cur.size() > 8) {
            nn.assign(cur.size() / 2, 0);
            node_layer(cur.data(), cur.size() / 16, nn.data());
            cur.swap(nn);
        }
        memcpy(r, cur.data(), 32);
        std::array<uint32_t, 8> a;
        memcpy(a.data(), r, 32);
        out.fri_roots.push_back(a);
        ch.mix_digest(r);
    };
    commit(layers[0]);
    uint32_t cur = S - 1;
    while (cur > log_last) {
        QM31 al = ch.draw_felt();
        std::vector<uint32_t> ltw;
        fold_twiddles(half_odds(cur), false, ltw);
        Soa4 nxt;
        fold_line(layers.back(), ltw, al, nxt);
        layers.push_back(std::move(nxt));
        cur--;
        commit(layers.back());
    }
    out.last.clear();
    const Soa4& L = layers.back();
    for (uint64_t i = 0; i < L.size(); i++)
        for (int c = 0; c < 4; c++) out.last.push_back(L.c[c][i]);
    if (lde_out) *lde_out = ev;
}
}  // namespace

// ================================================================ C ABI
extern "C" {

void orc_set_threads(uint32_t n) { g_threads.store(n); }
uint32_t orc_get_threads(void) { return n_threads(); }

uint32_t orc_m31_add(uint32_t a, uint32_t b) { return add(a, b); }
uint32_t orc_m31_sub(uint32_t a, uint32_t b) { return sub(a, b); }
uint32_t orc_m31_mul(uint32_t a, uint32_t b) { return mul(a, b); }
uint32_t orc_m31_inv(uint32_t a) { return inv(a); }
void orc_cm31_mul(const uint32_t a[2], const uint32_t b[2], uint32_t out[2]) {
    CM31 r = cmul({a[0], a[1]}, {b[0], b[1]});
    out[0] = r.a;
    out[1] = r.b;
}
void orc_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]) { qto(qmul(qfrom(a), qfrom(b)), out); }
void orc_qm31_inv(const uint32_t a[4], uint32_t out[4]) { qto(qinv(qfrom(a)), out); }

void orc_coset(int kind, uint32_t log_size, uint32_t out[4]) {
    Coset c = coset_of(kind, log_size);
    out[0] = c.initial.x;
    out[1] = c.initial.y;
    out[2] = c.step.x;
    out[3] = c.step.y;
}
void orc_coset_at(int kind, uint32_t log_size, uint64_t index, uint32_t out[2]) {
    Pt p = coset_of(kind, log_size).at(index);
    out[0] = p.x;
    out[1] = p.y;
}
void orc_circle_mul(const uint32_t a[2], const uint32_t b[2], uint32_t out[2]) {
    Pt r = pmul({a[0], a[1]}, {b[0], b[1]});
    out[0] = r.x;
    out[1] = r.y;
}
void orc_circle_repeated_double(const uint32_t a[2], uint32_t n, uint32_t out[2]) {
    Pt r = repeated_double({a[0], a[1]}, n);
    out[0] = r.x;
    out[1] = r.y;
}

void orc_ntt_twiddles(int kind, uint32_t log_n, uint32_t* line, uint32_t* circle, uint32_t* iline,
                      uint32_t* icircle, uint32_t* offsets, uint32_t* sizes) {
    NttTw tw;
    build_ntt_twiddles(coset_of(kind, log_n), tw, iline || icircle);
    if (line) memcpy(line, tw.line.data(), tw.line.size() * 4);
    if (circle) memcpy(circle, tw.circle.data(), tw.circle.size() * 4);
    if (iline) memcpy(iline, tw.iline.data(), tw.iline.size() * 4);
    if (icircle) memcpy(icircle, tw.icircle.data(), tw.icircle.size() * 4);
    if (offsets) memcpy(offsets, tw.offsets.data(), tw.offsets.size() * 4);
    if (sizes) memcpy(sizes, tw.sizes.data(), tw.sizes.size() * 4);
}
void orc_batch_inverse(const uint32_t* in, uint32_t* out, uint64_t n) { batch_inverse(in, out, n); }

void orc_ntt_layer(uint32_t* data, const uint32_t* twiddles, uint32_t layer_idx, uint64_t n, int forward) {
    ntt_layer(data, twiddles, layer_idx, n, forward != 0);
}
void orc_ntt_evaluate(uint32_t* data, const uint32_t* line, const uint32_t* circle, uint32_t log_n) {
    auto o = offsets_for(log_n);
    ntt_evaluate(data, line, circle, o.data(), log_n);
}
void orc_ntt_interpolate(uint32_t* data, const uint32_t* iline, const uint32_t* icircle, uint32_t log_n) {
    auto o = offsets_for(log_n);
    ntt_interpolate(data, iline, icircle, o.data(), log_n);
}
void orc_lde(const uint32_t* trace, uint32_t log_n, uint32_t log_blowup, uint32_t* eval) {
    uint32_t log_e = log_n + log_blowup;
    uint64_t n = 1ull << log_n, E = 1ull << log_e;
    NttTw ttw, etw;
    build_ntt_twiddles(half_coset(log_n), ttw, true);
    build_ntt_twiddles(log_e <= 30 ? half_coset(log_e) : subgroup(log_e), etw, false);
    memset(eval, 0, E * 4);
    memcpy(eval, trace, n * 4);
    ntt_interpolate(eval, ttw.iline.data(), ttw.icircle.data(), ttw.offsets.data(), log_n);
    ntt_evaluate(eval, etw.line.data(), etw.circle.data(), etw.offsets.data(), log_e);
}

void orc_blake2s(const uint8_t* msg, uint64_t len, uint8_t out[32]) { blake2s_full(msg, len, out); }
void orc_blake2s_single(const uint8_t* msg, uint64_t len, uint8_t out[32]) { blake2s_single(msg, len, out); }
void orc_blake2s_compress(uint32_t h[8], const uint32_t m[16], uint32_t t, uint32_t f) { compress(h, m, t, f); }

void orc_merkle_leaf_hashes(const uint32_t* const* cols, uint32_t n_cols, uint64_t n_leaves, uint32_t* out) {
    leaf_layer(cols, n_cols, n_leaves, out);
}
void orc_merkle_node_layer(const uint32_t* children, uint64_t n_parents, uint32_t* parents) {
    node_layer(children, n_parents, parents);
}
void orc_merkle_root(const uint32_t* const* cols, uint32_t n_cols, uint64_t n_leaves, uint32_t root[8]) {
    std::vector<uint32_t> cur(8 * n_leaves), nx;
    leaf_layer(cols, n_cols, n_leaves, cur.data());
    while (cur.size() > 8) {
        nx.assign(cur.size() / 2, 0);
        node_layer(cur.data(), cur.size() / 16, nx.data());
        cur.swap(nx);
    }
    memcpy(root, cur.data(), 32);
}
int orc_merkle_verify_path(const uint32_t root[8], const uint32_t leaf[8], uint64_t index, const uint32_t* path,
                           uint32_t depth) {
    std::vector<uint32_t> p(path, path + 8ull * depth);
    return verify_path(root, leaf, index, p) ? 1 : 0;
}

void orc_fibonacci_trace(uint32_t a, uint32_t b, uint32_t log_n, uint32_t* out) { fib_trace(a, b, log_n, out); }
void orc_fibonacci_quotient(const uint32_t* eval, uint64_t n, const uint32_t alpha[4], uint32_t* q0, uint32_t* q1,
                            uint32_t* q2, uint32_t* q3) {
    quotient(eval, n, qfrom(alpha), q0, q1, q2, q3);
}

void orc_fold_twiddles(int kind, uint32_t log_size, int extract_y, uint32_t* out) {
    std::vector<uint32_t> tw;
    fold_twiddles(coset_of(kind, log_size), extract_y != 0, tw);
    memcpy(out, tw.data(), tw.size() * 4);
}
void orc_fold_line(const uint32_t* const in[4], const uint32_t* tw, const uint32_t alpha[4], uint64_t half_n,
                   uint32_t* const out[4]) {
    QM31 al = qfrom(alpha);
    parallel_for(half_n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            QM31 f0 = {{in[0][2 * i], in[1][2 * i]}, {in[2][2 * i], in[3][2 * i]}};
            QM31 f1 = {{in[0][2 * i + 1], in[1][2 * i + 1]}, {in[2][2 * i + 1], in[3][2 * i + 1]}};
            QM31 r = fold_pair(f0, f1, al, tw[i]);
            out[0][i] = r.a.a;
            out[1][i] = r.a.b;
            out[2][i] = r.b.a;
            out[3][i] = r.b.b;
        }
    });
}
void orc_fold_circle_into_line(uint32_t* const dst[4], const uint32_t* const src[4], const uint32_t* tw,
                               const uint32_t alpha[4], uint64_t half_n) {
    QM31 al = qfrom(alpha), a2 = qmul(al, al);
    parallel_for(half_n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            QM31 f0 = {{src[0][2 * i], src[1][2 * i]}, {src[2][2 * i], src[3][2 * i]}};
            QM31 f1 = {{src[0][2 * i + 1], src[1][2 * i + 1]}, {src[2][2 * i + 1], src[3][2 * i + 1]}};
            QM31 prev = {{dst[0][i], dst[1][i]}, {dst[2][i], dst[3][i]}};
            QM31 r = qadd(qmul(prev, a2), fold_pair(f0, f1, al, tw[i]));
            dst[0][i] = r.a.a;
            dst[1][i] = r.a.b;
            dst[2][i] = r.b.a;
            dst[3][i] = r.b.b;
        }
    });
}

void orc_channel_init(orc_channel* ch) {
    memset(ch->state, 0, 32);
    ch->n_draws = 0;
}
static Channel to_ch(const orc_channel* c) {
    Channel ch;
    memcpy(ch.state, c->state, 32);
    ch.n_draws = c->n_draws;
    return ch;
}
static void from_ch(const Channel& ch, orc_channel* c) {
    memcpy(c->state, ch.state, 32);
    c->n_draws = ch.n_draws;
}
void orc_channel_mix_digest(orc_channel* c, const uint32_t digest[8]) {
    Channel ch = to_ch(c);
    ch.mix_digest(digest);
    from_ch(ch, c);
}
void orc_channel_mix_u64(orc_channel* c, uint64_t v) {
    Channel ch = to_ch(c);
    ch.mix_u64(v);
    from_ch(ch, c);
}
void orc_channel_mix_felts(orc_channel* c, const uint32_t* felts, uint64_t n_felts) {
    Channel ch = to_ch(c);
    ch.mix_felts(felts, n_felts);
    from_ch(ch, c);
}
void orc_channel_draw_felt(orc_channel* c, uint32_t out[4]) {
    Channel ch = to_ch(c);
    qto(ch.draw_felt(), out);
    from_ch(ch, c);
}
uint64_t orc_channel_draw_number(orc_channel* c, uint64_t bound) {
    Channel ch = to_ch(c);
    uint64_t r = ch.draw_number(bound);
    from_ch(ch, c);
    return r;
}

uint64_t orc_prove_fibonacci(uint32_t a, uint32_t b, uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer,
                             uint8_t* out, uint64_t cap, uint32_t* stats) {
    if (log_n < 1 || log_n + 2 > 30 || log_last_layer >= log_n + 2) return 0;
    Cache C;
    C.build(log_n, log_last_layer);
    std::vector<uint32_t> trace(1ull << log_n);
    fib_trace(a, b, log_n, trace.data());
    Proof pf;
    prove(a, b, C, n_queries, log_last_layer, pf, trace.data());
    std::vector<uint8_t> bytes = serialize(pf);
    if (out && cap >= bytes.size()) memcpy(out, bytes.data(), bytes.size());
    if (stats) {
        memcpy(stats, pf.trace_root, 32);
        memcpy(stats + 8, pf.quotient_root, 32);
        stats[16] = (uint32_t)pf.fri_roots.size();
    }
    return bytes.size();
}

int orc_prove_summary(uint32_t a, uint32_t b, uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer,
                      uint32_t* roots, uint32_t* last_layer, uint64_t* queries) {
    if (log_n < 1 || log_n + 2 > 31 || log_last_layer >= log_n + 2 || !roots || !last_layer || !queries) return -1;
    Proof pf;
    prove_summary(a % P, b % P, log_n, n_queries, log_last_layer, pf);
    memcpy(roots, pf.trace_root, 32);
    memcpy(roots + 8, pf.quotient_root, 32);
    for (size_t j = 0; j < pf.fri_roots.size(); j++) memcpy(roots + 16 + 8 * j, pf.fri_roots[j].data(), 32);
    memcpy(last_layer, pf.last_layer.data(), 4 * pf.last_layer.size());
    for (uint32_t i = 0; i < n_queries; i++) queries[i] = pf.queries[i];
    return 0;
}

void orc_ntt_twiddles_lean(int kind, uint32_t log_n, uint32_t* line, uint32_t* circle, uint32_t* iline,
                           uint32_t* icircle) {
    NttTw tw;
    build_ntt_twiddles_lean(coset_of(kind, log_n), tw, iline || icircle);
    if (line) memcpy(line, tw.line.data(), 4 * tw.line.size());
    if (circle) memcpy(circle, tw.circle.data(), 4 * tw.circle.size());
    if (iline) memcpy(iline, tw.iline.data(), 4 * tw.iline.size());
    if (icircle) memcpy(icircle, tw.icircle.data(), 4 * tw.icircle.size());
}

double orc_time_prove_fibonacci(uint32_t a, uint32_t b, uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer,
                                uint32_t reps) {
    if (log_n < 1 || log_n + 2 > 30) return -1.0;
    Cache C;
    C.build(log_n, log_last_layer);
    std::vector<uint32_t> trace(1ull << log_n);
    fib_trace(a, b, log_n, trace.data());
    Proof pf;
    auto t0 = std::chrono::steady_clock::now();
    for (uint32_t r = 0; r < reps; r++) prove(a, b, C, n_queries, log_last_layer, pf, trace.data());
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

int orc_time_prove_each(uint32_t a, uint32_t b, uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer,
                        uint32_t warmup, uint32_t reps, double* secs) {
    if (log_n < 1 || log_n + 2 > 30 || !secs) return -1;
    Cache C;
    C.build(log_n, log_last_layer);
    std::vector<uint32_t> trace(1ull << log_n);
    fib_trace(a, b, log_n, trace.data());
    Proof pf;
    for (uint32_t r = 0; r < warmup + reps; r++) {
        auto t0 = std::chrono::steady_clock::now();
        prove(a, b, C, n_queries, log_last_layer, pf, trace.data());
        auto t1 = std::chrono::steady_clock::now();
        if (r >= warmup) secs[r - warmup] = std::chrono::duration<double>(t1 - t0).count();
    }
    return 0;
}

int orc_verify_krak(const uint8_t* proof, uint64_t len, uint32_t n_queries, uint32_t log_last_layer) {
    Proof pf;
    if (!deserialize(proof, len, pf)) return 1;
    return verify(pf, n_queries, log_last_layer);
}


void orc_stwo_twiddles(uint32_t log_n, uint32_t* fwd, uint32_t* inv_out) {
    std::vector<uint32_t> f, i;
    stwo_twiddles(log_n, f, i);
    if (fwd) memcpy(fwd, f.data(), f.size() * 4);
    if (inv_out) memcpy(inv_out, i.data(), i.size() * 4);
}
void orc_stwo_ntt_evaluate(uint32_t* data, const uint32_t* twiddles, uint32_t log_n) {
    stwo_ntt(data, twiddles, log_n, true);
}
void orc_stwo_ntt_interpolate(uint32_t* data, const uint32_t* itwiddles, uint32_t log_n) {
    stwo_ntt(data, itwiddles, log_n, false);
}
void orc_lifted_leaves(const uint32_t* const* cols, const uint32_t* log_sizes, uint32_t n_cols, uint32_t lifting_log,
                       uint32_t* out) {
    lifted_leaves(cols, log_sizes, n_cols, lifting_log, out);
}
void orc_pack_leaves(const uint32_t* const* in4, uint32_t* const* out64, uint64_t n) {
    // cuda/pack_leaves.cu:22-38: out[c*16 + p][i] = in[c][16 i + p]
    for (uint32_t c = 0; c < 4; c++)
        for (uint64_t j = 0; j < n; j++) out64[c * 16 + (j % 16)][j / 16] = in4[c][j];
}
void orc_bit_reverse_qm31(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t log_n) {
    // cuda/bit_reverse_wide.cu:21-40
    for (uint64_t i = 0; i < n; i++) {
        uint64_t j = bitrev((uint32_t)i, log_n);
        for (int k = 0; k < 4; k++) out[4 * j + k] = in[4 * i + k];
    }
}
uint32_t orc_wide_commit(const uint32_t* const* trace, const uint32_t* log_sizes, uint32_t n_cols, uint32_t blowup,
                         uint32_t log_last, uint32_t* root8, uint32_t* fri_roots, uint32_t* last, uint32_t* beta4) {
    // columns of one lifted tree come in ascending log size (merkle_ops.rs:49)
    for (uint32_t c = 1; c < n_cols; c++)
        if (log_sizes[c] < log_sizes[c - 1]) return 0;
    std::vector<std::vector<uint32_t>> tr(n_cols);
    std::vector<uint32_t> logs(log_sizes, log_sizes + n_cols);
    for (uint32_t c = 0; c < n_cols; c++) tr[c].assign(trace[c], trace[c] + (1ull << logs[c]));
    WideOut o;
    wide_commit(tr, logs, blowup, log_last, o, nullptr);
    if (root8) memcpy(root8, o.root, 32);
    if (fri_roots)
        for (size_t j = 0; j < o.fri_roots.size(); j++) memcpy(fri_roots + 8 * j, o.fri_roots[j].data(), 32);
    if (last) memcpy(last, o.last.data(), o.last.size() * 4);
    if (beta4) memcpy(beta4, o.beta, 16);
    return (uint32_t)o.fri_roots.size();
}
double orc_time_wide_commit(const uint32_t* const* trace, const uint32_t* log_sizes, uint32_t n_cols, uint32_t blowup,
                            uint32_t log_last) {
    std::vector<std::vector<uint32_t>> tr(n_cols);
    std::vector<uint32_t> logs(log_sizes, log_sizes + n_cols);
    for (uint32_t c = 0; c < n_cols; c++) tr[c].assign(trace[c], trace[c] + (1ull << logs[c]));
    WideOut o;
    auto t0 = std::chrono::steady_clock::now();
    wide_commit(tr, logs, blowup, log_last, o, nullptr);
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

uint32_t orc_lifted_row(uint32_t leaf, uint32_t lifting_log, uint32_t col_log) {
    // cuda/merkle_leaves_lifted.cu:26-31
    if (col_log == lifting_log) return leaf;
    uint32_t shift = lifting_log - col_log + 1;
    return ((leaf >> shift) << 1) | (leaf & 1);
}

uint64_t orc_splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// ---- Blake2s proof-of-work grinding (src/channel.rs:140-168, cuda/grind.cu:13-70,
// cuda/grind_m31_output.cu:14-68, crates/cuda-backend/src/grind_ops.rs:20-34, 146-161)

// m31_reduce of a 32-bit word, step for step (cuda/include/m31.cuh:9-14): x mod p.
static uint32_t m31_reduce64(uint64_t x) {
    const uint32_t lo = (uint32_t)(x & P), hi = (uint32_t)(x >> 31);
    const uint32_t r = lo + hi;
    return r >= P ? r - P : r;
}

uint32_t orc_pow_reduce_word(uint32_t x) { return m31_reduce64(x); }

// prefixed_digest = Blake2s(0x12345678 LE || 12 zero bytes || state || pow_bits LE), 52 bytes
// (channel.rs:150-155). With m31_output every word of the digest is reduced mod p
// (the reduction grind_ops.rs:161 applies before the search).
void orc_pow_prefix(const uint32_t state[8], uint32_t pow_bits, int m31_output, uint32_t out[8]) {
    uint8_t in[52] = {0};
    const uint32_t prefix = 0x12345678u;
    memcpy(in, &prefix, 4);
    memcpy(in + 16, state, 32);
    memcpy(in + 48, &pow_bits, 4);
    uint8_t h[32];
    blake2s_full(in, 52, h);
    memcpy(out, h, 32);
    if (m31_output)
        for (int i = 0; i < 8; i++) out[i] = m31_reduce64(out[i]);
}

// Trailing zeros of the first 16 bytes (u128 LE) of Blake2s(prefixed_digest || nonce LE),
// 40 bytes (channel.rs:157-167, grind.cu:56-69); with m31_output the four words are
// reduced mod p first (grind_m31_output.cu:48-66).
uint32_t orc_pow_zeros(const uint32_t prefixed[8], uint64_t nonce, int m31_output) {
    uint8_t in[40];
    memcpy(in, prefixed, 32);
    memcpy(in + 32, &nonce, 8);
    uint8_t h[32];
    blake2s_full(in, 40, h);
    uint32_t w[4];
    memcpy(w, h, 16);
    for (int i = 0; i < 4; i++) {
        if (m31_output) w[i] = m31_reduce64(w[i]);
        if (w[i]) return 32u * i + (uint32_t)__builtin_ctz(w[i]);
    }
    return 128;
}

// Smallest nonce in [start, start + count) whose hash has >= pow_bits trailing zeros, or
// UINT64_MAX. Threads split the range; the minimum over them is the answer.
uint64_t orc_pow_search(const uint32_t prefixed[8], uint32_t pow_bits, int m31_output, uint64_t start,
                        uint64_t count) {
    std::atomic<uint64_t> best{UINT64_MAX};
    parallel_for(count, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            if (start + i >= best.load(std::memory_order_relaxed)) return;
            if (orc_pow_zeros(prefixed, start + i, m31_output) >= pow_bits) {
                uint64_t cur = best.load();
                while (start + i < cur && !best.compare_exchange_weak(cur, start + i)) {
                }
                return;
            }
        }
    });
    return best.load();
}

// ---- stwo DEEP quotient pieces (cuda/quotient.cu:25-253, cuda/accumulate_lift.cu:13-51;
// callers crates/cuda-backend/src/quotient_ops.rs:35-223, accumulation_ops.rs:46-88)

// result[row] = sum_i (c_i * col[col_indices[i]][row] - b_i)   (quotient.cu:25-80)
void orc_accumulate_numerators(const uint32_t* const* cols, const uint32_t* col_indices, const uint32_t* b,
                               const uint32_t* c, uint32_t n_batch, uint32_t n_rows, uint32_t* out0,
                               uint32_t* out1, uint32_t* out2, uint32_t* out3) {
    parallel_for(n_rows, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t row = lo; row < hi; row++) {
            QM31 acc = {{0, 0}, {0, 0}};
            for (uint32_t i = 0; i < n_batch; i++) {
                const uint32_t f = cols[col_indices[i]][row];
                acc = qadd(acc, qsub(qmul_m31(qfrom(c + 4 * i), f), qfrom(b + 4 * i)));
            }
            uint32_t v[4];
            qto(acc, v);
            out0[row] = v[0];
            out1[row] = v[1];
            out2[row] = v[2];
            out3[row] = v[3];
        }
    });
}

// quotient[row] = sum_j (numer_j[lifted(row)] - a_j * dy) * den_j^-1, with the CM31
// den_j = (Re(px) - dx) * Im(py) - (Re(py) - dy) * Im(px)          (quotient.cu:82-205)
void orc_quotients_combine(const uint32_t* spx, const uint32_t* spy, const uint32_t* fla,
                           const uint32_t* const* n0, const uint32_t* const* n1, const uint32_t* const* n2,
                           const uint32_t* const* n3, const uint32_t* acc_logs, uint32_t n_accs,
                           const uint32_t* dxs, const uint32_t* dys, uint32_t lifting_log, uint32_t n_rows,
                           uint32_t* out0, uint32_t* out1, uint32_t* out2, uint32_t* out3) {
    parallel_for(n_rows, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t row = lo; row < hi; row++) {
            const uint32_t dx = dxs[row], dy = dys[row];
            QM31 q = {{0, 0}, {0, 0}};
            for (uint32_t j = 0; j < n_accs; j++) {
                const QM31 px = qfrom(spx + 4 * j), py = qfrom(spy + 4 * j);
                const CM31 t1 = cmul(csub(px.a, {dx, 0}), py.b);
                const CM31 t2 = cmul(csub(py.a, {dy, 0}), px.b);
                const CM31 dinv = cinv(csub(t1, t2));
                const uint32_t r = lifting_log - acc_logs[j];
                const uint64_t li = ((row >> (r + 1)) << 1) + (row & 1);
                const uint32_t nv[4] = {n0[j][li], n1[j][li], n2[j][li], n3[j][li]};
                const QM31 full = qsub(qfrom(nv), qmul_m31(qfrom(fla + 4 * j), dy));
                q = qadd(q, {cmul(full.a, dinv), cmul(full.b, dinv)});
            }
            uint32_t v[4];
            qto(q, v);
            out0[row] = v[0];
            out1[row] = v[1];
            out2[row] = v[2];
            out3[row] = v[3];
        }
    });
}

// col[i] += curr[((i >> (log_ratio + 1)) << 1) | (i & 1)], QM31 elements of 4 words
// (accumulate_lift.cu:15-40)
void orc_accumulate_lift(uint32_t* col, const uint32_t* curr, uint32_t col_n, uint32_t log_ratio) {
    for (uint64_t i = 0; i < col_n; i++) {
        const uint64_t s = ((i >> (log_ratio + 1)) << 1) | (i & 1);
        const QM31 r = qadd(qfrom(col + 4 * i), qfrom(curr + 4 * s));
        qto(r, col + 4 * i);
    }
}

// Points of CanonicCoset(log).circle_domain() at bit_reverse_index(row) for every row
// (quotient_ops.rs:140-147): natural index k < n/2 is half_odds(log - 1).at(k), the
// rest are the conjugates (x, -y) of the first half.
void orc_circle_domain_brev(uint32_t log_n, uint32_t* xs, uint32_t* ys) {
    const uint64_t n = 1ull << log_n, half = n / 2;
    std::vector<uint32_t> hx, hy;
    all_points(half_odds(log_n - 1), hx, hy);
    for (uint64_t row = 0; row < n; row++) {
        const uint64_t k = bitrev((uint32_t)row, log_n);
        xs[row] = hx[k < half ? k : k - half];
        ys[row] = k < half ? hy[k] : neg(hy[k - half]);
    }
}

}  // extern "C"
