// Wave64 Blake2s Merkle engine with fused leaf producers.
//
// Tree semantics are the reference's (src/merkle.rs:28-104, cuda/blake2s.cu):
// complete binary tree, leaf = Blake2s(LE32 words), parent = Blake2s(L || R),
// no personalisation. What differs is the execution model:
//
//  * A "source" produces the items of one tree level. Leaf sources hash data
//    (one M31 column, four SoA columns, the Fibonacci quotient computed on the
//    fly from the LDE, or an FRI circle/line fold computed and stored on the
//    fly); the node source reads 8-word hashes produced by a previous kernel.
//  * Stage kernel: every wave owns 128 * 2^logc consecutive items. Each step
//    the 64 lanes compute 64 parents of 128 items (one chunk); chunks merge
//    pairwise like a binary counter, the left chunk waiting in the wave's LDS
//    stack, so every compression runs with all 64 lanes busy and no
//    workgroup barrier is ever needed. The last 64 nodes are reduced to one
//    with quad-cooperative compressions (4 lanes each), so the wave emits one
//    node, 7 + logc levels above its items.
//  * Tail kernel: one workgroup reduces the last <= 2048 items to the root,
//    writes it where the caller asks and advances the Fiat-Shamir transcript
//    (mix the root, draw the next challenge) in the same launch; its narrow
//    levels and the transcript hashes also run quad-cooperatively.
//  * Every node at level >= store_from is also written to a per-tree level
//    store; decommitment recomputes the 2^store_from-leaf subtree of a query
//    and reads the rest, so full trees are never written to HBM.
#pragma once
#include <type_traits>

#include "stage_plan.hpp"
#include "blake2s.hpp"
#include "field.hpp"
#include "transcript.hpp"

namespace vx {

// Orders this wave's LDS writes before its later LDS reads (the LDS executes a
// wave's requests in order; this keeps the compiler from reordering them).
#define VX_WAVE_SYNC()                                        \
    do {                                                      \
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); \
        __builtin_amdgcn_wave_barrier();                      \
    } while (0)

constexpr uint32_t MERKLE_TAIL_MAX = 2048;  // most items one tail-kernel block reduces
constexpr uint32_t MERKLE_STAGE_WAVES = 4;  // waves per stage workgroup
constexpr uint32_t MERKLE_STAGE_KEEP = 16;  // nodes a wave of a deep leaf stage hands on
constexpr uint32_t MERKLE_KEEP_LOGC = 4;    // chunk depth from which leaf stages do so

// Nodes of levels [from, depth] of a tree with 2^depth leaves, level-major.
struct LevelStore {
    uint32_t* base;
    uint32_t depth;
    uint32_t from;
    __device__ __forceinline__ bool want(uint32_t level) const { return base && level >= from && level <= depth; }
    __device__ __forceinline__ uint32_t* at(uint32_t level, uint64_t idx) const {
        uint64_t off = (1ull << (depth - from + 1)) - (1ull << (depth - level + 1));
        return base + 8ull * (off + idx);
    }
    __device__ __forceinline__ LevelStore shifted(uint64_t off) const {
        return {base ? base + off : nullptr, depth, from};
    }
};
__host__ __device__ inline uint64_t level_store_words(uint32_t depth, uint32_t from) {
    return from > depth ? 0 : 8ull * ((1ull << (depth - from + 1)) - 1);
}

__device__ __forceinline__ void store8(uint32_t* dst, const uint32_t h[8]) {
    uint4* d = (uint4*)dst;
    d[0] = make_uint4(h[0], h[1], h[2], h[3]);
    d[1] = make_uint4(h[4], h[5], h[6], h[7]);
}
__device__ __forceinline__ void load8(const uint32_t* src, uint32_t h[8]) {
    const uint4* s = (const uint4*)src;
    uint4 a = s[0], b = s[1];
    h[0] = a.x;
    h[1] = a.y;
    h[2] = a.z;
    h[3] = a.w;
    h[4] = b.x;
    h[5] = b.y;
    h[6] = b.z;
    h[7] = b.w;
}

// ---------------------------------------------------------------- sources
// pair(i, h): i even; h = parent of items i and i+1.
// leaf(i, h) / values(i, v): used by decommitment (leaf sources only).
// at(off): the same source for another proof of a batch, whose buffers sit
// `off` words further (batched provers give every proof an equal-size arena;
// shared tables such as twiddles do not move).

struct SrcNodes {
    const uint32_t* nodes;
    __device__ __forceinline__ SrcNodes at(uint64_t off) const { return {nodes + off}; }
    __device__ __forceinline__ void pair(uint64_t i, uint32_t h[8]) const {
        uint32_t l[8], r[8];
        load8(nodes + 8 * i, l);
        load8(nodes + 8 * i + 8, r);
        b2s_node(l, r, h);
    }
};

// One M31 column (the trace LDE); leaf = H(LE32(v)), t = 4.
struct SrcCol1 {
    const uint32_t* col;
    static constexpr uint32_t WIDTH = 1;
    __device__ __forceinline__ SrcCol1 at(uint64_t off) const { return {col + off}; }
    __device__ __forceinline__ void values(uint64_t i, uint32_t v[4]) const { v[0] = col[i]; }
    __device__ __forceinline__ void leaf(uint64_t i, uint32_t h[8]) const { b2s_leaf1(col[i], h); }
    __device__ __forceinline__ void pair(uint64_t i, uint32_t h[8]) const {
        uint2 v = *(const uint2*)(col + i);
        uint32_t l[8], r[8];
        b2s_leaf1(v.x, l);
        b2s_leaf1(v.y, r);
        b2s_node(l, r, h);
    }
};

// Four SoA columns (a QM31 column); leaf = H(16 bytes), t = 16.
struct SrcSoa4 {
    const uint32_t* c0;
    const uint32_t* c1;
    const uint32_t* c2;
    const uint32_t* c3;
    static constexpr uint32_t WIDTH = 4;
    __device__ __forceinline__ SrcSoa4 at(uint64_t off) const { return {c0 + off, c1 + off, c2 + off, c3 + off}; }
    __device__ __forceinline__ void values(uint64_t i, uint32_t v[4]) const {
        v[0] = c0[i];
        v[1] = c1[i];
        v[2] = c2[i];
        v[3] = c3[i];
    }
    __device__ __forceinline__ void leaf(uint64_t i, uint32_t h[8]) const {
        b2s_leaf4(c0[i], c1[i], c2[i], c3[i], h);
    }
    __device__ __forceinline__ void pair(uint64_t i, uint32_t h[8]) const {
        uint2 a = *(const uint2*)(c0 + i), b = *(const uint2*)(c1 + i);
        uint2 c = *(const uint2*)(c2 + i), d = *(const uint2*)(c3 + i);
        uint32_t l[8], r[8];
        b2s_leaf4(a.x, b.x, c.x, d.x, l);
        b2s_leaf4(a.y, b.y, c.y, d.y, r);
        b2s_node(l, r, h);
    }
};

// Fibonacci quotient evaluated on the fly from the LDE t (size mask+1):
// q[i] = alpha * (t[i+2] - (t[i+1] + t[i])), indices mod E (constraint.cu:10-36).
__device__ __forceinline__ QM31 quotient_at(uint32_t t0, uint32_t t1, uint32_t t2, QM31 alpha) {
    return qmul_m31(alpha, msub(t2, madd(t1, t0)));
}
__device__ __forceinline__ QM31 load_alpha(const uint32_t* a) { return {{a[0], a[1], a[2], a[3]}}; }

struct SrcQuotient {
    const uint32_t* t;
    uint64_t mask;
    const uint32_t* alpha;  // device, 4 words
    static constexpr uint32_t WIDTH = 4;
    __device__ __forceinline__ SrcQuotient at(uint64_t off) const { return {t + off, mask, alpha + off}; }
    __device__ __forceinline__ QM31 q(uint64_t i, QM31 al) const {
        return quotient_at(t[i], t[(i + 1) & mask], t[(i + 2) & mask], al);
    }
    __device__ __forceinline__ void values(uint64_t i, uint32_t v[4]) const {
        QM31 r = q(i, load_alpha(alpha));
        for (int k = 0; k < 4; k++) v[k] = r.v[k];
    }
    __device__ __forceinline__ void leaf(uint64_t i, uint32_t h[8]) const {
        QM31 r = q(i, load_alpha(alpha));
        b2s_leaf4(r.v[0], r.v[1], r.v[2], r.v[3], h);
    }
    __device__ __forceinline__ void pair(uint64_t i, uint32_t h[8]) const {
        QM31 al = load_alpha(alpha);
        uint2 a = *(const uint2*)(t + i);
        uint2 b = *(const uint2*)(t + ((i + 2) & mask));
        QM31 q0 = quotient_at(a.x, a.y, b.x, al);
        QM31 q1 = quotient_at(a.y, b.x, b.y, al);
        uint32_t l[8], r[8];
        b2s_leaf4(q0.v[0], q0.v[1], q0.v[2], q0.v[3], l);
        b2s_leaf4(q1.v[0], q1.v[1], q1.v[2], q1.v[3], r);
        b2s_node(l, r, h);
    }
};

// First FRI layer: circle fold of the quotient (recomputed from the LDE),
// stored to out and hashed (fri.cu:44-78 with a zero-initialised destination,
// so the layer equals the fold f' itself).
struct SrcCircleFold {
    const uint32_t* t;
    uint64_t mask;
    const uint32_t* alpha_q;
    const uint32_t* alpha_f;
    const uint32_t* tw;
    uint32_t* o0;
    uint32_t* o1;
    uint32_t* o2;
    uint32_t* o3;
    __device__ __forceinline__ SrcCircleFold at(uint64_t off) const {
        return {t + off, mask, alpha_q + off, alpha_f + off, tw, o0 + off, o1 + off, o2 + off, o3 + off};
    }
    // With q(i) = aq * d(i), d(i) = t[i+2] - t[i+1] - t[i] an M31 value, the fold
    // (q0 + q1) + af * ((q0 - q1) * tw) equals aq * (d0 + d1) + (af * aq) * ((d0 - d1) * tw):
    // nine M31 products per output instead of twenty-one.
    __device__ __forceinline__ static QM31 fold_d(uint32_t d0, uint32_t d1, QM31 aq, QM31 g, uint32_t tw) {
        const uint32_t s = madd(d0, d1), u = mmul(msub(d0, d1), tw);
        QM31 r;
#pragma unroll
        for (int k = 0; k < 4; k++) r.v[k] = madd(mmul(aq.v[k], s), mmul(g.v[k], u));
        return r;
    }
    __device__ __forceinline__ void pair(uint64_t i, uint32_t h[8]) const {
        const QM31 aq = load_alpha(alpha_q), g = qmul(load_alpha(alpha_f), aq);
        uint64_t j = 2 * i;  // quotient index of the first of four inputs
        uint2 a = *(const uint2*)(t + j);
        uint2 b = *(const uint2*)(t + ((j + 2) & mask));
        uint2 c = *(const uint2*)(t + ((j + 4) & mask));
        const uint32_t d0 = msub(b.x, madd(a.y, a.x)), d1 = msub(b.y, madd(b.x, a.y));
        const uint32_t d2 = msub(c.x, madd(b.y, b.x)), d3 = msub(c.y, madd(c.x, b.y));
        uint2 w = *(const uint2*)(tw + i);
        QM31 f0 = fold_d(d0, d1, aq, g, w.x);
        QM31 f1 = fold_d(d2, d3, aq, g, w.y);
        *(uint2*)(o0 + i) = make_uint2(f0.v[0], f1.v[0]);
        *(uint2*)(o1 + i) = make_uint2(f0.v[1], f1.v[1]);
        *(uint2*)(o2 + i) = make_uint2(f0.v[2], f1.v[2]);
        *(uint2*)(o3 + i) = make_uint2(f0.v[3], f1.v[3]);
        uint32_t l[8], r[8];
        b2s_leaf4(f0.v[0], f0.v[1], f0.v[2], f0.v[3], l);
        b2s_leaf4(f1.v[0], f1.v[1], f1.v[2], f1.v[3], r);
        b2s_node(l, r, h);
    }
};

// Line fold of the previous layer, stored to out and hashed (fri.cu:10-39).
struct SrcLineFold {
    const uint32_t* i0;
    const uint32_t* i1;
    const uint32_t* i2;
    const uint32_t* i3;
    const uint32_t* alpha;
    const uint32_t* tw;
    uint32_t* o0;
    uint32_t* o1;
    uint32_t* o2;
    uint32_t* o3;
    __device__ __forceinline__ SrcLineFold at(uint64_t off) const {
        return {i0 + off, i1 + off, i2 + off, i3 + off, alpha + off, tw, o0 + off, o1 + off, o2 + off, o3 + off};
    }
    __device__ __forceinline__ void pair(uint64_t i, uint32_t h[8]) const {
        QM31 al = load_alpha(alpha);
        uint64_t j = 2 * i;
        uint4 a = *(const uint4*)(i0 + j), b = *(const uint4*)(i1 + j);
        uint4 c = *(const uint4*)(i2 + j), d = *(const uint4*)(i3 + j);
        uint2 w = *(const uint2*)(tw + i);
        QM31 f0 = fold_pair({{a.x, b.x, c.x, d.x}}, {{a.y, b.y, c.y, d.y}}, al, w.x);
        QM31 f1 = fold_pair({{a.z, b.z, c.z, d.z}}, {{a.w, b.w, c.w, d.w}}, al, w.y);
        *(uint2*)(o0 + i) = make_uint2(f0.v[0], f1.v[0]);
        *(uint2*)(o1 + i) = make_uint2(f0.v[1], f1.v[1]);
        *(uint2*)(o2 + i) = make_uint2(f0.v[2], f1.v[2]);
        *(uint2*)(o3 + i) = make_uint2(f0.v[3], f1.v[3]);
        uint32_t l[8], r[8];
        b2s_leaf4(f0.v[0], f0.v[1], f0.v[2], f0.v[3], l);
        b2s_leaf4(f1.v[0], f1.v[1], f1.v[2], f1.v[3], r);
        b2s_node(l, r, h);
    }
};

// Lifted multi-column leaves (cuda/merkle_leaves_lifted.cu:26-98): leaf i hashes,
// with standard streaming Blake2s (full blocks non-final, the last block final,
// t = bytes so far), the LE32 words of every column in the given order, column
// c read at row i when its log size equals the tree's, else at
// ((i >> sh) << 1) | (i & 1) with sh = S - log_size + 1.
struct LiftCol {
    const uint32_t* p;
    uint32_t sh;  // 0: full-size column
    uint32_t pad;
};
// The column table is read-only for the whole launch and every lane reads the
// same entry: reading it through the constant address space lets the compiler
// use scalar loads, and the column data through the global one plain vector
// loads (a generic pointer would make them flat loads).
typedef const __attribute__((address_space(4))) LiftCol ConstLiftCol;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) u32x2 GlobalV2;
typedef const __attribute__((address_space(1))) u32x4 GlobalV4;
__device__ __forceinline__ ConstLiftCol* const_table(const LiftCol* t) { return (ConstLiftCol*)t; }

struct SrcLifted {
    static constexpr uint32_t STAGE_LOGC_MAX = 4;
    const LiftCol* cols;  // device, hashing order
    uint32_t n_cols;
    __device__ __forceinline__ SrcLifted at(uint64_t) const { return *this; }
    // leaves i (even) and i + 1 read rows r and r + 1 of every column
    __device__ __forceinline__ void load_block(ConstLiftCol* tab, uint32_t c0, uint64_t i, uint2 v[16]) const {
        const uint32_t nw = n_cols - c0 < 16 ? n_cols - c0 : 16;
#pragma unroll
        for (uint32_t k = 0; k < 16; k++) {
            v[k] = make_uint2(0, 0);
            if (k < nw) {
                const uint32_t* p = tab[c0 + k].p;
                const uint32_t sh = tab[c0 + k].sh;
                const uint64_t r = sh ? ((i >> sh) << 1) : i;
                const u32x2 x = *(GlobalV2*)(p + r);
                v[k] = make_uint2(x.x, x.y);
            }
        }
    }
    __device__ __forceinline__ void pair(uint64_t i, uint32_t h[8]) const {
        ConstLiftCol* tab = const_table(cols);
        uint32_t ha[8], hb[8];
        b2s_init(ha);
        b2s_init(hb);
        uint2 cur[16];
        load_block(tab, 0, i, cur);
        for (uint32_t c0 = 0;; c0 += 16) {
            // the next block's loads are in flight while this block is hashed
            const bool last = c0 + 16 >= n_cols;
            uint2 nxt[16];
            if (!last) load_block(tab, c0 + 16, i, nxt);
            uint32_t ma[16], mb[16];
#pragma unroll
            for (uint32_t k = 0; k < 16; k++) {
                ma[k] = cur[k].x;
                mb[k] = cur[k].y;
            }
            const uint32_t nw = n_cols - c0 < 16 ? n_cols - c0 : 16;
            const uint32_t t = 4 * (c0 + nw), f = last ? 0xFFFFFFFFu : 0u;
            b2s_compress(ha, ma, t, f);
            b2s_compress(hb, mb, t, f);
            if (last) break;
#pragma unroll
            for (uint32_t k = 0; k < 16; k++) cur[k] = nxt[k];
        }
        b2s_node(ha, hb, h);
    }
};

// x < 2^64 -> a value < 2^33 congruent to it mod p: bits 0..30, plus bits
// 31..62 (times 2^31 = 1), plus bit 63 (times 2^63 = 2 mod p).
__device__ __forceinline__ uint64_t fold62(uint64_t x) {
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    return (uint64_t)(lo & P) + __builtin_amdgcn_alignbit(hi, lo, 31) + 2u * (hi >> 31);
}
// x < 2^64 -> x mod p
__device__ __forceinline__ uint32_t reduce64(uint64_t x) {
    x = fold62(fold62(x));  // < 2^31 + 4
    return canon((uint32_t)x);
}

// First FRI layer of the wide commitment: f = sum_c beta^c * lifted_c is formed
// on the fly from the LDE columns, circle-folded pairwise (2j, 2j + 1) with
// alpha and tw[j], stored to out and hashed as an SoA4 leaf.
struct SrcWideFold {
    static constexpr uint32_t STAGE_LOGC_MAX = 4;
    const LiftCol* cols;
    uint32_t n_cols;
    const uint32_t* beta_pows;  // device, 4 words per column
    const uint32_t* alpha;
    const uint32_t* tw;
    uint32_t* o0;
    uint32_t* o1;
    uint32_t* o2;
    uint32_t* o3;
    __device__ __forceinline__ SrcWideFold at(uint64_t) const { return *this; }
    // Column c's values at f indices j .. j + 3 (j a multiple of 4): a full-size
    // column gives four rows, a lifted one rows r, r + 1 (repeated for j + 2, j + 3).
    __device__ __forceinline__ u32x4 load_col(ConstLiftCol* tab, uint32_t c, uint64_t j) const {
        const uint32_t* p = tab[c].p;
        const uint32_t sh = tab[c].sh;
        if (sh == 0) return *(GlobalV4*)(p + j);
        const u32x2 v = *(GlobalV2*)(p + ((j >> sh) << 1));
        return u32x4{v.x, v.y, 0u, 0u};
    }
    __device__ __forceinline__ void pair(uint64_t i, uint32_t h[8]) const {
        const uint64_t j = 2 * i;  // f index of the first of four inputs (a multiple of 4)
        ConstLiftCol* tab = const_table(cols);
        typedef const __attribute__((address_space(4))) uint32_t ConstU32;
        ConstU32* bp = (ConstU32*)beta_pows;
        // sum_c beta^c * v_c with 64-bit products accumulated unreduced: three
        // products (< 2^62 each) fit on top of a folded (< 2^33) accumulator,
        // so accumulators are folded after every group of three columns.
        // Full-size columns feed af (four points), lifted ones ag (two).
        uint64_t af[4][4], ag[2][4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            for (int k = 0; k < 4; k++) af[k][m] = 0;
            ag[0][m] = ag[1][m] = 0;
        }
        constexpr uint32_t G = 3;
        u32x4 cur[G], nxt[G];
#pragma unroll
        for (uint32_t k = 0; k < G; k++)
            if (k < n_cols) cur[k] = load_col(tab, k, j);
        for (uint32_t c0 = 0; c0 < n_cols; c0 += G) {
            // the next group's loads are in flight while this group accumulates
#pragma unroll
            for (uint32_t k = 0; k < G; k++)
                if (c0 + G + k < n_cols) nxt[k] = load_col(tab, c0 + G + k, j);
            bool tf = false, tl = false;
#pragma unroll
            for (uint32_t k = 0; k < G; k++) {
                const uint32_t c = c0 + k;
                if (c >= n_cols) break;
                const uint32_t w[4] = {bp[4 * c], bp[4 * c + 1], bp[4 * c + 2], bp[4 * c + 3]};
                const u32x4 v = cur[k];
                if (tab[c].sh == 0) {
                    tf = true;
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        af[0][m] += (uint64_t)w[m] * v.x;
                        af[1][m] += (uint64_t)w[m] * v.y;
                        af[2][m] += (uint64_t)w[m] * v.z;
                        af[3][m] += (uint64_t)w[m] * v.w;
                    }
                } else {
                    tl = true;
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        ag[0][m] += (uint64_t)w[m] * v.x;
                        ag[1][m] += (uint64_t)w[m] * v.y;
                    }
                }
            }
            if (tf)
#pragma unroll
                for (int m = 0; m < 4; m++)
                    for (int k = 0; k < 4; k++) af[k][m] = fold62(af[k][m]);
            if (tl)
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    ag[0][m] = fold62(ag[0][m]);
                    ag[1][m] = fold62(ag[1][m]);
                }
#pragma unroll
            for (uint32_t k = 0; k < G; k++) cur[k] = nxt[k];
        }
        QM31 f[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
            for (int m = 0; m < 4; m++) f[k].v[m] = reduce64(af[k][m] + ag[k & 1][m]);
        const QM31 al = load_alpha(alpha);
        const uint2 w = *(const uint2*)(tw + i);
        const QM31 g0 = fold_pair(f[0], f[1], al, w.x);
        const QM31 g1 = fold_pair(f[2], f[3], al, w.y);
        *(uint2*)(o0 + i) = make_uint2(g0.v[0], g1.v[0]);
        *(uint2*)(o1 + i) = make_uint2(g0.v[1], g1.v[1]);
        *(uint2*)(o2 + i) = make_uint2(g0.v[2], g1.v[2]);
        *(uint2*)(o3 + i) = make_uint2(g0.v[3], g1.v[3]);
        uint32_t l[8], r[8];
        b2s_leaf4(g0.v[0], g0.v[1], g0.v[2], g0.v[3], l);
        b2s_leaf4(g1.v[0], g1.v[1], g1.v[2], g1.v[3], r);
        b2s_node(l, r, h);
    }
};

// ---------------------------------------------------------------- narrow levels
// Reduces n (a power of two <= 64) nodes stored node-major ([node][8 words]) in
// the LDS buffer `a` to `stop` nodes (a power of two, default one) with the
// quads of the calling wave, ping-ponging with `b`. The nodes are at tree level
// `lvl`, node 0 having index idx0 there. Every lane of the wave must call this;
// returns the buffer holding the remaining nodes.
__device__ __forceinline__ uint32_t* narrow_reduce(uint32_t* a, uint32_t* b, uint32_t n, uint32_t lvl, uint64_t idx0,
                                                   const LevelStore& st, uint32_t stop = 1) {
    const uint32_t lane = threadIdx.x & 63, quad = lane >> 2, l = lane & 3;
    while (n > stop) {
        const uint32_t half = n / 2;
        for (uint32_t q0 = 0; q0 < half; q0 += 16) {
            const uint32_t q = q0 + quad;
            const uint32_t qq = q < half ? q : half - 1;  // whole quads stay active for the DPP permutes
            uint32_t hl, hh;
            b2s_quad_init(hl, hh);
            b2s_compress_quad(a + 16 * qq, 64u, 0xFFFFFFFFu, hl, hh);
            if (q < half) {
                b[8 * q + l] = hl;
                b[8 * q + 4 + l] = hh;
                if (st.want(lvl + 1)) {
                    uint32_t* dst = st.at(lvl + 1, (idx0 >> 1) + q);
                    dst[l] = hl;
                    dst[4 + l] = hh;
                }
            }
        }
        VX_WAVE_SYNC();
        uint32_t* t = a;
        a = b;
        b = t;
        n = half;
        lvl++;
        idx0 >>= 1;
    }
    return a;
}

// Block-wide form of narrow_reduce: n (a power of two) nodes stored node-major
// in LDS buffer a are reduced to one, every quad of the block taking one
// compression per level (ping-ponging with b), so a level of up to
// THREADS / 4 compressions costs one quad-cooperative compression of latency.
// The nodes are at tree level lvl, node 0 having index idx0 there; it stops at
// `stop` nodes (a power of two >= 1). All threads of the block must call it;
// returns the buffer holding the remaining nodes.
template <uint32_t THREADS>
__device__ __forceinline__ uint32_t* block_quad_reduce(uint32_t* a, uint32_t* b, uint32_t& n, uint32_t& lvl,
                                                       uint64_t& idx0, const LevelStore& st, uint32_t stop) {
    constexpr uint32_t QUADS = THREADS / 4;
    const uint32_t quad = threadIdx.x >> 2, l = threadIdx.x & 3;
    const uint32_t wave_q0 = (threadIdx.x >> 6) * 16;  // first quad of this wave
    while (n > stop) {
        const uint32_t half = n / 2;
        for (uint32_t q0 = 0; q0 < half; q0 += QUADS) {
            if (q0 + wave_q0 >= half) continue;  // wave-uniform: no quad of this wave has a node
            const uint32_t q = q0 + quad;
            const uint32_t qq = q < half ? q : half - 1;  // whole quads stay active for the DPP permutes
            uint32_t hl, hh;
            b2s_quad_init(hl, hh);
            b2s_compress_quad(a + 16 * qq, 64u, 0xFFFFFFFFu, hl, hh);
            if (q < half) {
                b[8 * q + l] = hl;
                b[8 * q + 4 + l] = hh;
                if (st.want(lvl + 1)) {
                    uint32_t* dst = st.at(lvl + 1, (idx0 >> 1) + q);
                    dst[l] = hl;
                    dst[4 + l] = hh;
                }
            }
        }
        __syncthreads();
        uint32_t* t = a;
        a = b;
        b = t;
        n = half;
        lvl++;
        idx0 >>= 1;
    }
    return a;
}
template <uint32_t THREADS>
__device__ __forceinline__ uint32_t* block_quad_reduce(uint32_t* a, uint32_t* b, uint32_t n, uint32_t lvl,
                                                       uint64_t idx0, const LevelStore& st) {
    return block_quad_reduce<THREADS>(a, b, n, lvl, idx0, st, 1u);
}

// Fiat-Shamir step after a commitment, by one whole wave (quad 0 is the one whose
// results are kept): optional reset, mix_digest(root) and n_felts draw_felt
// calls into alpha_dst[4k..4k+4) (channel.rs:30-39, 67-86). root and msg are
// LDS (8 and 16 words).
__device__ __forceinline__ void quad_transcript(DevChannel* ch, const uint32_t* root, uint32_t* msg, bool init,
                                                uint32_t* alpha_dst, uint32_t n_felts = 1) {
    const uint32_t lane = threadIdx.x & 63, l = lane & 3;
    if (init && lane == 0) ch->failed = 0;
    if (lane < 8) msg[lane] = init ? 0u : ch->state[lane];
    if (lane >= 8 && lane < 16) msg[lane] = root[lane - 8];
    VX_WAVE_SYNC();
    uint32_t hl, hh;
    b2s_quad_init(hl, hh);
    b2s_compress_quad(msg, 64u, 0xFFFFFFFFu, hl, hh);
    VX_WAVE_SYNC();
    // new state words l and 4 + l live in every quad; publish quad 0's
    if (lane < 4) {
        msg[l] = hl;
        msg[4 + l] = hh;
    }
    VX_WAVE_SYNC();
    if (lane < 8) ch->state[lane] = msg[lane];
    uint32_t n_draws = 0;
    for (uint32_t k = 0; alpha_dst && k < n_felts; k++) {
        bool got = false;
        for (int tries = 0; tries < 64; tries++) {
            if (lane < 8) msg[8 + lane] = lane == 0 ? n_draws : 0u;
            VX_WAVE_SYNC();
            uint32_t wl, wh;
            b2s_quad_init(wl, wh);
            b2s_compress_quad(msg, 37u, 0xFFFFFFFFu, wl, wh);
            n_draws++;
            const uint64_t ok = __ballot(wl < 2u * P && wh < 2u * P);
            VX_WAVE_SYNC();
            if ((ok & 0xFull) == 0xFull) {
                if (lane < 4) alpha_dst[4 * k + l] = canon(wl);
                got = true;
                break;
            }
        }
        if (!got) {
            if (lane == 0) ch->failed = 1;
            break;
        }
    }
    if (lane == 0) ch->n_draws = n_draws;
}

// ---------------------------------------------------------------- stage kernel
// LDS per wave: up to logc_max pending chunks + 1 exchange chunk, each 64
// nodes stored word-major ([word][lane]) so chunk writes are conflict-free.
// A source may cap logc_max (STAGE_LOGC_MAX) to trade tree depth per wave for
// occupancy: memory-heavy leaf sources need more waves in flight.
template <class S, class = void>
struct StageLogcMax {
    static constexpr uint32_t value = MERKLE_LOGC_MAX;
};
template <class S>
struct StageLogcMax<S, std::void_t<decltype(S::STAGE_LOGC_MAX)>> {
    static constexpr uint32_t value = S::STAGE_LOGC_MAX;
};
// The tuned depth table is timed on single-column leaves, whose stage is
// compute-bound; a source that caps its depth is memory-heavy and keeps the
// occupancy rule of pick_logc (a tuned depth of 4-5 leaves it one wave per
// SIMD, with its leaf loads exposed).
template <class S>
constexpr bool stage_uses_plan() {
    return StageLogcMax<S>::value == MERKLE_LOGC_MAX;
}
template <class S>
constexpr uint32_t stage_wave_lds() {
    return (StageLogcMax<S>::value + 1) * 8 * 64;
}

// blockIdx.y selects the proof of a batch: every per-proof buffer (items,
// out_nodes, level store) lies blockIdx.y * bs words past proof 0's.
// A wave hands on `keep` (1 or MERKLE_STAGE_KEEP) nodes of its subtree: the
// last levels of its 64 -> 1 reduction leave most of the wave's quads idle, and
// the next stage computes those nodes one per lane instead.
template <class Src>
__global__ __launch_bounds__(64 * MERKLE_STAGE_WAVES) void merkle_stage_kernel(Src src0, uint32_t item_level,
                                                                             uint64_t n_items, uint32_t logc,
                                                                             uint32_t* out_nodes, LevelStore st0,
                                                                             uint64_t bs, uint32_t keep) {
    constexpr uint32_t WAVE_LDS = stage_wave_lds<Src>();
    __shared__ uint32_t lds[MERKLE_STAGE_WAVES * WAVE_LDS];
    const uint64_t boff = (uint64_t)blockIdx.y * bs;
    const Src src = src0.at(boff);
    const LevelStore st = st0.shifted(boff);
    out_nodes += boff;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t* wl = lds + w * WAVE_LDS;
    uint32_t* xchg = wl + StageLogcMax<Src>::value * 8 * 64;
    const uint64_t wave_id = (uint64_t)blockIdx.x * MERKLE_STAGE_WAVES + w;
    const uint64_t base = wave_id << (7 + logc);
    if (base >= n_items) return;  // wave-uniform
    const uint32_t n_chunks = 1u << logc;
    uint32_t h[8];
    for (uint32_t c = 0; c < n_chunks; c++) {
        const uint64_t i = base + 128ull * c + 2 * lane;
        src.pair(i, h);
        uint32_t lvl = item_level + 1, d = 0, k = c;
        if (st.want(lvl)) store8(st.at(lvl, i >> 1), h);
        while (k & 1) {
            // left sibling chunk waits in slot d, right chunk is h
#pragma unroll
            for (int q = 0; q < 8; q++) xchg[q * 64 + lane] = h[q];
            VX_WAVE_SYNC();
            const uint32_t* srcc = lane < 32 ? wl + d * 8 * 64 : xchg;
            const uint32_t li = (lane & 31) * 2;
            uint32_t l[8], r[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                l[q] = srcc[q * 64 + li];
                r[q] = srcc[q * 64 + li + 1];
            }
            VX_WAVE_SYNC();
            b2s_node(l, r, h);
            lvl++;
            d++;
            k >>= 1;
            if (st.want(lvl)) {
                uint64_t idx = (base >> (lvl - item_level)) + 64ull * (c >> d) + lane;
                store8(st.at(lvl, idx), h);
            }
        }
        if (c + 1 < n_chunks) {
#pragma unroll
            for (int q = 0; q < 8; q++) wl[d * 8 * 64 + q * 64 + lane] = h[q];
            VX_WAVE_SYNC();
        }
    }
    // the wave's 64 top nodes (level item_level + 1 + logc) -> `keep` nodes
#pragma unroll
    for (int q = 0; q < 8; q++) wl[lane * 8 + q] = h[q];
    VX_WAVE_SYNC();
    const uint32_t* top = narrow_reduce(wl, wl + 512, 64, item_level + 1 + logc, wave_id * 64, st, keep);
    for (uint32_t k = lane; k < 8 * keep; k += 64) out_nodes[8 * keep * wave_id + k] = top[k];
}

// ---------------------------------------------------------------- small line folds
// Leaf stage of a small line-fold layer (fri.cu:10-39 and the SoA4 leaf rule):
// one leaf per quad, hashed quad-cooperatively, 16 * R leaves a wave reduced to
// one node. On the few waves such a layer has (one per SIMD), a wave's chain is
// then one quad compression for the leaves and log2(16 R) node levels, where
// the lane-per-pair form (128 leaves a wave) takes three lane compressions and
// seven quad rounds. Value i of the layer: every lane of its quad folds it,
// lane l stores and hashes coordinate l.
constexpr uint32_t LINE_QUAD_WAVES = 4;  // waves per workgroup
template <uint32_t R>
__global__ __launch_bounds__(64 * LINE_QUAD_WAVES) void line_fold_quad_stage_kernel(SrcLineFold src0, uint64_t n_items,
                                                                                  uint32_t* out_nodes, LevelStore st0,
                                                                                  uint64_t bs) {
    constexpr uint32_t LPW = 16 * R;  // leaves a wave
    __shared__ uint32_t lds[LINE_QUAD_WAVES][2 * 8 * LPW + 16 * 16];
    const uint64_t boff = (uint64_t)blockIdx.y * bs;
    const SrcLineFold src = src0.at(boff);
    const LevelStore st = st0.shifted(boff);
    out_nodes += boff;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6, quad = lane >> 2, l = lane & 3;
    const uint64_t wave_id = (uint64_t)blockIdx.x * LINE_QUAD_WAVES + w;
    const uint64_t base = wave_id * LPW;
    if (base >= n_items) return;  // wave-uniform
    uint32_t* na = lds[w];                         // LPW nodes, node-major
    uint32_t* nb = na + 8 * LPW;                   // the next level
    uint32_t* msg = nb + 8 * LPW + 16 * quad;      // the quad's leaf message
    const QM31 al = load_alpha(src.alpha);
    uint32_t* out = l == 0 ? src.o0 : l == 1 ? src.o1 : l == 2 ? src.o2 : src.o3;
    for (uint32_t r = 0; r < R; r++) {
        const uint64_t i = base + 16 * r + quad;
        const uint2 a = *(const uint2*)(src.i0 + 2 * i), b = *(const uint2*)(src.i1 + 2 * i);
        const uint2 c = *(const uint2*)(src.i2 + 2 * i), d = *(const uint2*)(src.i3 + 2 * i);
        const QM31 f = fold_pair({{a.x, b.x, c.x, d.x}}, {{a.y, b.y, c.y, d.y}}, al, src.tw[i]);
        const uint32_t fl = l == 0 ? f.v[0] : l == 1 ? f.v[1] : l == 2 ? f.v[2] : f.v[3];
        out[i] = fl;
        msg[l] = fl;
        msg[4 + l] = 0;
        msg[8 + l] = 0;
        msg[12 + l] = 0;
        VX_WAVE_SYNC();
        uint32_t hl, hh;
        b2s_quad_init(hl, hh);
        b2s_compress_quad(msg, 16u, 0xFFFFFFFFu, hl, hh);
        const uint32_t k = 16 * r + quad;
        na[8 * k + l] = hl;
        na[8 * k + 4 + l] = hh;
        VX_WAVE_SYNC();  // the message is rewritten next round; the leaves are read below
    }
    const uint32_t* top = narrow_reduce(na, nb, LPW, 0, base, st);
    if (lane < 8) out_nodes[8 * wave_id + lane] = top[lane];
}

// ---------------------------------------------------------------- tail kernel
struct TailOut {
    uint32_t* root_dst;      // 8 words (e.g. inside the proof buffer); may be null
    DevChannel* channel;     // transcript to advance; may be null
    uint32_t* alpha_dst;     // challenges drawn after mixing the root; may be null
    uint32_t init_channel;   // reset the transcript before mixing
    uint32_t n_felts = 1;    // challenges drawn (4 words each)
};

constexpr uint32_t TAIL_THREADS = 1024;

// Block b reduces items [b * n_items, (b + 1) * n_items) to one root (the
// tiled form used by the reference-ABI entry points runs many blocks; the
// prover runs one). n_items is a power of two in [2, MERKLE_TAIL_MAX].
// blockIdx.y selects the proof of a batch as in merkle_stage_kernel; the root,
// transcript and challenge destinations move with it.
template <class Src>
__global__ __launch_bounds__(TAIL_THREADS) void merkle_tail_kernel(Src src0, uint32_t item_level, uint32_t n_items,
                                                                  LevelStore st0, TailOut out, uint64_t bs) {
    constexpr uint32_t W = MERKLE_TAIL_MAX / 2;
    constexpr uint32_t QUADS = TAIL_THREADS / 4;
    const uint64_t boff = (uint64_t)blockIdx.y * bs;
    const Src src = src0.at(boff);
    const LevelStore st = st0.shifted(boff);
    if (boff) {
        if (out.root_dst) out.root_dst += boff;
        if (out.channel) out.channel = (DevChannel*)((uint32_t*)out.channel + boff);
        if (out.alpha_dst) out.alpha_dst += boff;
    }
    __shared__ uint32_t nodes[8 * W];  // wide levels, word-major: word q of node j at q * W + j
    __shared__ uint4 qa[4 * QUADS];    // quad levels, node-major: up to 2 * QUADS nodes
    __shared__ uint4 qb[2 * QUADS];
    __shared__ uint32_t na[8 * 64], nb[8 * 64];
    __shared__ uint32_t msg[16];
    const uint32_t tid = threadIdx.x;
    const uint64_t item_base = (uint64_t)blockIdx.x * n_items;
    uint32_t n = n_items / 2;
    uint32_t lvl = item_level + 1;
    for (uint32_t j = tid; j < n; j += TAIL_THREADS) {
        uint32_t h[8];
        src.pair(item_base + 2ull * j, h);
        if (st.want(lvl)) store8(st.at(lvl, (item_base >> 1) + j), h);
#pragma unroll
        for (int q = 0; q < 8; q++) nodes[q * W + j] = h[q];
    }
    __syncthreads();
    // wide levels (more compressions than the block has quads): one per thread
    while (n > 2 * QUADS) {
        const uint32_t half = n / 2;
        uint32_t l[8], r[8];
        if (tid < half) {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                l[q] = nodes[q * W + 2 * tid];
                r[q] = nodes[q * W + 2 * tid + 1];
            }
        }
        __syncthreads();
        lvl++;
        if (tid < half) {
            uint32_t h[8];
            b2s_node(l, r, h);
#pragma unroll
            for (int q = 0; q < 8; q++) nodes[q * W + tid] = h[q];
            if (st.want(lvl)) store8(st.at(lvl, (item_base >> (lvl - item_level)) + tid), h);
        }
        __syncthreads();
        n = half;
    }
    // levels of 32 to 256 compressions: quad-cooperative across the whole block
    uint64_t idx0 = item_base >> (lvl - item_level);
    uint32_t* cur = na;
    if (n > 64) {
        uint32_t* qaw = (uint32_t*)qa;
        for (uint32_t j = tid; j < n; j += TAIL_THREADS)
#pragma unroll
            for (int q = 0; q < 8; q++) qaw[8 * j + q] = nodes[q * W + j];
        __syncthreads();
        cur = block_quad_reduce<TAIL_THREADS>(qaw, (uint32_t*)qb, n, lvl, idx0, st, 64u);
    } else {
        for (uint32_t j = tid; j < n; j += TAIL_THREADS)
#pragma unroll
            for (int q = 0; q < 8; q++) na[8 * j + q] = nodes[q * W + j];
        __syncthreads();
    }
    if (tid >= 64) return;  // one wave finishes: narrow levels and transcript
    if (cur != na) {
        for (uint32_t k = tid; k < 8 * n; k += 64) na[k] = cur[k];
        VX_WAVE_SYNC();
    }
    uint32_t* root = narrow_reduce(na, nb, n, lvl, idx0, st);
    if (out.root_dst && tid < 8) out.root_dst[8ull * blockIdx.x + tid] = root[tid];
    if (out.channel) quad_transcript(out.channel, root, msg, out.init_channel != 0, out.alpha_dst, out.n_felts);
}

}  // namespace vx
