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
//    workgroup barrier is ever needed. The wave emits 64 nodes, 2^(logc+1)
//    levels up.
//  * Tail kernel: one workgroup reduces the last <= 4096 items to the root,
//    writes it where the caller asks and advances the Fiat-Shamir transcript
//    (mix the root, draw the next challenge) in the same launch.
//  * Every node at level >= store_from is also written to a per-tree level
//    store; decommitment recomputes the 2^store_from-leaf subtree of a query
//    and reads the rest, so full trees are never written to HBM.
#pragma once
#include "blake2s.hpp"
#include "field.hpp"
#include "transcript.hpp"

namespace vx {

constexpr uint32_t MERKLE_TAIL_MAX = 4096;  // items reduced by the tail kernel
constexpr uint32_t MERKLE_LOGC_MAX = 5;     // chunks per wave in a stage: up to 32
constexpr uint32_t MERKLE_STAGE_WAVES = 4;  // waves per stage workgroup

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

struct SrcNodes {
    const uint32_t* nodes;
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
    __device__ __forceinline__ void pair(uint64_t i, uint32_t h[8]) const {
        QM31 aq = load_alpha(alpha_q), af = load_alpha(alpha_f);
        uint64_t j = 2 * i;  // quotient index of the first of four inputs
        uint2 a = *(const uint2*)(t + j);
        uint2 b = *(const uint2*)(t + ((j + 2) & mask));
        uint2 c = *(const uint2*)(t + ((j + 4) & mask));
        QM31 q0 = quotient_at(a.x, a.y, b.x, aq);
        QM31 q1 = quotient_at(a.y, b.x, b.y, aq);
        QM31 q2 = quotient_at(b.x, b.y, c.x, aq);
        QM31 q3 = quotient_at(b.y, c.x, c.y, aq);
        uint2 w = *(const uint2*)(tw + i);
        QM31 f0 = fold_pair(q0, q1, af, w.x);
        QM31 f1 = fold_pair(q2, q3, af, w.y);
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

// ---------------------------------------------------------------- stage kernel
// LDS per wave: MERKLE_LOGC_MAX pending chunks + 1 exchange chunk, each 64
// nodes stored word-major ([word][lane]) so chunk writes are conflict-free.
// Orders this wave's LDS writes before its later LDS reads (the LDS executes a
// wave's requests in order; this keeps the compiler from reordering them).
#define VX_WAVE_SYNC()                                        \
    do {                                                      \
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); \
        __builtin_amdgcn_wave_barrier();                      \
    } while (0)

constexpr uint32_t STAGE_SLOTS = MERKLE_LOGC_MAX + 1;
constexpr uint32_t STAGE_WAVE_LDS = STAGE_SLOTS * 8 * 64;

template <class Src>
__global__ __launch_bounds__(64 * MERKLE_STAGE_WAVES) void merkle_stage_kernel(Src src, uint32_t item_level,
                                                                             uint64_t n_items, uint32_t logc,
                                                                             uint32_t* out_nodes, LevelStore st) {
    __shared__ uint32_t lds[MERKLE_STAGE_WAVES * STAGE_WAVE_LDS];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t* wl = lds + w * STAGE_WAVE_LDS;
    uint32_t* xchg = wl + MERKLE_LOGC_MAX * 8 * 64;
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
    store8(out_nodes + 8 * (wave_id * 64 + lane), h);
}

// ---------------------------------------------------------------- tail kernel
struct TailOut {
    uint32_t* root_dst;      // 8 words (e.g. inside the proof buffer); may be null
    DevChannel* channel;     // transcript to advance; may be null
    uint32_t* alpha_dst;     // challenge drawn after mixing the root; may be null
    uint32_t init_channel;   // reset the transcript before mixing
};

constexpr uint32_t TAIL_THREADS = 1024;

// Block b reduces items [b * n_items, (b + 1) * n_items) to one root (the
// tiled form used by the reference-ABI entry points runs many blocks; the
// prover runs one). n_items is a power of two in [2, MERKLE_TAIL_MAX].
template <class Src>
__global__ __launch_bounds__(TAIL_THREADS) void merkle_tail_kernel(Src src, uint32_t item_level, uint32_t n_items,
                                                                  LevelStore st, TailOut out) {
    // nodes at one level, word-major: word q of node j at q * 2048 + j
    __shared__ uint32_t nodes[8 * (MERKLE_TAIL_MAX / 2)];
    const uint32_t tid = threadIdx.x;
    const uint64_t item_base = (uint64_t)blockIdx.x * n_items;
    uint32_t n = n_items / 2;
    uint32_t lvl = item_level + 1;
    for (uint32_t j = tid; j < n; j += TAIL_THREADS) {
        uint32_t h[8];
        src.pair(item_base + 2ull * j, h);
        if (st.want(lvl)) store8(st.at(lvl, (item_base >> 1) + j), h);
#pragma unroll
        for (int q = 0; q < 8; q++) nodes[q * 2048 + j] = h[q];
    }
    __syncthreads();
    while (n > 1) {
        const uint32_t half = n / 2;
        uint32_t l[8], r[8];
        if (tid < half) {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                l[q] = nodes[q * 2048 + 2 * tid];
                r[q] = nodes[q * 2048 + 2 * tid + 1];
            }
        }
        __syncthreads();
        lvl++;
        if (tid < half) {
            uint32_t h[8];
            b2s_node(l, r, h);
#pragma unroll
            for (int q = 0; q < 8; q++) nodes[q * 2048 + tid] = h[q];
            if (st.want(lvl)) store8(st.at(lvl, (item_base >> (lvl - item_level)) + tid), h);
        }
        __syncthreads();
        n = half;
    }
    if (tid == 0) {
        uint32_t root[8];
#pragma unroll
        for (int q = 0; q < 8; q++) root[q] = nodes[q * 2048];
        if (out.root_dst)
            for (int q = 0; q < 8; q++) out.root_dst[8ull * blockIdx.x + q] = root[q];
        if (out.channel) {
            if (out.init_channel) {
                for (int q = 0; q < 8; q++) out.channel->state[q] = 0;
                out.channel->n_draws = 0;
                out.channel->failed = 0;
            }
            ch_mix_digest(out.channel, root);
            if (out.alpha_dst) ch_draw_felt(out.channel, out.alpha_dst);
        }
    }
}

}  // namespace vx
