// Native circle NTT and its twiddles for gfx950.
//
// Transform semantics are those of cuda/circle_ntt.cu:8-50,181-249: layer b
// (butterfly stride 2^b) pairs i0 = (h << (b+1)) + l with i1 = i0 + 2^b and
// uses twiddle table[h]; b >= 1 reads line layer b-1, b = 0 the circle table.
// Evaluate runs b = k-1 .. 0, interpolate runs b = 0 .. k-1 and scales by
// 2^((30k) mod 31).
//
// Execution is MI355X-shaped instead of one launch per layer: a transform is
// cut into a few passes, each applying several consecutive layers to tiles
// held in LDS. The pass holding the low strides works on contiguous tiles of
// up to 2^14 words (64 KB LDS); passes over higher strides gather 2^m rows of
// 2^c contiguous words (c = 4, 5, 6: 64- to 256-byte segments) so every HBM access stays
// coalesced. A 2^26-point evaluate is 3 passes instead of 26.
//
// The same pass kernel runs the stwo-basis transform (cuda/circle_ntt_stwo.cu):
// layer b >= 1 reads the flat stwo buffer at n/2 - n/2^b, and the circle layer
// derives its twiddle for pair t from the first line layer as
// [y, -y, -x, x][t % 4] with (x, y) = (T[2(t/4)], T[2(t/4) + 1]).
#include <algorithm>
#include <functional>
#include <mutex>

#include "common.hpp"
#include "kernels.hpp"
#include "ntt_plan.hpp"

namespace vx {

// ---------------------------------------------------------------- host circle math
namespace {
constexpr uint32_t HP = 0x7FFFFFFFu;
inline uint32_t hmul(uint32_t a, uint32_t b) {
    uint64_t p = (uint64_t)a * b;
    uint32_t r = (uint32_t)(p & HP) + (uint32_t)(p >> 31);
    return r >= HP ? r - HP : r;
}
inline uint32_t hadd(uint32_t a, uint32_t b) {
    uint32_t r = a + b;
    return r >= HP ? r - HP : r;
}
inline uint32_t hsub(uint32_t a, uint32_t b) { return a >= b ? a - b : a + HP - b; }
}  // namespace

HostPt host_pmul(HostPt a, HostPt b) {
    return {hsub(hmul(a.x, b.x), hmul(a.y, b.y)), hadd(hmul(a.x, b.y), hmul(a.y, b.x))};
}
HostPt host_repeated_double(HostPt p, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) p = host_pmul(p, p);
    return p;
}
HostPt host_generator() { return {2u, 1268011823u}; }  // src/circle.rs:26-29
void host_coset(int kind, uint32_t k, HostPt& initial, HostPt& step) {
    HostPt g = host_generator();
    step = host_repeated_double(g, 31 - k);
    if (kind == 0)
        initial = {1, 0};
    else if (kind == 1)
        initial = host_repeated_double(g, 30 - k);
    else
        initial = host_repeated_double(g, k >= 29 ? 0 : 29 - k);
}
static HostPt host_pow(HostPt p, uint64_t e) {
    HostPt r = {1, 0};
    while (e) {
        if (e & 1) r = host_pmul(r, p);
        p = host_pmul(p, p);
        e >>= 1;
    }
    return r;
}

// ---------------------------------------------------------------- pass kernel
struct NttPassArgs {
    uint32_t* dst;
    const uint32_t* src;
    uint32_t* const* cols;        // batched: column blockIdx.y is written here
    const uint32_t* const* srcs;  // batched: and read from here (null: in place)
    // twiddle table of each layer of this pass, row bit rb = 0 .. m-1 (global
    // stride bit s + rb): the line layer s + rb, or at stride bit 0 the circle
    // table (the stwo variant derives its circle twiddles from the first line
    // layer, which tw[0] then points at). Set on the host, so the kernel reads
    // one pointer per layer instead of looking offsets up.
    const uint32_t* tw[NTT_LOW_MAX];
    // the LDE bridge's inverse layers (the interpolation's tables), row bit rb
    const uint32_t* tw2[NTT_LOW_MAX];
    uint64_t src_n;  // words of src that are read; the rest is zero
    uint32_t s;      // lowest global stride bit of this pass
    uint32_t k;      // log2 of the transform size
    uint32_t scale;  // multiply on store when non-zero
    uint32_t vec;    // every column pointer is 16-byte aligned: move tiles 4 words per lane
    uint32_t zero_top2;  // forward, top pass, src_n = n / 4: the two top layers only copy
};

// LDS word index with one pad word per 8, which keeps the strided radix-8
// accesses below free of bank conflicts.
__device__ __forceinline__ uint32_t lpad(uint32_t e) { return e + (e >> 3); }

// Word h / word pair h of a twiddle table whose address is wave-uniform. A
// layer's table holds at most 2^30 words (transforms of up to 2^31 points), so
// the byte offset is a 32-bit value: the load takes the table in scalar
// registers plus one 32-bit offset per lane instead of a 64-bit address
// computed per lane.
__device__ __forceinline__ uint32_t ldw(const uint32_t* T, uint32_t h) {
    return *(const uint32_t*)((const char*)T + (h << 2));
}
__device__ __forceinline__ uint2 ldw2(const uint32_t* T, uint32_t h) {
    return *(const uint2*)((const char*)T + (h << 3));
}

// Threads of a pass over 2^M x 2^C tiles: one radix-8 group per thread, 64 to 256.
template <int M, int C>
constexpr uint32_t ntt_threads() {
    constexpr uint32_t groups = 1u << (M + C > 3 ? M + C - 3 : 0);
    return groups >= 256 ? 256 : (groups < 64 ? 64 : groups);
}

// Twiddle h of a layer whose table is T; circle0: the layer is stride bit 0.
template <bool STWO>
__device__ __forceinline__ uint32_t ntt_twiddle(const uint32_t* T, bool circle0, uint32_t h) {
    if (!STWO || !circle0) return ldw(T, h);
    const uint2 pr = ldw2(T, h >> 2);
    const uint32_t x = pr.x, y = pr.y;
    switch (h & 3) {
        case 0: return y;
        case 1: return mneg(y);
        case 2: return mneg(x);
        default: return x;
    }
}

// One radix-2^R round over row bits [J, J + R) of an LDS tile of 2^M rows x
// 2^C columns. Each group of 2^R values (fixed column and other row bits) is
// loaded into registers, transformed through R layers, and written back.
// J and R are compile-time, so the 2^R LDS addresses of a group are one base
// plus constant offsets, and the twiddle of butterfly d in layer t is one
// per-layer base index plus the constant d >> (t + 1).
// With zero_top set (the first round of a forward transform whose input is
// zero past its first quarter), the values of the group's two upper row bits
// are zero on entry, so the top two layers only copy: (x, 0) -> (x, x). Those
// rounds load two values and replicate them instead of running 8 butterflies.
// ALT: the layers' twiddle tables are a.tw2 (the LDE bridge's inverse layers)
template <bool FWD, int M, int C, int R, int J, bool STWO, bool ALT = false>
__device__ __forceinline__ void ntt_round(uint32_t* tile, const NttPassArgs& a, uint64_t hi, uint32_t tid,
                                          bool zero_top = false) {
    constexpr uint32_t GROUPS = 1u << (M + C - R);
    constexpr uint32_t STEP = 1u << (J + C);  // tile index distance between rows d and d + 1
    constexpr uint32_t NT = ntt_threads<M, C>();
    constexpr uint32_t ITER = (GROUPS + NT - 1) / NT;
    // Thread t takes groups g = t + it * NT; a group is (row_hi, row_lo, q) =
    // (g >> (C+J), the J bits above C, the C low bits) and starts at tile index
    // ((row_hi << (J+R) | row_lo) << C) + q. With NT a multiple of 2^(C+J) the
    // next group of a thread has the same column and low row bits and row_hi
    // larger by NT >> (C+J); otherwise (NT < 2^(C+J), t < 2^(C+J)) group it has
    // row_hi = it / P and its low bits t + (it % P) * NT, P = 2^(C+J) / NT.
    // Either way a thread's groups sit at constant distances from its first in
    // the tile and in each layer's twiddle table, which the unrolled loop folds
    // into instruction offsets.
    constexpr uint32_t CJ = 1u << (C + J);
    constexpr bool WIDE = NT >= CJ;
    constexpr uint32_t P = WIDE ? 1u : CJ / NT;
    const uint32_t g0 = tid;
    const uint32_t q = g0 & ((1u << C) - 1);
    const uint32_t rest = g0 >> C;
    const uint32_t row_lo = rest & ((1u << J) - 1);
    const uint32_t row_hi0 = rest >> J;  // 0 unless WIDE
    const uint32_t e00 = ((((row_hi0 << (J + R)) | row_lo)) << C) + q;
#pragma unroll
    for (uint32_t it = 0; it < ITER; it++) {
        if (GROUPS % NT != 0 && g0 + it * NT >= GROUPS) break;  // fewer groups than threads
        const uint32_t e0 = WIDE ? e00 + it * (NT << R) : e00 + (it % P) * NT + ((it / P) << (J + R + C));
        const uint32_t row_hi = WIDE ? row_hi0 + it * (NT / CJ) : it / P;
        uint32_t v[1 << R];
        uint32_t addr[1 << R];
#pragma unroll
        for (int d = 0; d < (1 << R); d++) {
            // lpad is linear in d once the row step is a multiple of 8 words
            addr[d] = (J + C >= 3) ? lpad(e0) + (uint32_t)d * (STEP + STEP / 8) : lpad(e0 + (uint32_t)d * STEP);
        }
        if (FWD && R == 3 && zero_top) {
            v[0] = tile[addr[0]];
            v[1] = tile[addr[1]];
#pragma unroll
            for (int d = 2; d < (1 << R); d++) v[d] = v[d & 1];
        } else {
#pragma unroll
            for (int d = 0; d < (1 << R); d++) v[d] = tile[addr[d]];
        }
#pragma unroll
        for (int tt = 0; tt < R; tt++) {
            const int t = FWD ? (R - 1 - tt) : tt;  // layer inside the group
            if (FWD && R == 3 && t >= 1 && zero_top) continue;
            const uint32_t rb = J + t;                 // row bit of the layer
            const uint32_t* T = ALT ? a.tw2[rb] : a.tw[rb];
            // only the contiguous pass (stride bits from 0) has a stwo variant,
            // so the circle layer is row bit 0, known at compile time
            const bool circle0 = STWO && rb == 0;
            // h(d) = (hi << (M - rb - 1)) + (row >> (rb + 1)) = hb + (d >> (t + 1))
            const uint32_t hb = (uint32_t)(hi << (M - rb - 1)) + (row_hi << (R - t - 1));
            if (STWO && R == 3 && J == 0 && t == 0) {
                // the circle layer in a radix-8 round: its four butterflies
                // (d = 0, 2, 4, 6) take h = hb .. hb + 3 with hb a multiple of
                // 4, i.e. [y, -y, -x, x] of the pair at hb / 2; a negated
                // twiddle swaps the two outputs (forward) or negates the
                // difference (inverse), so no negation is computed
                const uint2 xy = ldw2(T, hb >> 2);
#pragma unroll
                for (int d = 0; d < 8; d += 2) {
                    const uint32_t w = d < 4 ? xy.y : xy.x;
                    const bool neg = d == 2 || d == 4;
                    uint32_t& x0 = v[d];
                    uint32_t& x1 = v[d + 1];
                    if (FWD) {
                        const uint32_t m = mmul(x1, w);
                        const uint32_t s0 = madd(x0, m), s1 = msub(x0, m);
                        x0 = neg ? s1 : s0;
                        x1 = neg ? s0 : s1;
                    } else {
                        const uint32_t s0 = madd(x0, x1);
                        x1 = mmul(neg ? msub(x1, x0) : msub(x0, x1), w);
                        x0 = s0;
                    }
                }
                continue;
            }
#pragma unroll
            for (int d = 0; d < (1 << R); d++) {
                if (d & (1 << t)) continue;
                const uint32_t w = ntt_twiddle<STWO>(T, circle0, hb + ((uint32_t)d >> (t + 1)));
                uint32_t& x0 = v[d];
                uint32_t& x1 = v[d | (1 << t)];
                if (FWD) {
                    uint32_t m = mmul(x1, w);
                    x1 = msub(x0, m);
                    x0 = madd(x0, m);
                } else {
                    uint32_t s0 = madd(x0, x1);
                    x1 = mmul(msub(x0, x1), w);
                    x0 = s0;
                }
            }
        }
#pragma unroll
        for (int d = 0; d < (1 << R); d++) tile[addr[d]] = v[d];
    }
}

// Radix-8 rounds over the M layers of a tile, top down (forward) or bottom up
// (inverse), unrolled at compile time.
// tid is the thread's index in the block (see the replicating pass for why it
// is a parameter).
template <bool FWD, int M, int C, bool STWO, int TOP>
__device__ __forceinline__ void fwd_rounds(uint32_t* tile, const NttPassArgs& a, uint64_t hi, uint32_t tid) {
    if constexpr (TOP > 0) {
        constexpr int R = TOP >= 3 ? 3 : TOP;
        // only the pass's first round (its top three row bits) may skip its top two layers
        ntt_round<FWD, M, C, R, TOP - R, STWO>(tile, a, hi, tid, TOP == M && a.zero_top2 != 0);
        __syncthreads();
        fwd_rounds<FWD, M, C, STWO, TOP - R>(tile, a, hi, tid);
    }
}
template <bool FWD, int M, int C, bool STWO, int LO, bool ALT = false>
__device__ __forceinline__ void inv_rounds(uint32_t* tile, const NttPassArgs& a, uint64_t hi, uint32_t tid) {
    if constexpr (LO < M) {
        constexpr int R = M - LO >= 3 ? 3 : M - LO;
        ntt_round<FWD, M, C, R, LO, STWO, ALT>(tile, a, hi, tid);
        __syncthreads();
        inv_rounds<FWD, M, C, STWO, LO + R, ALT>(tile, a, hi, tid);
    }
}

// Applies the M layers with stride bits [s, s + M) of a size-2^k transform to
// the tile of 2^M rows (stride 2^s) by 2^C contiguous columns owned by this block.
// Tiles move between HBM and LDS 4 words per lane (rows of >= 4 contiguous
// words, or the contiguous tile of the first pass) when the buffers allow it.
//
// REPL (forward only): the pass is the top pass of an evaluation whose input is
// zero past its first quarter, with two layers above the tile's M. Those two
// only copy the first quarter into all four, so the block loads its tile from
// the first quarter and, for each quarter q, runs the M layers with q as the
// row bits above the tile and stores the result to quarter q. The tile is
// kept in registers between quarters (or reloaded, with word moves); quarter 0
// is done last, so in place it overwrites its input only after every reload.
// Tile element e (row e >> C, column e & (2^C - 1)) lies at word base + (row <<
// s) + column of its column buffer (< 2^31 words, so a 32-bit index; the
// address itself needs 64 bits).
template <int M, int C>
__device__ __forceinline__ uint32_t tile_word(uint32_t base, uint32_t s, uint32_t e) {
    return base + ((e >> C) << s) + (e & ((1u << C) - 1));
}
template <bool FWD, int M, int C>
__device__ __forceinline__ void tile_load(uint32_t* tile, const uint32_t* src, uint32_t base, uint32_t s,
                                          uint64_t src_n, bool vec, uint32_t tid) {
    constexpr uint32_t TILE = 1u << (M + C);
    constexpr uint32_t NT = ntt_threads<M, C>();
    if (vec) {
#pragma unroll
        for (uint32_t it = 0; it < (TILE + 4 * NT - 1) / (4 * NT); it++) {
            const uint32_t e = 4 * tid + it * 4 * NT;
            if (TILE % (4 * NT) != 0 && e >= TILE) break;
            const uint32_t g = tile_word<M, C>(base, s, e);
            const uint4 x = g < src_n ? *(const uint4*)(src + g) : make_uint4(0, 0, 0, 0);
            const uint32_t p = lpad(e);  // e % 4 == 0: the four words stay within one padded group
            tile[p] = x.x;
            tile[p + 1] = x.y;
            tile[p + 2] = x.z;
            tile[p + 3] = x.w;
        }
    } else {
#pragma unroll
        for (uint32_t it = 0; it < (TILE + NT - 1) / NT; it++) {
            const uint32_t e = tid + it * NT;
            if (TILE % NT != 0 && e >= TILE) break;
            const uint32_t g = tile_word<M, C>(base, s, e);
            tile[lpad(e)] = g < src_n ? src[g] : 0u;
        }
    }
}
template <bool FWD, int M, int C>
__device__ __forceinline__ void tile_store(const uint32_t* tile, uint32_t* dst, uint32_t base, uint32_t s,
                                           uint32_t scale, bool vec, uint32_t tid) {
    constexpr uint32_t TILE = 1u << (M + C);
    constexpr uint32_t NT = ntt_threads<M, C>();
    if (vec) {
#pragma unroll
        for (uint32_t it = 0; it < (TILE + 4 * NT - 1) / (4 * NT); it++) {
            const uint32_t e = 4 * tid + it * 4 * NT;
            if (TILE % (4 * NT) != 0 && e >= TILE) break;
            const uint32_t g = tile_word<M, C>(base, s, e);
            const uint32_t p = lpad(e);
            uint4 x = make_uint4(tile[p], tile[p + 1], tile[p + 2], tile[p + 3]);
            if (scale) x = make_uint4(mmul(x.x, scale), mmul(x.y, scale), mmul(x.z, scale), mmul(x.w, scale));
            *(uint4*)(dst + g) = x;
        }
    } else {
#pragma unroll
        for (uint32_t it = 0; it < (TILE + NT - 1) / NT; it++) {
            const uint32_t e = tid + it * NT;
            if (TILE % NT != 0 && e >= TILE) break;
            const uint32_t g = tile_word<M, C>(base, s, e);
            const uint32_t v = tile[lpad(e)];
            dst[g] = scale ? mmul(v, scale) : v;
        }
    }
}

template <bool FWD, int M, int C, bool STWO, bool REPL>
__global__ __launch_bounds__(256) void ntt_pass_kernel(NttPassArgs a) {
    static_assert(!STWO || C == 0, "the stwo variant is the contiguous pass (stride bits from 0)");
    static_assert(!REPL || FWD, "only an evaluation replicates its input");
    __shared__ uint32_t tile[(1u << (M + C)) + (1u << (M + C)) / 8];
    uint32_t* dst = a.cols ? a.cols[blockIdx.y] : a.dst;
    const uint32_t* src = a.cols ? (a.srcs ? a.srcs[blockIdx.y] : a.cols[blockIdx.y]) : a.src;
    const uint32_t s = a.s;
    constexpr uint32_t TILE = 1u << (M + C);
    const uint32_t mid_bits = s - C;
    const uint32_t blk = blockIdx.x;
    const uint32_t mid = blk & ((1u << mid_bits) - 1);
    const uint32_t hi = blk >> mid_bits;  // REPL: the grid holds no row bits above the tile, hi = 0
    const uint32_t base = (hi << (s + M)) + (mid << C);  // < 2^k <= 2^31 words
    constexpr bool VEC_OK = (C >= 2 || C == 0) && TILE >= 4;
    const bool vec = VEC_OK && a.vec;

    if constexpr (REPL) {
        // With vector moves each thread keeps its part of the quarter's tile in
        // registers and puts it back into LDS for the next quarter (reloading it
        // from memory would miss an L2 that the whole grid's tiles overflow);
        // word moves reload it.
        constexpr uint32_t NT = ntt_threads<M, C>();
        constexpr uint32_t KV = (TILE + 4 * NT - 1) / (4 * NT);
        uint4 keep[KV];
        auto put = [&](uint32_t tid) {
#pragma unroll
            for (uint32_t it = 0; it < KV; it++) {
                const uint32_t e = 4 * tid + it * 4 * NT;
                if (TILE % (4 * NT) != 0 && e >= TILE) break;
                const uint32_t p = lpad(e);
                tile[p] = keep[it].x;
                tile[p + 1] = keep[it].y;
                tile[p + 2] = keep[it].z;
                tile[p + 3] = keep[it].w;
            }
        };
        if (vec) {
#pragma unroll
            for (uint32_t it = 0; it < KV; it++) {
                const uint32_t e = 4 * threadIdx.x + it * 4 * NT;
                if (TILE % (4 * NT) != 0 && e >= TILE) break;
                const uint32_t g = tile_word<M, C>(base, s, e);
                keep[it] = g < a.src_n ? *(const uint4*)(src + g) : make_uint4(0, 0, 0, 0);
            }
            put(threadIdx.x);
        } else {
            tile_load<FWD, M, C>(tile, src, base, s, a.src_n, false, threadIdx.x);
        }
        __syncthreads();
#pragma unroll 1
        for (int q = 3; q >= 0; q--) {
            // Addresses within the tile do not depend on the quarter: left to
            // itself the compiler computes all of them once, before the loop,
            // and holds them in registers (hundreds for the largest tiles); a
            // thread index it cannot see through keeps them inside the loop.
            uint32_t tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            if (q != 3) {
                if (vec)
                    put(tid);
                else
                    tile_load<FWD, M, C>(tile, src, base, s, a.src_n, false, tid);
                __syncthreads();
            }
            fwd_rounds<FWD, M, C, STWO, M>(tile, a, (uint64_t)q, tid);
            tile_store<FWD, M, C>(tile, dst, base + ((uint32_t)q << (s + M)), s, 0u, vec, tid);
            __syncthreads();  // the next quarter overwrites the tile
        }
    } else {
        tile_load<FWD, M, C>(tile, src, base, s, a.src_n, vec, threadIdx.x);
        __syncthreads();
        if constexpr (FWD)
            fwd_rounds<FWD, M, C, STWO, M>(tile, a, hi, threadIdx.x);
        else
            inv_rounds<FWD, M, C, STWO, 0>(tile, a, hi, threadIdx.x);
        tile_store<FWD, M, C>(tile, dst, base, s, a.scale, vec, threadIdx.x);
    }
}

// The bridge of a blowup-4 LDE. When the interpolation's top pass and the
// evaluation's replicating top pass cover the same tile (rows s .. s+M-1 of
// the first quarter, 2^C columns), one pass runs both: the tile of
// interpolation input is loaded, its M inverse layers run with the
// interpolation's twiddles (tw2), the coefficients are scaled by the
// interpolation's 1/n into registers, and each quarter q runs the M forward
// layers with the evaluation's twiddles (tw, row bits above the tile = q)
// and is stored, as in the replicating pass. The quarter's coefficients never
// go to memory. Vector moves only (an LDE's column buffers are 16-byte
// aligned); the driver uses the two passes otherwise.
template <int M, int C, bool STWO>
__global__ __launch_bounds__(256) void ntt_lde_bridge_kernel(NttPassArgs a) {
    static_assert(C >= 2 || C == 0, "vector moves need rows of at least 4 words");
    constexpr uint32_t TILE = 1u << (M + C);
    constexpr uint32_t NT = ntt_threads<M, C>();
    constexpr uint32_t KV = (TILE + 4 * NT - 1) / (4 * NT);
    __shared__ uint32_t tile[TILE + TILE / 8];
    uint32_t* dst = a.cols ? a.cols[blockIdx.y] : a.dst;
    const uint32_t* src = a.cols ? (a.srcs ? a.srcs[blockIdx.y] : a.cols[blockIdx.y]) : a.src;
    const uint32_t s = a.s;
    const uint32_t mid = blockIdx.x;  // the grid holds no row bits above the tile
    const uint32_t base = mid << C;
    tile_load<false, M, C>(tile, src, base, s, a.src_n, true, threadIdx.x);
    __syncthreads();
    inv_rounds<false, M, C, STWO, 0, true>(tile, a, 0, threadIdx.x);
    uint4 keep[KV];
#pragma unroll
    for (uint32_t it = 0; it < KV; it++) {
        const uint32_t e = 4 * threadIdx.x + it * 4 * NT;
        if (TILE % (4 * NT) != 0 && e >= TILE) break;
        const uint32_t p = lpad(e);
        uint4 x = make_uint4(tile[p], tile[p + 1], tile[p + 2], tile[p + 3]);
        if (a.scale) x = make_uint4(mmul(x.x, a.scale), mmul(x.y, a.scale), mmul(x.z, a.scale), mmul(x.w, a.scale));
        keep[it] = x;
    }
#pragma unroll 1
    for (int q = 3; q >= 0; q--) {
        uint32_t tid = threadIdx.x;  // hidden from the compiler, as in the replicating pass
        asm volatile("" : "+v"(tid));
#pragma unroll
        for (uint32_t it = 0; it < KV; it++) {
            const uint32_t e = 4 * tid + it * 4 * NT;
            if (TILE % (4 * NT) != 0 && e >= TILE) break;
            const uint32_t p = lpad(e);
            tile[p] = keep[it].x;
            tile[p + 1] = keep[it].y;
            tile[p + 2] = keep[it].z;
            tile[p + 3] = keep[it].w;
        }
        __syncthreads();
        fwd_rounds<true, M, C, STWO, M>(tile, a, (uint64_t)q, tid);
        tile_store<true, M, C>(tile, dst, base + ((uint32_t)q << (s + M)), s, 0u, true, tid);
        __syncthreads();  // the next quarter overwrites the tile
    }
}

namespace {
std::mutex g_plan_mu;
std::vector<Pass> g_plan[32];      // tuned or set plans; empty: default split
std::vector<Pass> g_lde_plan[32];  // the same for blowup-4 evaluations; empty: g_plan's
}  // namespace

static std::vector<Pass> plan_passes(uint32_t k) {
    {
        std::lock_guard<std::mutex> g(g_plan_mu);
        if (k < 32 && !g_plan[k].empty()) return g_plan[k];
    }
    return default_plan(k);
}
static std::vector<Pass> lde_plan_passes(uint32_t k) {
    {
        std::lock_guard<std::mutex> g(g_plan_mu);
        if (k < 32 && !g_lde_plan[k].empty()) return g_lde_plan[k];
    }
    return plan_passes(k);
}

template <bool FWD, int M, int C, bool STWO, bool REPL>
static void launch_tile(const NttPassArgs& a, uint32_t k, uint32_t n_cols, hipStream_t st) {
    dim3 grid((uint32_t)(1ull << (k - (REPL ? 2 : 0) - M - C)), n_cols);
    ntt_pass_kernel<FWD, M, C, STWO, REPL><<<grid, ntt_threads<M, C>(), 0, st>>>(a);
}

// m: rows of the tile (a replicating pass's layers less its two copy layers)
template <bool FWD, int C, int M, bool STWO, bool REPL>
static void dispatch_m(uint32_t m, const NttPassArgs& a, uint32_t k, uint32_t n_cols, hipStream_t st) {
    if constexpr (M >= 1) {
        if (m == (uint32_t)M) {
            launch_tile<FWD, M, C, STWO, REPL>(a, k, n_cols, st);
            return;
        }
        dispatch_m<FWD, C, M - 1, STWO, REPL>(m, a, k, n_cols, st);
    } else {
        set_error(-2, "ntt: unsupported pass shape");
    }
}

// Only the contiguous pass (c == 0) contains the circle layer, so only it has a
// stwo variant; strided passes differ by their layer offsets alone.
template <bool FWD, bool REPL>
static void launch_pass_dir(bool stwo, const NttPassArgs& a, uint32_t m, uint32_t c, uint32_t k, uint32_t n_cols,
                            hipStream_t st) {
    if (c == 0) {
        if (stwo)
            dispatch_m<FWD, 0, NTT_LOW_MAX, true, REPL>(m, a, k, n_cols, st);
        else
            dispatch_m<FWD, 0, NTT_LOW_MAX, false, REPL>(m, a, k, n_cols, st);
    } else if (c == 4) {
        dispatch_m<FWD, 4, NTT_HIGH_MAX4, false, REPL>(m, a, k, n_cols, st);
    } else if (c == 6) {
        dispatch_m<FWD, 6, NTT_HIGH_MAX6, false, REPL>(m, a, k, n_cols, st);
    } else {
        dispatch_m<FWD, 5, NTT_HIGH_MAX5, false, REPL>(m, a, k, n_cols, st);
    }
}
static void launch_pass(bool fwd, bool stwo, const NttPassArgs& a, const Pass& p, uint32_t k, uint32_t n_cols,
                        hipStream_t st) {
    if (p.r && !fwd) {
        set_error(-2, "ntt: a replicating pass is forward only");
        return;
    }
    if (p.r)
        launch_pass_dir<true, true>(stwo, a, pass_rows(p), p.c, k, n_cols, st);
    else if (fwd)
        launch_pass_dir<true, false>(stwo, a, p.m, p.c, k, n_cols, st);
    else
        launch_pass_dir<false, false>(stwo, a, p.m, p.c, k, n_cols, st);
    check_launch("ntt_pass_kernel");
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// Twiddle table of layer b (stride 2^b) of a transform.
static const uint32_t* layer_table(const NttTables& t, uint32_t b) {
    return b ? t.line + t.offs[b - 1] : (t.stwo ? t.line : t.circle);
}
// 1/n of an inverse transform of 2^k points, 0 when it is 1 (no scaling).
static uint32_t inverse_scale(uint32_t k) {
    const uint32_t ex = (30u * k) % 31u;
    return ex == 0 ? 0u : (1u << ex);
}

// Runs the passes ps (covering some or all layers of a 2^t.log_n transform:
// all of them, or for the LDE bridge the passes below its top pass, whose
// inverse then leaves the 1/n scaling to the bridge: last_scale = false).
static void run_ntt_plan(bool fwd, uint32_t* dst, const uint32_t* src, uint64_t src_n, uint32_t* const* cols,
                         const uint32_t* const* srcs, uint32_t n_cols, const NttTables& t, hipStream_t st,
                         const std::vector<Pass>& ps, bool vec, bool last_scale = true) {
    uint32_t k = t.log_n;
    const uint32_t inv_n = inverse_scale(k);
    for (size_t i = 0; i < ps.size(); i++) {
        size_t pi = fwd ? ps.size() - 1 - i : i;
        NttPassArgs a = {};
        a.cols = cols;
        a.srcs = i == 0 ? srcs : nullptr;
        a.dst = dst;
        a.src = i == 0 ? src : dst;
        a.src_n = i == 0 ? src_n : (1ull << k);
        for (uint32_t rb = 0; rb < ps[pi].m && rb < NTT_LOW_MAX; rb++) a.tw[rb] = layer_table(t, ps[pi].s + rb);
        a.s = ps[pi].s;
        a.k = k;
        a.vec = vec ? 1u : 0u;
        a.scale = (!fwd && last_scale && i + 1 == ps.size()) ? inv_n : 0u;
        // the top pass of an evaluation of n/4 coefficients: its two top layers copy
        const bool quarter = fwd && i == 0 && ps[pi].s + ps[pi].m == k && k >= 2 && src_n == (1ull << k) / 4;
        if (ps[pi].r && !quarter) {
            set_error(-2, "ntt: a replicating pass needs an input that fills the first quarter");
            return;
        }
        a.zero_top2 = (quarter && !ps[pi].r && ps[pi].m >= 3) ? 1u : 0u;
        launch_pass(fwd, t.stwo, a, ps[pi], k, n_cols, st);
    }
    if (!fwd && ps.empty() && src != dst && src_n) {
        check(hipMemcpyAsync(dst, src, 4, hipMemcpyDeviceToDevice, st), "ntt copy");
    }
}
static void run_ntt(bool fwd, uint32_t* dst, const uint32_t* src, uint64_t src_n, uint32_t* const* cols,
                    const uint32_t* const* srcs, uint32_t n_cols, const NttTables& t, hipStream_t st, bool vec) {
    const bool lde = fwd && t.log_n >= 2 && src_n == (1ull << t.log_n) / 4;
    run_ntt_plan(fwd, dst, src, src_n, cols, srcs, n_cols, t, st, lde ? lde_plan_passes(t.log_n) : plan_passes(t.log_n),
                 vec);
}

namespace {
bool g_lde_bridge = true;
}
template <int C, int M, bool STWO>
static void dispatch_bridge(uint32_t m, const NttPassArgs& a, uint32_t ke, uint32_t n_cols, hipStream_t st) {
    if constexpr (M >= 1) {
        if (m == (uint32_t)M) {
            dim3 grid((uint32_t)(1ull << (ke - 2 - M - C)), n_cols);
            ntt_lde_bridge_kernel<M, C, STWO><<<grid, ntt_threads<M, C>(), 0, st>>>(a);
            return;
        }
        dispatch_bridge<C, M - 1, STWO>(m, a, ke, n_cols, st);
    } else {
        set_error(-2, "ntt: unsupported bridge shape");
    }
}

// A blowup-4 LDE of columns: interpolate 2^kt points (from srcs, or in place)
// into the first quarter of each column, evaluate 2^(kt+2) points in place.
// When the evaluation plan ends in a replicating pass over layers [s, kt+2),
// the interpolation's layers [s, kt) run in the same pass (the bridge): the
// interpolation takes its plan's passes below s if they end at s, else the
// default split of those layers.
static void lde_run(uint32_t* const* cols, uint32_t n_cols, const NttTables& ti, const NttTables& te, hipStream_t st,
                    const uint32_t* const* srcs, bool src_aligned, const std::vector<Pass>& pe,
                    const std::function<void(const char*)>& mark) {
    auto marked = [&](const char* what) {
        if (mark) mark(what);
    };
    const uint32_t kt = ti.log_n, ke = te.log_n;
    bool bridge = false;
    {
        std::lock_guard<std::mutex> g(g_plan_mu);
        bridge = g_lde_bridge;
    }
    bridge = bridge && n_cols && kt && ke == kt + 2 && ke >= 3 && ti.stwo == te.stwo && !pe.empty() && pe.back().r == 1 &&
             (pe.back().c == 0 || pe.back().c >= 2) && (pe.back().s > 0 || src_aligned);
    if (!bridge) {
        if (n_cols && kt) {
            ntt_inverse_batch(cols, n_cols, ti, st, srcs, src_aligned);
            marked("interpolate");
            if (ke == kt + 2 && !pe.empty())
                run_ntt_plan(true, nullptr, nullptr, 1ull << kt, cols, nullptr, n_cols, te, st, pe, true);
            else
                ntt_forward_batch(cols, n_cols, te, st, 1ull << kt, true);
            marked("evaluate");
        }
        return;
    }
    const Pass& top = pe.back();
    std::vector<Pass> pil;
    {
        const std::vector<Pass> pi = plan_passes(kt);
        uint32_t covered = 0;
        for (const Pass& x : pi) {
            if (covered + x.m > top.s) break;
            pil.push_back(x);
            covered += x.m;
        }
        if (covered != top.s) pil = default_plan(top.s);
    }
    const std::vector<Pass> pel(pe.begin(), pe.end() - 1);
    if (!pil.empty()) run_ntt_plan(false, nullptr, nullptr, 1ull << kt, cols, srcs, n_cols, ti, st, pil, src_aligned, false);
    marked("interpolate");  // with the bridge: the interpolation's layers below it
    const Pass& pt = top;
    NttPassArgs a = {};
    a.cols = cols;
    a.srcs = pil.empty() ? srcs : nullptr;
    for (uint32_t rb = 0; rb + 2 < pt.m && rb < NTT_LOW_MAX; rb++) {
        a.tw[rb] = layer_table(te, pt.s + rb);
        a.tw2[rb] = layer_table(ti, pt.s + rb);
    }
    a.src_n = 1ull << kt;
    a.s = pt.s;
    a.k = ke;
    a.vec = 1u;
    a.scale = inverse_scale(kt);
    const uint32_t rows = pt.m - 2;
    if (pt.c == 0) {
        if (te.stwo)
            dispatch_bridge<0, NTT_LOW_MAX, true>(rows, a, ke, n_cols, st);
        else
            dispatch_bridge<0, NTT_LOW_MAX, false>(rows, a, ke, n_cols, st);
    } else if (pt.c == 4) {
        dispatch_bridge<4, NTT_HIGH_MAX4, false>(rows, a, ke, n_cols, st);
    } else if (pt.c == 6) {
        dispatch_bridge<6, NTT_HIGH_MAX6, false>(rows, a, ke, n_cols, st);
    } else {
        dispatch_bridge<5, NTT_HIGH_MAX5, false>(rows, a, ke, n_cols, st);
    }
    check_launch("ntt_lde_bridge_kernel");
    if (!pel.empty()) run_ntt_plan(true, nullptr, nullptr, 1ull << ke, cols, nullptr, n_cols, te, st, pel, true);
    marked("evaluate");  // with the bridge: the bridge and the evaluation's lower passes
}
void lde_batch(uint32_t* const* cols, uint32_t n_cols, const NttTables& ti, const NttTables& te, hipStream_t st,
               const uint32_t* const* srcs, bool src_aligned, const std::function<void(const char*)>& mark) {
    lde_run(cols, n_cols, ti, te, st, srcs, src_aligned, te.log_n >= 3 ? lde_plan_passes(te.log_n) : std::vector<Pass>{},
            mark);
}

void ntt_forward(uint32_t* dst, const uint32_t* src, uint64_t src_n, const NttTables& t, hipStream_t s) {
    if (t.log_n == 0) {
        if (src != dst) check(hipMemcpyAsync(dst, src, 4, hipMemcpyDeviceToDevice, s), "ntt copy");
        return;
    }
    run_ntt(true, dst, src, src_n, nullptr, nullptr, 1, t, s, aligned16(dst) && aligned16(src));
}
void ntt_inverse(uint32_t* dst, const uint32_t* src, const NttTables& t, hipStream_t s) {
    run_ntt(false, dst, src, 1ull << t.log_n, nullptr, nullptr, 1, t, s, aligned16(dst) && aligned16(src));
}
void ntt_forward_batch(uint32_t* const* cols, uint32_t n_cols, const NttTables& t, hipStream_t s, uint64_t src_n,
                       bool aligned) {
    if (n_cols && t.log_n)
        run_ntt(true, nullptr, nullptr, src_n ? src_n : (1ull << t.log_n), cols, nullptr, n_cols, t, s, aligned);
}
void ntt_inverse_batch(uint32_t* const* cols, uint32_t n_cols, const NttTables& t, hipStream_t s,
                       const uint32_t* const* srcs, bool aligned) {
    if (n_cols && t.log_n) run_ntt(false, nullptr, nullptr, 1ull << t.log_n, cols, srcs, n_cols, t, s, aligned);
}

NttTables stwo_tables(const uint32_t* flat, uint32_t log_n) {
    NttTables t;
    t.line = flat;
    t.circle = nullptr;
    t.log_n = log_n;
    t.stwo = true;
    uint64_t half = (1ull << log_n) / 2;
    for (uint32_t b = 1; b < log_n && b <= 32; b++) t.offs[b - 1] = (uint32_t)(half - (half >> (b - 1)));
    return t;
}

// ---------------------------------------------------------------- twiddle kernels
// Reference-ABI kernels (cuda/twiddles.cu): coordinates of initial * step^i.
__global__ void coset_points_kernel(CPt initial, CPt step, uint32_t* ox, uint32_t* oy, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    CPt p = cp_pow_mul(initial, step, i);
    ox[i] = p.x;
    oy[i] = p.y;
}
// Points of a geometric progression: thread t of T visits i = t + k * T for
// k < PROG_K (i < n), one point product per step after a single power.
constexpr uint32_t PROG_K = 16;
template <class F>
__device__ __forceinline__ void progression(CPt initial, CPt step, uint64_t n, F&& f) {
    const uint64_t T = (n + PROG_K - 1) / PROG_K;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const CPt stride = cp_pow_mul(CPt{1, 0}, step, T);
    CPt p = cp_pow_mul(initial, step, t);
#pragma unroll 4
    for (uint32_t k = 0; k < PROG_K; k++) {
        const uint64_t i = t + (uint64_t)k * T;
        if (i >= n) break;
        f(i, p);
        p = cpmul(p, stride);
    }
}
inline uint32_t progression_blocks(uint64_t n, uint32_t threads) {
    return (uint32_t)ceil_div((n + PROG_K - 1) / PROG_K, threads);
}

// Fold-twiddle sources of a domain of n = 2^log_n points: out[i] = coordinate
// of initial * step^(bitrev(2i)). The exponents bitrev(2i) are exactly the
// j < n/2, so the points are made in natural order and stored to
// i = bitrev(j) / 2.
__global__ void fold_sources_pow2_kernel(CPt initial, CPt step, uint32_t* out, uint32_t half_n, uint32_t log_n,
                                         int use_y) {
    progression(initial, step, half_n, [&](uint64_t j, CPt p) {
        out[brev_bits((uint32_t)j, log_n) >> 1] = use_y ? p.y : p.x;
    });
}
__global__ void fold_sources_kernel(CPt initial, CPt step, uint32_t* out, uint32_t half_n, uint32_t log_n,
                                    int use_y) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= half_n) return;
    CPt p = cp_pow_mul(initial, step, brev_bits(i << 1, log_n));
    out[i] = use_y ? p.y : p.x;
}
__global__ void squash_kernel(const uint32_t* in, uint32_t* tw_out, uint32_t* sq_out, uint32_t half_n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= half_n) return;
    uint32_t x = in[2 * i];
    if (tw_out) tw_out[i] = x;
    uint32_t x2 = mmul(x, x);
    sq_out[i] = msub(madd(x2, x2), 1u);
}

// Closed form of the line twiddles: layer l, entry i equals
// x(initial^(2^l) * step^(i * 2^(2l+1))), i.e. the x-coordinate that the
// extract-and-squash chain of circle.rs:394-425 produces for that slot
// (each squash doubles the point and keeps every second one).
// Per-layer starting point, step and word offset of the line twiddles, passed
// by value (at most 31 layers).
struct LineTwiddleMeta {
    CPt base[32];
    CPt step[32];
    uint32_t offs[32];
};
__global__ void line_twiddles_kernel(LineTwiddleMeta m, uint32_t* line, uint64_t n) {
    const uint32_t l = blockIdx.y;  // layer l holds n / 2^(l+1) words
    const uint64_t len = n >> (l + 1);
    uint32_t* dst = line + m.offs[l];
    progression(m.base[l], m.step[l], len, [&](uint64_t i, CPt p) { dst[i] = p.x; });
}
__global__ void circle_twiddles_kernel(CPt initial, CPt step, uint32_t* circle, uint32_t half_n) {
    progression(initial, step, half_n, [&](uint64_t i, CPt p) { circle[i] = p.y; });
}

// Montgomery batch inverse with the outputs of cuda/batch_inverse.cu:8-41.
// That kernel runs one thread per 64-word chunk whose prefix products skip zero
// words but start from the chunk's first word as it is: zeros map to zero,
// other words to their inverses, and a chunk whose first word is zero maps
// entirely to zero (its product is zero). Here thread t of T inverts the BI_K
// words t + k * T, so each load and store of a wave is contiguous and the
// prefix products stay in registers, and the chunk rule is applied per word.
constexpr uint32_t BI_K = 32;
__global__ __launch_bounds__(128) void batch_inverse_kernel(const uint32_t* in, uint32_t* out, uint64_t n) {
    const uint64_t T = (n + BI_K - 1) / BI_K;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    uint32_t v[BI_K], pre[BI_K];
    uint32_t acc = 1;
#pragma unroll
    for (uint32_t k = 0; k < BI_K; k++) {
        const uint64_t i = t + (uint64_t)k * T;
        v[k] = i < n ? in[i] : 0;
        pre[k] = acc;  // product of the non-zero v[0 .. k-1]
        if (v[k]) acc = mmul(acc, v[k]);
    }
    uint32_t iv = minv(acc);  // 1 / product of all non-zero v
#pragma unroll
    for (int k = BI_K - 1; k >= 0; k--) {
        const uint64_t i = t + (uint64_t)k * T;
        if (i < n) out[i] = (v[k] && in[i & ~63ull]) ? mmul(iv, pre[k]) : 0;
        if (v[k]) iv = mmul(iv, v[k]);
    }
}

void batch_inverse(const uint32_t* in, uint32_t* out, uint64_t n, hipStream_t s) {
    if (!n) return;
    batch_inverse_kernel<<<ceil_div((n + BI_K - 1) / BI_K, 128), 128, 0, s>>>(in, out, n);
    check_launch("batch_inverse_kernel");
}

void build_ntt_tables(CPt initial, CPt step, uint32_t log_n, uint32_t* line, uint32_t* circle, uint32_t* iline,
                      uint32_t* icircle, uint32_t offs[32], hipStream_t s) {
    uint64_t n = 1ull << log_n;
    uint32_t n_line = log_n > 0 ? log_n - 1 : 0;
    for (int i = 0; i < 32; i++) offs[i] = 0;
    uint64_t off = 0, cur = n;
    LineTwiddleMeta meta = {};
    HostPt hi = {initial.x, initial.y}, hs = {step.x, step.y};
    for (uint32_t l = 0; l < n_line; l++) {
        offs[l] = (uint32_t)off;
        meta.offs[l] = (uint32_t)off;
        HostPt b = host_repeated_double(hi, l);                 // initial^(2^l)
        HostPt st = host_pow(hs, 1ull << (2 * l + 1));          // step^(2^(2l+1))
        meta.base[l] = {b.x, b.y};
        meta.step[l] = {st.x, st.y};
        off += cur / 2;
        cur /= 2;
    }
    if (n_line) {
        // n - 1 line words in total; the last entry of each layer is unused by
        // the transform but is part of the reference layout.
        line_twiddles_kernel<<<dim3(progression_blocks(n / 2, 256), n_line), 256, 0, s>>>(meta, line, n);
        check_launch("line_twiddles_kernel");
        // Word n-2 of the n-1 word layout belongs to no layer; keep it defined (zero)
        // so the batch inverse below never reads an undefined value.
        check(hipMemsetAsync(line + off, 0, 4 * (n - 1 - off), s), "hipMemsetAsync");
    }
    if (n >= 2) {
        circle_twiddles_kernel<<<progression_blocks(n / 2, 256), 256, 0, s>>>(initial, step, circle, (uint32_t)(n / 2));
        check_launch("circle_twiddles_kernel");
    }
    if (iline && n_line) batch_inverse(line, iline, n - 1, s);
    if (icircle && n >= 2) batch_inverse(circle, icircle, n / 2, s);
}

void build_fold_twiddles(CPt initial, CPt step, uint32_t log_n, bool use_y, uint32_t* out, uint32_t* scratch,
                         hipStream_t s) {
    uint32_t half = (uint32_t)((1ull << log_n) / 2);
    if (!half) return;
    fold_sources_pow2_kernel<<<progression_blocks(half, 256), 256, 0, s>>>(initial, step, scratch, half, log_n,
                                                                           use_y ? 1 : 0);
    check_launch("fold_sources_pow2_kernel");
    batch_inverse(scratch, out, half, s);
}

// ---------------------------------------------------------------- single layer / bit reverse
__global__ void ntt_layer_kernel(uint32_t* d, const uint32_t* tw, uint32_t li, uint32_t half_n, int fwd) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= half_n) return;
    uint32_t stride = 1u << li, h = t >> li, l = t & (stride - 1);
    uint32_t i0 = (h << (li + 1)) + l, i1 = i0 + stride;
    uint32_t v0 = d[i0], v1 = d[i1], w = tw[h];
    if (fwd) {
        uint32_t x = mmul(v1, w);
        d[i0] = madd(v0, x);
        d[i1] = msub(v0, x);
    } else {
        d[i0] = madd(v0, v1);
        d[i1] = mmul(msub(v0, v1), w);
    }
}
__global__ void bit_reverse_kernel(uint32_t* d, uint32_t log_n, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t j = brev_bits(i, log_n);
    if (i < j) {
        uint32_t t = d[i];
        d[i] = d[j];
        d[j] = t;
    }
}

static bool tables_from_abi(const uint32_t* tw, const uint32_t* circ, const uint32_t* h_offs, uint32_t n_line,
                            uint32_t n, NttTables& t) {
    if (!is_pow2(n)) {
        set_error(-2, "circle NTT: n must be a power of two");
        return false;
    }
    t.log_n = ilog2(n);
    uint32_t expect = t.log_n ? t.log_n - 1 : 0;
    if (n_line != expect) {
        set_error(-2, "circle NTT: n_line_layers must be log2(n) - 1");
        return false;
    }
    t.line = tw;
    t.circle = circ;
    for (uint32_t l = 0; l < 32; l++) t.offs[l] = l < n_line ? h_offs[l] : 0;
    return true;
}

}  // namespace vx

using namespace vx;

extern "C" {

void cuda_circle_ntt_evaluate(uint32_t* d_data, const uint32_t* d_twiddles, const uint32_t* d_circle_twids,
                              const uint32_t* h_layer_offsets, const uint32_t* /*h_layer_sizes*/,
                              uint32_t n_line_layers, uint32_t n) {
    NttTables t;
    if (!tables_from_abi(d_twiddles, d_circle_twids, h_layer_offsets, n_line_layers, n, t)) return;
    hipStream_t s = current_stream();
    ntt_forward(d_data, d_data, n, t, s);
    check(hipStreamSynchronize(s), "cuda_circle_ntt_evaluate");
}
void cuda_circle_ntt_interpolate(uint32_t* d_data, const uint32_t* d_itwiddles, const uint32_t* d_circle_itwids,
                                 const uint32_t* h_layer_offsets, const uint32_t* /*h_layer_sizes*/,
                                 uint32_t n_line_layers, uint32_t n) {
    NttTables t;
    if (!tables_from_abi(d_itwiddles, d_circle_itwids, h_layer_offsets, n_line_layers, n, t)) return;
    hipStream_t s = current_stream();
    ntt_inverse(d_data, d_data, t, s);
    check(hipStreamSynchronize(s), "cuda_circle_ntt_interpolate");
}
void cuda_circle_ntt_evaluate_batch(uint32_t** d_columns, const uint32_t* d_twiddles, const uint32_t* d_circle_twids,
                                    const uint32_t* h_layer_offsets, const uint32_t* /*h_layer_sizes*/,
                                    uint32_t n_line_layers, uint32_t n, uint32_t n_cols) {
    NttTables t;
    if (!tables_from_abi(d_twiddles, d_circle_twids, h_layer_offsets, n_line_layers, n, t)) return;
    hipStream_t s = current_stream();
    if (t.log_n) ntt_forward_batch(d_columns, n_cols, t, s);
    check(hipStreamSynchronize(s), "cuda_circle_ntt_evaluate_batch");
}
void cuda_circle_ntt_interpolate_batch(uint32_t** d_columns, const uint32_t* d_itwiddles,
                                       const uint32_t* d_circle_itwids, const uint32_t* h_layer_offsets,
                                       const uint32_t* /*h_layer_sizes*/, uint32_t n_line_layers, uint32_t n,
                                       uint32_t n_cols) {
    NttTables t;
    if (!tables_from_abi(d_itwiddles, d_circle_itwids, h_layer_offsets, n_line_layers, n, t)) return;
    hipStream_t s = current_stream();
    if (t.log_n) ntt_inverse_batch(d_columns, n_cols, t, s);
    check(hipStreamSynchronize(s), "cuda_circle_ntt_interpolate_batch");
}
void cuda_circle_ntt_layer(uint32_t* d_data, const uint32_t* d_twiddles, uint32_t layer_idx, uint32_t n,
                           int32_t forward) {
    uint32_t half = n / 2;
    if (!half) return;
    ntt_layer_kernel<<<ceil_div(half, 256), 256, 0, current_stream()>>>(d_data, d_twiddles, layer_idx, half,
                                                                         forward);
    check_launch("cuda_circle_ntt_layer");
}
int32_t vx_ntt_plan_set(uint32_t log_n, uint32_t n_passes, const uint32_t* m, const uint32_t* c) {
    if (log_n >= 32 || n_passes > 8 || (n_passes && (!m || !c))) {
        set_error(-2, "vx_ntt_plan_set: bad arguments");
        return -1;
    }
    std::vector<Pass> ps;
    uint32_t s0 = 0;
    for (uint32_t i = 0; i < n_passes; i++) {
        ps.push_back({s0, m[i], c[i]});
        s0 += m[i];
    }
    if (n_passes && !plan_valid(log_n, ps)) {
        set_error(-2, "vx_ntt_plan_set: passes must cover every layer with supported shapes");
        return -1;
    }
    std::lock_guard<std::mutex> g(g_plan_mu);
    g_plan[log_n] = ps;  // empty: back to the default split
    return 0;
}

uint32_t vx_ntt_plan_get(uint32_t log_n, uint32_t* m, uint32_t* c) {
    if (log_n >= 32) return 0;
    const auto ps = plan_passes(log_n);
    for (size_t i = 0; i < ps.size() && i < 8; i++) {
        if (m) m[i] = ps[i].m;
        if (c) c[i] = ps[i].c;
    }
    return (uint32_t)ps.size();
}

void vx_ntt_plan_reset(void) {
    std::lock_guard<std::mutex> g(g_plan_mu);
    for (auto& p : g_plan) p.clear();
    for (auto& p : g_lde_plan) p.clear();
}

void vx_ntt_set_lde_bridge(int32_t enable) {
    std::lock_guard<std::mutex> g(g_plan_mu);
    g_lde_bridge = enable != 0;
}

int32_t vx_ntt_lde_plan_set(uint32_t log_n, uint32_t n_passes, const uint32_t* m, const uint32_t* c,
                            const uint32_t* r) {
    if (log_n >= 32 || n_passes > 8 || (n_passes && (!m || !c || !r))) {
        set_error(-2, "vx_ntt_lde_plan_set: bad arguments");
        return -1;
    }
    std::vector<Pass> ps;
    uint32_t s0 = 0;
    for (uint32_t i = 0; i < n_passes; i++) {
        ps.push_back({s0, m[i], c[i], r[i]});
        s0 += m[i];
    }
    if (n_passes && !plan_valid(log_n, ps, true)) {
        set_error(-2, "vx_ntt_lde_plan_set: passes must cover every layer with supported shapes");
        return -1;
    }
    std::lock_guard<std::mutex> g(g_plan_mu);
    g_lde_plan[log_n] = ps;  // empty: the general plan of the size
    return 0;
}

uint32_t vx_ntt_lde_plan_get(uint32_t log_n, uint32_t* m, uint32_t* c, uint32_t* r) {
    if (log_n >= 32) return 0;
    const auto ps = lde_plan_passes(log_n);
    for (size_t i = 0; i < ps.size() && i < 8; i++) {
        if (m) m[i] = ps[i].m;
        if (c) c[i] = ps[i].c;
        if (r) r[i] = ps[i].r;
    }
    return (uint32_t)ps.size();
}

int32_t vx_ntt_autotune(uint32_t log_min, uint32_t log_max, float* best_ms) {
    if (log_min < 2 || log_max < log_min || log_max > 28) {
        set_error(-2, "vx_ntt_autotune: sizes must satisfy 2 <= log_min <= log_max <= 28");
        return -1;
    }
    const uint64_t N = 1ull << log_max;
    // blowup-4 evaluations are timed as the provers run them: a batch of columns
    // (one per proof or per committed column) of about 2^26 output words
    constexpr uint32_t LDE_LOG_WORDS = 26, LDE_MAX_COLS = 1024;
    const uint64_t words = std::max<uint64_t>(N, 1ull << LDE_LOG_WORDS);
    hipStream_t st = nullptr;
    uint32_t *data = nullptr, *line = nullptr, *circle = nullptr, *line2 = nullptr, *circle2 = nullptr;
    uint32_t** d_cols = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate") &&
              check(hipMalloc((void**)&data, 4 * words), "hipMalloc") &&
              check(hipMalloc((void**)&d_cols, sizeof(uint32_t*) * LDE_MAX_COLS), "hipMalloc") &&
              check(hipMalloc((void**)&line2, 4 * N / 4), "hipMalloc") &&
              check(hipMalloc((void**)&circle2, 4 * N / 8 + 4), "hipMalloc") &&
              check(hipMalloc((void**)&line, 4 * N), "hipMalloc") &&
              check(hipMalloc((void**)&circle, 4 * N / 2), "hipMalloc") && check(hipEventCreate(&e0), "hipEvent") &&
              check(hipEventCreate(&e1), "hipEvent");
    for (uint32_t k = log_min; ok && k <= log_max; k++) {
        HostPt hi, hs;
        host_coset(1, k, hi, hs);
        NttTables t;
        t.line = line;
        t.circle = circle;
        t.log_n = k;
        build_ntt_tables(CPt{hi.x, hi.y}, CPt{hs.x, hs.y}, k, line, circle, nullptr, nullptr, t.offs, st);
        const uint32_t n_cols = k >= LDE_LOG_WORDS ? 1u : std::min<uint32_t>(LDE_MAX_COLS, 1u << (LDE_LOG_WORDS - k));
        ok = check(hipMemsetAsync(data, 0x11, 4ull * n_cols << k, st), "hipMemsetAsync");  // canonical M31 words
        std::vector<uint32_t*> h_cols(n_cols);
        for (uint32_t c = 0; c < n_cols; c++) h_cols[c] = data + ((uint64_t)c << k);
        ok = ok && check(hipMemcpyAsync(d_cols, h_cols.data(), sizeof(uint32_t*) * n_cols, hipMemcpyHostToDevice, st),
                         "hipMemcpyAsync") && check(hipStreamSynchronize(st), "sync");
        // tables of a 2^(k-2) transform for the LDE's interpolation (any
        // canonical words time the same; these are the forward ones of that size)
        NttTables ti;
        if (k >= 3) {
            HostPt hi2, hs2;
            host_coset(1, k - 2, hi2, hs2);
            ti.line = line2;
            ti.circle = circle2;
            ti.log_n = k - 2;
            build_ntt_tables(CPt{hi2.x, hi2.y}, CPt{hs2.x, hs2.y}, k - 2, line2, circle2, nullptr, nullptr, ti.offs, st);
        }
        // the general table on one full input, then the blowup-4 table on a batch
        // of LDEs (interpolation with the tuned plan of 2^(k-2), evaluation with
        // the candidate, bridged when the candidate allows)
        for (int lde = 0; lde < 2 && ok; lde++) {
            if (lde && k < 3) break;
            const uint64_t src_n = 1ull << k;
            const auto cands = lde ? lde_plan_candidates(k) : plan_candidates(k);
            float best = 1e30f;
            size_t bi = 0;
            auto run = [&](const std::vector<Pass>& ps) {
                if (lde)
                    lde_run(d_cols, n_cols, ti, t, st, nullptr, true, ps, nullptr);
                else
                    run_ntt_plan(true, data, data, src_n, nullptr, nullptr, 1, t, st, ps, true);
            };
            for (size_t ci = 0; ok && ci < cands.size(); ci++) {
                run(cands[ci]);  // warm-up
                const int reps = 3;
                ok = check(hipEventRecord(e0, st), "hipEventRecord");
                for (int r = 0; r < reps; r++) run(cands[ci]);
                ok = ok && check(hipEventRecord(e1, st), "hipEventRecord") && check(hipEventSynchronize(e1), "sync");
                float ms = 0.f;
                ok = ok && check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime") && peek_error() == 0;
                if (ok && ms / reps < best) {
                    best = ms / reps;
                    bi = ci;
                }
            }
            if (ok) {
                std::lock_guard<std::mutex> g(g_plan_mu);
                (lde ? g_lde_plan : g_plan)[k] = cands[bi];
            }
            if (best_ms && !lde) best_ms[k - log_min] = best;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamSynchronize(st);
    if (data) (void)hipFree(data);
    if (d_cols) (void)hipFree(d_cols);
    if (line2) (void)hipFree(line2);
    if (circle2) (void)hipFree(circle2);
    if (line) (void)hipFree(line);
    if (circle) (void)hipFree(circle);
    if (st) (void)hipStreamDestroy(st);
    return ok ? 0 : -1;
}

void cuda_bit_reverse_m31(uint32_t* data, uint32_t log_n) {
    uint32_t n = 1u << log_n;
    bit_reverse_kernel<<<ceil_div(n, 256), 256, 0, current_stream()>>>(data, log_n, n);
    check_launch("cuda_bit_reverse_m31");
}

void cuda_compute_fold_twiddle_sources(uint32_t ix, uint32_t iy, uint32_t sx, uint32_t sy, uint32_t* output,
                                       uint32_t n, uint32_t log_n, int32_t extract_y) {
    cuda_compute_fold_twiddle_sources_stream(ix, iy, sx, sy, output, n, log_n, extract_y, current_stream());
}
void cuda_compute_fold_twiddle_sources_stream(uint32_t ix, uint32_t iy, uint32_t sx, uint32_t sy, uint32_t* output,
                                              uint32_t n, uint32_t log_n, int32_t extract_y, void* stream) {
    uint32_t half = n / 2;
    if (!half) return;
    fold_sources_kernel<<<ceil_div(half, 256), 256, 0, (hipStream_t)stream>>>(CPt{ix, iy}, CPt{sx, sy}, output,
                                                                              half, log_n, extract_y);
    check_launch("cuda_compute_fold_twiddle_sources");
}
void cuda_batch_inverse_m31(const uint32_t* input, uint32_t* output, uint32_t n) {
    batch_inverse(input, output, n, current_stream());
}
void cuda_batch_inverse_m31_stream(const uint32_t* input, uint32_t* output, uint32_t n, void* stream) {
    batch_inverse(input, output, n, (hipStream_t)stream);
}
void cuda_compute_coset_points(uint32_t ix, uint32_t iy, uint32_t sx, uint32_t sy, uint32_t* ox, uint32_t* oy,
                               uint32_t n) {
    if (!n) return;
    coset_points_kernel<<<ceil_div(n, 256), 256, 0, current_stream()>>>(CPt{ix, iy}, CPt{sx, sy}, ox, oy, n);
    check_launch("cuda_compute_coset_points");
}
void cuda_squash_x(const uint32_t* input, uint32_t* output, uint32_t out_n) {
    if (!out_n) return;
    squash_kernel<<<ceil_div(out_n, 256), 256, 0, current_stream()>>>(input, nullptr, output, out_n);
    check_launch("cuda_squash_x");
}
void cuda_extract_and_squash(const uint32_t* input, uint32_t* tw_out, uint32_t* sq_out, uint32_t half_n) {
    if (!half_n) return;
    squash_kernel<<<ceil_div(half_n, 256), 256, 0, current_stream()>>>(input, tw_out, sq_out, half_n);
    check_launch("cuda_extract_and_squash");
}

}  // extern "C"
