// Blake2s-256 compression for gfx950 (RFC 7693; the reference's single-block
// helper is src/channel.rs:222-298 and its GPU copy cuda/include/blake2s.cuh).
//
// All state lives in VGPRs. Rotations lower to v_alignbit_b32 and the
// three-input adds to v_add3_u32, so one G costs 12 VALU ops; the feed-forward
// is one v_bitop3_b32 per word: a compression is 960 + 8 VALU ops. The round schedule is written out with compile-time message
// indices so no message word is ever addressed dynamically.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vx {

constexpr uint32_t B2S_IV0 = 0x6A09E667u, B2S_IV1 = 0xBB67AE85u, B2S_IV2 = 0x3C6EF372u, B2S_IV3 = 0xA54FF53Au;
constexpr uint32_t B2S_IV4 = 0x510E527Fu, B2S_IV5 = 0x9B05688Cu, B2S_IV6 = 0x1F83D9ABu, B2S_IV7 = 0x5BE0CD19u;
// Parameter block word 0: digest length 32, key length 0, fanout 1, depth 1.
constexpr uint32_t B2S_H0 = B2S_IV0 ^ 0x01010020u;

struct Digest {
    uint32_t w[8];
};

#define VX_ROTR(x, n) __builtin_rotateright32((x), (n))
// x ^ y ^ z in one gfx950 instruction (v_bitop3_b32, truth table 0x96)
#define VX_XOR3(x, y, z) __builtin_amdgcn_bitop3_b32((x), (y), (z), 0x96)
#define VX_G(a, b, c, d, x, y)      \
    do {                            \
        a = a + b + (x);            \
        d = VX_ROTR(d ^ a, 16);     \
        c = c + d;                  \
        b = VX_ROTR(b ^ c, 12);     \
        a = a + b + (y);            \
        d = VX_ROTR(d ^ a, 8);      \
        c = c + d;                  \
        b = VX_ROTR(b ^ c, 7);      \
    } while (0)
#define VX_ROUND(m, s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15) \
    do {                                                                                     \
        VX_G(v0, v4, v8, v12, m[s0], m[s1]);                                                 \
        VX_G(v1, v5, v9, v13, m[s2], m[s3]);                                                 \
        VX_G(v2, v6, v10, v14, m[s4], m[s5]);                                                \
        VX_G(v3, v7, v11, v15, m[s6], m[s7]);                                                \
        VX_G(v0, v5, v10, v15, m[s8], m[s9]);                                                \
        VX_G(v1, v6, v11, v12, m[s10], m[s11]);                                              \
        VX_G(v2, v7, v8, v13, m[s12], m[s13]);                                               \
        VX_G(v3, v4, v9, v14, m[s14], m[s15]);                                               \
    } while (0)

// h <- compress(h, m, t, f). m must be indexable with constants only.
__device__ __forceinline__ void b2s_compress(uint32_t h[8], const uint32_t m[16], uint32_t t, uint32_t f) {
    uint32_t v0 = h[0], v1 = h[1], v2 = h[2], v3 = h[3], v4 = h[4], v5 = h[5], v6 = h[6], v7 = h[7];
    uint32_t v8 = B2S_IV0, v9 = B2S_IV1, v10 = B2S_IV2, v11 = B2S_IV3;
    uint32_t v12 = B2S_IV4 ^ t, v13 = B2S_IV5, v14 = B2S_IV6 ^ f, v15 = B2S_IV7;
    VX_ROUND(m, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
    VX_ROUND(m, 14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3);
    VX_ROUND(m, 11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4);
    VX_ROUND(m, 7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8);
    VX_ROUND(m, 9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13);
    VX_ROUND(m, 2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9);
    VX_ROUND(m, 12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11);
    VX_ROUND(m, 13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10);
    VX_ROUND(m, 6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5);
    VX_ROUND(m, 10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0);
    h[0] = VX_XOR3(h[0], v0, v8);
    h[1] = VX_XOR3(h[1], v1, v9);
    h[2] = VX_XOR3(h[2], v2, v10);
    h[3] = VX_XOR3(h[3], v3, v11);
    h[4] = VX_XOR3(h[4], v4, v12);
    h[5] = VX_XOR3(h[5], v5, v13);
    h[6] = VX_XOR3(h[6], v6, v14);
    h[7] = VX_XOR3(h[7], v7, v15);
}

__device__ __forceinline__ void b2s_init(uint32_t h[8]) {
    h[0] = B2S_H0;
    h[1] = B2S_IV1;
    h[2] = B2S_IV2;
    h[3] = B2S_IV3;
    h[4] = B2S_IV4;
    h[5] = B2S_IV5;
    h[6] = B2S_IV6;  // no personalisation on any hash of this path (cuda/blake2s.cu:9-10)
    h[7] = B2S_IV7;
}

// Single-block hash of up to 16 words (t = 4 * n_words), the Merkle leaf rule
// (cuda/blake2s.cu:45-60, merkle.rs:538-556). Words at index >= n_words must be 0.
__device__ __forceinline__ void b2s_leaf(const uint32_t m[16], uint32_t n_words, uint32_t out[8]) {
    b2s_init(out);
    b2s_compress(out, m, 4u * n_words, 0xFFFFFFFFu);
}

// Merkle node: H(left || right), t = 64 (cuda/blake2s.cu:65-89).
__device__ __forceinline__ void b2s_node(const uint32_t l[8], const uint32_t r[8], uint32_t out[8]) {
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        m[i] = l[i];
        m[8 + i] = r[i];
    }
    b2s_init(out);
    b2s_compress(out, m, 64u, 0xFFFFFFFFu);
}

// Leaf of one M31 word (single-column trace tree, t = 4).
__device__ __forceinline__ void b2s_leaf1(uint32_t v, uint32_t out[8]) {
    uint32_t m[16] = {v, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    b2s_init(out);
    b2s_compress(out, m, 4u, 0xFFFFFFFFu);
}

// Leaf of one QM31 (SoA4 trees, t = 16).
__device__ __forceinline__ void b2s_leaf4(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t out[8]) {
    uint32_t m[16] = {a, b, c, d, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    b2s_init(out);
    b2s_compress(out, m, 16u, 0xFFFFFFFFu);
}

// ---------------------------------------------------------------- quad-cooperative form
// Four lanes (a quad: lanes 4q..4q+3 of a wave) compute one compression; lane
// l holds column l of the 4x4 state (v[l], v[4+l], v[8+l], v[12+l]). The
// diagonal step rotates rows 1..3 by 1..3 lanes with DPP quad permutes and
// rotates them back afterwards, so a compression is ~400 instructions per lane
// instead of ~980: used where a tree level is too narrow to fill a wave, cutting
// its latency. The message lives in LDS (16 words per quad); every lane of the
// quad holds all 16 words in registers and selects the ones its two G
// functions of each round use.

// Message schedule of round r (RFC 7693, section 2.7).
constexpr uint8_t B2S_SIGMA[10][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0},
};

// value from lane (l + k) % 4 of the same quad. A quad permute reads only lanes
// of its own quad, so no lane ever reads an invalid source and bound_ctrl
// changes nothing in the result; set, it lets the compiler fold the move into
// the add or xor that consumes it (a DPP operand), saving about 40 of the ~450
// instructions of a compression.
__device__ __forceinline__ uint32_t quad_rot1(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x39, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t quad_rot2(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t quad_rot3(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x93, 0xF, 0xF, true);
}

// mask ? a : b for a mask of all ones or all zeros: one v_bfi_b32
__device__ __forceinline__ uint32_t b2s_pick(uint32_t mask, uint32_t a, uint32_t b) {
    return __builtin_amdgcn_bitop3_b32(mask, a, b, 0xCA);
}

// Message word sigma[R][2l + J] of quad lane l, chosen among the four
// candidates of the quad with the lane's bit masks (m0: bit 0 of l set, m1:
// bit 1): three selects on registers, independent of the round's state.
template <int R, int J>
__device__ __forceinline__ uint32_t b2s_quad_word(const uint32_t m[16], uint32_t m0, uint32_t m1) {
    constexpr int w0 = B2S_SIGMA[R][J], w1 = B2S_SIGMA[R][2 + J], w2 = B2S_SIGMA[R][4 + J], w3 = B2S_SIGMA[R][6 + J];
    return b2s_pick(m1, b2s_pick(m0, m[w3], m[w2]), b2s_pick(m0, m[w1], m[w0]));
}

template <int R>
__device__ __forceinline__ void b2s_quad_round(const uint32_t m[16], uint32_t m0, uint32_t m1, uint32_t& a,
                                               uint32_t& b, uint32_t& c, uint32_t& d) {
    if constexpr (R < 10) {
        // column step: lane l runs G on column l; diagonal step: on diagonal l
        VX_G(a, b, c, d, (b2s_quad_word<R, 0>(m, m0, m1)), (b2s_quad_word<R, 1>(m, m0, m1)));
        b = quad_rot1(b);
        c = quad_rot2(c);
        d = quad_rot3(d);
        VX_G(a, b, c, d, (b2s_quad_word<R, 8>(m, m0, m1)), (b2s_quad_word<R, 9>(m, m0, m1)));
        b = quad_rot3(b);
        c = quad_rot2(c);
        d = quad_rot1(d);
        b2s_quad_round<R + 1>(m, m0, m1, a, b, c, d);
    }
}

// Lane l: in hl = h[l], hh = h[4+l]; out the same words of the new chaining value.
// msg: the quad's 16 message words (LDS). All four lanes of the quad must be active.
// Every lane reads the whole message into registers before the first round and
// picks the words of its G functions by selects, so no LDS access waits inside
// the dependent chain of the rounds.
__device__ __forceinline__ void b2s_compress_quad(const uint32_t* msg, uint32_t t, uint32_t f, uint32_t& hl,
                                                  uint32_t& hh) {
    const uint32_t l = threadIdx.x & 3;
    const uint32_t m0 = 0u - (l & 1u), m1 = 0u - ((l >> 1) & 1u);
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 16; i++) m[i] = msg[i];
    uint32_t a = hl, b = hh;
    uint32_t c = b2s_pick(m1, b2s_pick(m0, B2S_IV3, B2S_IV2), b2s_pick(m0, B2S_IV1, B2S_IV0));
    uint32_t d = b2s_pick(m1, b2s_pick(m0, B2S_IV7, B2S_IV6 ^ f), b2s_pick(m0, B2S_IV5, B2S_IV4 ^ t));
    b2s_quad_round<0>(m, m0, m1, a, b, c, d);
    hl = VX_XOR3(hl, a, c);
    hh = VX_XOR3(hh, b, d);
}

// Initial chaining words of lane l (IV with the parameter block folded in).
__device__ __forceinline__ void b2s_quad_init(uint32_t& hl, uint32_t& hh) {
    const uint32_t l = threadIdx.x & 3;
    hl = l == 0 ? B2S_H0 : l == 1 ? B2S_IV1 : l == 2 ? B2S_IV2 : B2S_IV3;
    hh = l == 0 ? B2S_IV4 : l == 1 ? B2S_IV5 : l == 2 ? B2S_IV6 : B2S_IV7;
}

}  // namespace vx
