// Fiat-Shamir channel evaluated on the device (src/channel.rs:15-206).
//
// The prover keeps the transcript in device memory and advances it from the
// kernel that produces each Merkle root, so the FRI layer loop never waits for
// the host. Byte layout: the 32-byte state is handled as the 8 little-endian
// words Blake2s produced it as, which is exactly how channel.rs concatenates it.
#pragma once
#include "blake2s.hpp"
#include "field.hpp"

namespace vx {

struct DevChannel {
    uint32_t state[8];
    uint32_t n_draws;
    uint32_t failed;  // set if draw_felt exhausted its retry budget
    uint32_t pad[6];
};

// mix_digest: state = H(state || digest) (channel.rs:30-39)
__device__ __forceinline__ void ch_mix_digest(DevChannel* ch, const uint32_t d[8]) {
    uint32_t m[16], h[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        m[i] = ch->state[i];
        m[8 + i] = d[i];
    }
    b2s_init(h);
    b2s_compress(h, m, 64u, 0xFFFFFFFFu);
#pragma unroll
    for (int i = 0; i < 8; i++) ch->state[i] = h[i];
    ch->n_draws = 0;
}

// squeeze: H(state || LE32(n_draws) || 0x00), 37 bytes (channel.rs:192-199)
__device__ __forceinline__ void ch_squeeze(const uint32_t st[8], uint32_t n_draws, uint32_t out[8]) {
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 8; i++) m[i] = st[i];
    m[8] = n_draws;
#pragma unroll
    for (int i = 9; i < 16; i++) m[i] = 0;
    b2s_init(out);
    b2s_compress(out, m, 37u, 0xFFFFFFFFu);
}

// draw_felt with rejection until all 8 words < 2p (channel.rs:67-86). The
// rejection probability per squeeze is about 2^-29, so 64 tries never run out
// in practice; if they do, `failed` is raised and the host reports an error.
__device__ __forceinline__ void ch_draw_felt(DevChannel* ch, uint32_t out[4]) {
    for (int tries = 0; tries < 64; tries++) {
        uint32_t w[8];
        ch_squeeze(ch->state, ch->n_draws, w);
        ch->n_draws++;
        bool ok = true;
#pragma unroll
        for (int i = 0; i < 8; i++) ok = ok && (w[i] < 2u * P);
        if (ok) {
#pragma unroll
            for (int i = 0; i < 4; i++) out[i] = canon(w[i]);
            return;
        }
    }
    ch->failed = 1;
    for (int i = 0; i < 4; i++) out[i] = 0;
}

// mix_felts: state = Blake2s(state || felts), standard multi-block hashing
// (channel.rs:48-61 via the blake2 crate). words: 4 per QM31.
__device__ __forceinline__ void ch_mix_felts(DevChannel* ch, const uint32_t* words, uint32_t n_words) {
    uint32_t h[8];
    b2s_init(h);
    const uint32_t total_words = 8 + n_words;
    const uint32_t total_bytes = 4 * total_words;
    const uint32_t nb = (total_bytes + 63) / 64;
    for (uint32_t b = 0; b < nb; b++) {
        uint32_t m[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            uint32_t gi = 16 * b + i;
            m[i] = gi < 8 ? ch->state[gi] : (gi < total_words ? words[gi - 8] : 0u);
        }
        bool last = b + 1 == nb;
        b2s_compress(h, m, last ? total_bytes : 64u * (b + 1), last ? 0xFFFFFFFFu : 0u);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) ch->state[i] = h[i];
    ch->n_draws = 0;
}

}  // namespace vx
