// Blake2s proof-of-work nonce search (cuda/grind.cu:13-98,
// cuda/grind_m31_output.cu:14-89; the check itself is src/channel.rs:140-168).
//
// nonce is valid iff the first 16 bytes of Blake2s(prefixed_digest || nonce LE)
// (40 bytes, one block: t = 40, final), read as a little-endian u128, have at
// least pow_bits trailing zeros; the M31 variant reduces those four words mod p
// first. Work per nonce is one compression and no memory: the search is bound
// by VALU issue. The message words 0..7 are the same for every nonce (SGPRs)
// and words 10..15 are zero, so round 0's column step and three of its four
// diagonal G functions are wave-uniform and leave the vector ALU (they run on
// the scalar unit); for pow_bits <= 32 only word 0 of the digest is formed,
// which drops the two diagonal G functions of the last round and part of the
// other two. About 840 VALU ops a nonce remain, against 968 for a full
// compression.
//
// The grid is persistent: one workgroup of 256 per CU, so one wave per SIMD,
// each lane hashing NONCES_PER_LANE nonces a pass. Nonces are dealt
// wave-contiguously: pass k gives wave w the 256 nonces start + k * stride +
// 256 * w + [0, 256). Before each pass a wave reads the current best nonce and
// stops once its next nonce is not below it, so the final value is the
// smallest valid nonce of [start, start + count) (or what result already held,
// if smaller), whatever order the waves ran in: the reference returns the
// first nonce any thread found, a valid one but not always the smallest.
// Why one wave per SIMD: with several, the SIMD issues for its oldest wave
// first, the waves stop advancing evenly, and once a valid nonce is found the
// lagging waves still have to cover every nonce below it: a 26-bit search then
// hashed 1.4x (2 waves per SIMD) to 2.8x (8 waves) the nonces it needed, for a
// few percent more issue rate; one wave with four nonces per lane hashes within
// 1 % of the nonces needed at 96 % of the VALU issue rate (measured on MI355X).
#include <atomic>

#include "blake2s.hpp"
#include "common.hpp"
#include "field.hpp"
#include "transcript.hpp"

namespace vx {
namespace {

constexpr uint32_t GRIND_THREADS = 256;
constexpr uint32_t GRIND_BLOCKS_PER_CU = 1;
constexpr uint32_t NONCES_PER_LANE = 4;
constexpr uint32_t NONCES_PER_BLOCK = GRIND_THREADS * NONCES_PER_LANE;
constexpr uint32_t POW_PREFIX = 0x12345678u;  // Blake2sChannel POW_PREFIX (grind_ops.rs:24)

// m31_reduce of a 32-bit word (cuda/include/m31.cuh:9-14): x mod p.
__device__ __forceinline__ uint32_t reduce_word(uint32_t x) {
    const uint32_t r = (x & P) + (x >> 31);
    return r >= P ? r - P : r;
}

// prefixed_digest = Blake2s(POW_PREFIX LE || 12 zero bytes || state || pow_bits LE),
// 52 bytes (channel.rs:146-151); the M31 channel reduces every word of it
// (grind_ops.rs:157-161). Uniform across the wave: it runs on the scalar unit.
__device__ __forceinline__ void pow_prefix(const uint32_t st[8], uint32_t pow_bits, bool m31, uint32_t out[8]) {
    const uint32_t m[16] = {POW_PREFIX, 0, 0, 0, st[0], st[1], st[2], st[3],
                            st[4],      st[5], st[6], st[7], pow_bits, 0, 0, 0};
    b2s_init(out);
    b2s_compress(out, m, 52u, 0xFFFFFFFFu);
    if (m31)
#pragma unroll
        for (int i = 0; i < 8; i++) out[i] = reduce_word(out[i]);
}

// Whether nonce passes: trailing zeros of digest words 0..3 (u128 LE) >= pow_bits.
// WIDE = false serves pow_bits <= 32 from word 0 alone (mask32 = low pow_bits bits).
template <bool M31, bool WIDE>
__device__ __forceinline__ bool pow_valid(const uint32_t pre[8], uint64_t nonce, uint32_t pow_bits, uint32_t mask32) {
    const uint32_t m[16] = {pre[0], pre[1], pre[2], pre[3], pre[4], pre[5], pre[6], pre[7],
                            (uint32_t)nonce, (uint32_t)(nonce >> 32), 0, 0, 0, 0, 0, 0};
    uint32_t h[8];
    b2s_init(h);
    b2s_compress(h, m, 40u, 0xFFFFFFFFu);
    if (!WIDE) {
        const uint32_t w0 = M31 ? reduce_word(h[0]) : h[0];
        return (w0 & mask32) == 0;
    }
    uint32_t tz = 128;
#pragma unroll
    for (int i = 3; i >= 0; i--) {
        const uint32_t w = M31 ? reduce_word(h[i]) : h[i];
        if (w) tz = 32u * i + (uint32_t)__builtin_ctz(w);
    }
    return tz >= pow_bits;
}

__device__ __forceinline__ uint64_t load_best(uint64_t* result) {
    return __hip_atomic_load(result, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// FROM_STATE: words is a channel state (8 words) and each wave forms the
// prefix itself; otherwise words is the prefixed digest.
template <bool M31, bool WIDE, bool FROM_STATE>
__global__ __launch_bounds__(GRIND_THREADS) void grind_kernel(const uint32_t* __restrict__ words,
                                                              uint64_t* __restrict__ result, uint32_t pow_bits,
                                                              uint64_t start, uint64_t count) {
    uint32_t pre[8];
    if (FROM_STATE) {
        uint32_t st[8];
#pragma unroll
        for (int i = 0; i < 8; i++) st[i] = words[i];
        pow_prefix(st, pow_bits, M31, pre);
    } else {
#pragma unroll
        for (int i = 0; i < 8; i++) pre[i] = words[i];
    }
    const uint32_t mask32 = pow_bits >= 32 ? 0xFFFFFFFFu : (1u << pow_bits) - 1u;
    const uint64_t stride = (uint64_t)gridDim.x * NONCES_PER_BLOCK;
    const uint32_t lane = threadIdx.x & 63u;
    // first offset of this wave: 64 * NONCES_PER_LANE consecutive nonces a pass
    uint64_t w = (uint64_t)blockIdx.x * NONCES_PER_BLOCK + (threadIdx.x >> 6) * (64u * NONCES_PER_LANE);
    if (w >= count) return;
    uint64_t best = load_best(result);
    while (start + w < best) {
        const uint64_t next_best = load_best(result);  // in flight during the hashes
        uint64_t found = ~0ull;
#pragma unroll
        for (int j = NONCES_PER_LANE - 1; j >= 0; j--) {  // the smallest valid one of the lane wins
            const uint64_t off = w + 64u * j + lane;
            if (off < count && pow_valid<M31, WIDE>(pre, start + off, pow_bits, mask32)) found = start + off;
        }
        if (found != ~0ull) __hip_atomic_fetch_min(result, found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w += stride;
        if (w >= count) break;
        best = next_best;
    }
}

__global__ void grind_init_kernel(uint64_t* result) {
    if (threadIdx.x == 0) *result = ~0ull;
}

// mix_u64 of the found nonce (channel.rs:117-128): state = H(state || lo || hi),
// 40 bytes; n_draws = 0. Nothing changes when no nonce was found.
__global__ void grind_mix_kernel(DevChannel* ch, const uint64_t* result) {
    if (threadIdx.x != 0) return;
    const uint64_t nonce = *result;
    if (nonce == ~0ull) return;
    uint32_t m[16] = {};
#pragma unroll
    for (int i = 0; i < 8; i++) m[i] = ch->state[i];
    m[8] = (uint32_t)nonce;
    m[9] = (uint32_t)(nonce >> 32);
    uint32_t h[8];
    b2s_init(h);
    b2s_compress(h, m, 40u, 0xFFFFFFFFu);
#pragma unroll
    for (int i = 0; i < 8; i++) ch->state[i] = h[i];
    ch->n_draws = 0;
}

uint32_t persistent_blocks() {
    static std::atomic<int> cached[64];
    int dev = 0, cus = 0;
    if (!check(hipGetDevice(&dev), "hipGetDevice") || dev < 0 || dev >= 64) return 0;
    cus = cached[dev].load(std::memory_order_relaxed);
    if (!cus) {
        if (!check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute"))
            return 0;
        cached[dev].store(cus, std::memory_order_relaxed);
    }
    return (uint32_t)cus * GRIND_BLOCKS_PER_CU;
}

template <bool M31, bool FROM_STATE>
bool launch_search(const uint32_t* words, uint64_t* result, uint32_t pow_bits, uint64_t start, uint64_t count,
                   hipStream_t s, const char* what) {
    if (count > ~0ull - start) count = ~0ull - start;  // nonce 2^64 - 1 is the "none" value
    if (!count) return true;
    const uint32_t cap = persistent_blocks();
    if (!cap) return false;
    const uint64_t need = count / NONCES_PER_BLOCK + (count % NONCES_PER_BLOCK != 0);
    const uint32_t blocks = need < cap ? (uint32_t)need : cap;
    if (pow_bits <= 32)
        grind_kernel<M31, false, FROM_STATE><<<blocks, GRIND_THREADS, 0, s>>>(words, result, pow_bits, start, count);
    else
        grind_kernel<M31, true, FROM_STATE><<<blocks, GRIND_THREADS, 0, s>>>(words, result, pow_bits, start, count);
    return check_launch(what);
}

}  // namespace
}  // namespace vx

using namespace vx;

extern "C" {

void cuda_grind_pow(const uint32_t* prefixed_digest, uint64_t* result, uint32_t pow_bits, uint64_t batch_offset,
                    uint32_t n_threads) {
    if (!prefixed_digest || !result) {
        set_error(hipErrorInvalidValue, "cuda_grind_pow: null pointer");
        return;
    }
    launch_search<false, false>(prefixed_digest, result, pow_bits, batch_offset, n_threads, current_stream(),
                                "cuda_grind_pow");
}

void cuda_grind_pow_m31_output(const uint32_t* prefixed_digest, uint64_t* result, uint32_t pow_bits,
                               uint64_t batch_offset, uint32_t n_threads) {
    if (!prefixed_digest || !result) {
        set_error(hipErrorInvalidValue, "cuda_grind_pow_m31_output: null pointer");
        return;
    }
    launch_search<true, false>(prefixed_digest, result, pow_bits, batch_offset, n_threads, current_stream(),
                               "cuda_grind_pow_m31_output");
}

int32_t vx_grind_blake2s(void* channel, uint32_t pow_bits, int32_t m31_output, uint64_t start, uint64_t count,
                         uint64_t* result, int32_t mix, void* stream) {
    if (!channel || !result) {
        set_error(hipErrorInvalidValue, "vx_grind_blake2s: null pointer");
        return -1;
    }
    if (pow_bits > 128) {  // no 128-bit value has more trailing zeros: the search could never end
        set_error(-2, "vx_grind_blake2s: pow_bits above 128 can never be met");
        return -1;
    }
    hipStream_t s = (hipStream_t)stream;
    grind_init_kernel<<<1, 64, 0, s>>>(result);
    if (!check_launch("grind_init_kernel")) return -1;
    const uint32_t* st = ((const DevChannel*)channel)->state;
    const bool ok = m31_output
                        ? launch_search<true, true>(st, result, pow_bits, start, count, s, "vx_grind_blake2s")
                        : launch_search<false, true>(st, result, pow_bits, start, count, s, "vx_grind_blake2s");
    if (!ok) return -1;
    if (mix) {
        grind_mix_kernel<<<1, 64, 0, s>>>((DevChannel*)channel, result);
        if (!check_launch("grind_mix_kernel")) return -1;
    }
    return 0;
}

}  // extern "C"
