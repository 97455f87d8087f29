/*
 * Oracle — CPU restatement of VortexSTARK's Fibonacci commit path
 * (circle LDE -> Blake2s Merkle commit -> FRI folding -> Fiat-Shamir queries
 * -> decommitment -> KRAK v1 bytes -> verifier).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: the parity tests,
 * smoke() and the CPU-baseline leg of bench.py load it to compare against or
 * to time the CPU path. Product code never links, loads or calls it.
 *
 * Every function restates the reference file:line named in its comment in
 * oracle.cpp. Paths are relative to the VortexSTARK source tree.
 *
 * All arithmetic is exact M31 / CM31 / QM31 (p = 2^31 - 1); buffers are u32.
 */
#ifndef VORTEX_ORACLE_H
#define VORTEX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- threads (CPU baseline) ---- */
void orc_set_threads(uint32_t n); /* 0 = hardware_concurrency */
uint32_t orc_get_threads(void);

/* ---- field (src/field/{m31,cm31,qm31}.rs) ---- */
uint32_t orc_m31_add(uint32_t a, uint32_t b);
uint32_t orc_m31_sub(uint32_t a, uint32_t b);
uint32_t orc_m31_mul(uint32_t a, uint32_t b);
uint32_t orc_m31_inv(uint32_t a);
void orc_cm31_mul(const uint32_t a[2], const uint32_t b[2], uint32_t out[2]);
void orc_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]);
void orc_qm31_inv(const uint32_t a[4], uint32_t out[4]);

/* ---- circle (src/circle.rs) ----
 * kind: 0 = subgroup, 1 = half_coset, 2 = half_odds.
 * out = {initial.x, initial.y, step.x, step.y}. */
void orc_coset(int kind, uint32_t log_size, uint32_t out[4]);
void orc_coset_at(int kind, uint32_t log_size, uint64_t index, uint32_t out[2]);
void orc_circle_mul(const uint32_t a[2], const uint32_t b[2], uint32_t out[2]);
void orc_circle_repeated_double(const uint32_t a[2], uint32_t n, uint32_t out[2]);

/* ---- native-NTT twiddles (circle.rs:357-451, twiddles.cu, batch_inverse.cu) ----
 * coset of size n = 2^log_n given by kind.
 * line: n-1 words, circle: n/2 words, offsets/sizes: log_n-1 words each (may be NULL).
 * iline/icircle: inverses (may be NULL). */
void orc_ntt_twiddles(int kind, uint32_t log_n, uint32_t* line, uint32_t* circle,
                      uint32_t* iline, uint32_t* icircle, uint32_t* offsets, uint32_t* sizes);
void orc_batch_inverse(const uint32_t* in, uint32_t* out, uint64_t n);

/* ---- native NTT (cuda/circle_ntt.cu) ---- */
void orc_ntt_layer(uint32_t* data, const uint32_t* twiddles, uint32_t layer_idx, uint64_t n, int forward);
void orc_ntt_evaluate(uint32_t* data, const uint32_t* line, const uint32_t* circle, uint32_t log_n);
void orc_ntt_interpolate(uint32_t* data, const uint32_t* iline, const uint32_t* icircle, uint32_t log_n);
/* trace (2^log_n) -> eval (2^(log_n+log_blowup)) on half_coset domains */
void orc_lde(const uint32_t* trace, uint32_t log_n, uint32_t log_blowup, uint32_t* eval);

/* ---- Blake2s (RFC 7693; channel.rs:222-298; cuda/include/blake2s.cuh) ---- */
void orc_blake2s(const uint8_t* msg, uint64_t len, uint8_t out[32]);          /* standard multi-block */
void orc_blake2s_single(const uint8_t* msg, uint64_t len, uint8_t out[32]);   /* reference 1-block helper */
void orc_blake2s_compress(uint32_t h[8], const uint32_t m[16], uint32_t t, uint32_t f);

/* ---- Merkle (src/merkle.rs, cuda/blake2s.cu) ----
 * cols: array of n_cols column pointers, each n_leaves words (n_leaves a power of two).
 * leaf = Blake2s(LE32 words of min(n_cols,16) columns). */
void orc_merkle_leaf_hashes(const uint32_t* const* cols, uint32_t n_cols, uint64_t n_leaves, uint32_t* out);
void orc_merkle_node_layer(const uint32_t* children, uint64_t n_parents, uint32_t* parents);
void orc_merkle_root(const uint32_t* const* cols, uint32_t n_cols, uint64_t n_leaves, uint32_t root[8]);
int orc_merkle_verify_path(const uint32_t root[8], const uint32_t leaf[8], uint64_t index,
                           const uint32_t* path, uint32_t depth);

/* ---- Fibonacci AIR (src/air.rs, cuda/constraint.cu) ---- */
void orc_fibonacci_trace(uint32_t a, uint32_t b, uint32_t log_n, uint32_t* out);
void orc_fibonacci_quotient(const uint32_t* eval, uint64_t n, const uint32_t alpha[4],
                            uint32_t* q0, uint32_t* q1, uint32_t* q2, uint32_t* q3);

/* ---- FRI (src/fri.rs, cuda/fri.cu, cuda/twiddles.cu:23-52) ----
 * domain kind as orc_coset; extract_y: 1 = y (circle fold), 0 = x (line fold). */
void orc_fold_twiddles(int kind, uint32_t log_size, int extract_y, uint32_t* out);
void orc_fold_line(const uint32_t* const in[4], const uint32_t* tw, const uint32_t alpha[4],
                   uint64_t half_n, uint32_t* const out[4]);
void orc_fold_circle_into_line(uint32_t* const dst[4], const uint32_t* const src[4], const uint32_t* tw,
                               const uint32_t alpha[4], uint64_t half_n);

/* ---- Channel (src/channel.rs) ---- */
typedef struct {
    uint8_t state[32];
    uint32_t n_draws;
} orc_channel;
void orc_channel_init(orc_channel* ch);
void orc_channel_mix_digest(orc_channel* ch, const uint32_t digest[8]);
void orc_channel_mix_felts(orc_channel* ch, const uint32_t* felts, uint64_t n_felts);
void orc_channel_mix_u64(orc_channel* ch, uint64_t v);
void orc_channel_draw_felt(orc_channel* ch, uint32_t out[4]);
uint64_t orc_channel_draw_number(orc_channel* ch, uint64_t bound);

/* ---- whole proof (src/prover.rs:1721-1978, stark_cli.rs:907-975) ----
 * Returns the KRAK v1 byte length; writes bytes when out != NULL and cap suffices.
 * stats (may be NULL), filled with: [0..8) trace root, [8..16) quotient root,
 * [16] number of FRI layers. */
uint64_t orc_prove_fibonacci(uint32_t a, uint32_t b, uint32_t log_n, uint32_t n_queries,
                             uint32_t log_last_layer, uint8_t* out, uint64_t cap, uint32_t* stats);
/* Commitments of the same proof without its trees or decommitments, in
 * bounded host memory (root-only Merkle commitments, quotient recomputed from
 * the LDE, FRI layers freed as the next is made): roots = trace root,
 * quotient root, then one root per FRI layer (8 words each, log_n + 2 -
 * log_last_layer layers); last_layer = 4 << log_last_layer words; queries =
 * n_queries indices. Up to log_n = 29 (the 2^31-point domain is subgroup(31)). */
int orc_prove_summary(uint32_t a, uint32_t b, uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer,
                      uint32_t* roots, uint32_t* last_layer, uint64_t* queries);
/* NTT twiddles built layer from layer (no all-points arrays); same words as
 * orc_ntt_twiddles. */
void orc_ntt_twiddles_lean(int kind, uint32_t log_n, uint32_t* line, uint32_t* circle, uint32_t* iline,
                           uint32_t* icircle);
/* Phase timing variant used by the CPU baseline: same work as orc_prove_fibonacci
 * from an already generated trace with cached twiddles; returns seconds spent. */
double orc_time_prove_fibonacci(uint32_t a, uint32_t b, uint32_t log_n, uint32_t n_queries,
                                uint32_t log_last_layer, uint32_t reps);
/* Each of `reps` proofs timed on its own (steady_clock) after `warmup` untimed
 * ones, twiddles built once beforehand; secs gets reps entries. */
int orc_time_prove_each(uint32_t a, uint32_t b, uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer,
                        uint32_t warmup, uint32_t reps, double* secs);
/* Verifier (src/verifier.rs:39-310), parameterised by (n_queries, log_last_layer).
 * Returns 0 on accept, a positive reason code on reject. */
int orc_verify_krak(const uint8_t* proof, uint64_t len, uint32_t n_queries, uint32_t log_last_layer);

/* ---- stwo-basis pieces for the wide/lifted commit ---- */
uint32_t orc_lifted_row(uint32_t leaf, uint32_t lifting_log, uint32_t col_log);
/* twiddles of a stwo-basis NTT of size 2^log_n (src/ntt.rs:262-324): 2^(log_n-1) words each */
void orc_stwo_twiddles(uint32_t log_n, uint32_t* fwd, uint32_t* inv_out);
void orc_stwo_ntt_evaluate(uint32_t* data, const uint32_t* twiddles, uint32_t log_n);
void orc_stwo_ntt_interpolate(uint32_t* data, const uint32_t* itwiddles, uint32_t log_n);
/* lifted leaf hashes of columns in the given order (merkle_leaves_lifted.cu) */
void orc_lifted_leaves(const uint32_t* const* cols, const uint32_t* log_sizes, uint32_t n_cols, uint32_t lifting_log,
                       uint32_t* out);
void orc_pack_leaves(const uint32_t* const* in4, uint32_t* const* out64, uint64_t n);
void orc_bit_reverse_qm31(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t log_n);
/* batched wide commitment (stwo LDE + lifted tree + FRI on a beta-combination); log_sizes
 * ascending; returns the number of FRI roots written to fri_roots (8 words each; 0 when the
 * sizes are not ascending); last gets 4 << log_last words. */
uint32_t orc_wide_commit(const uint32_t* const* trace, const uint32_t* log_sizes, uint32_t n_cols, uint32_t blowup,
                         uint32_t log_last, uint32_t* root8, uint32_t* fri_roots, uint32_t* last, uint32_t* beta4);
double orc_time_wide_commit(const uint32_t* const* trace, const uint32_t* log_sizes, uint32_t n_cols, uint32_t blowup,
                            uint32_t log_last);

/* ---- Blake2s proof-of-work grinding (channel.rs:140-168, cuda/grind*.cu) ---- */
uint32_t orc_pow_reduce_word(uint32_t x); /* m31_reduce of one digest word (m31.cuh:9-14) */
/* m31_output: every word of the digest reduced mod p, as grind_ops.rs:157-161 states stwo's
 * reduce_to_m31 does; stwo is absent, so that all eight words are reduced is parity unpinned
 * (the digest and the reduction itself are pinned by hashlib and x mod p). */
void orc_pow_prefix(const uint32_t state[8], uint32_t pow_bits, int m31_output, uint32_t out[8]);
uint32_t orc_pow_zeros(const uint32_t prefixed[8], uint64_t nonce, int m31_output);
/* smallest nonce in [start, start + count) with >= pow_bits zeros, else UINT64_MAX */
uint64_t orc_pow_search(const uint32_t prefixed[8], uint32_t pow_bits, int m31_output, uint64_t start,
                        uint64_t count);

/* ---- stwo DEEP quotient pieces (cuda/quotient.cu, cuda/accumulate_lift.cu) ---- */
void orc_accumulate_numerators(const uint32_t* const* cols, const uint32_t* col_indices, const uint32_t* b,
                               const uint32_t* c, uint32_t n_batch, uint32_t n_rows, uint32_t* out0,
                               uint32_t* out1, uint32_t* out2, uint32_t* out3);
void orc_quotients_combine(const uint32_t* spx, const uint32_t* spy, const uint32_t* fla,
                           const uint32_t* const* n0, const uint32_t* const* n1, const uint32_t* const* n2,
                           const uint32_t* const* n3, const uint32_t* acc_logs, uint32_t n_accs,
                           const uint32_t* dxs, const uint32_t* dys, uint32_t lifting_log, uint32_t n_rows,
                           uint32_t* out0, uint32_t* out1, uint32_t* out2, uint32_t* out3);
void orc_accumulate_lift(uint32_t* col, const uint32_t* curr, uint32_t col_n, uint32_t log_ratio);
void orc_circle_domain_brev(uint32_t log_n, uint32_t* xs, uint32_t* ys);

/* ---- splitmix64 input generator ---- */
uint64_t orc_splitmix64(uint64_t* state);

#ifdef __cplusplus
}
#endif
#endif
