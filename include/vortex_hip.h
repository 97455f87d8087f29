/*
 * libvortex_hip — MI355X (gfx950) commit pipeline for VortexSTARK's
 * Fibonacci prover: circle LDE -> Blake2s Merkle commit -> FRI folding.
 *
 * C ABI only: plain pointers, sizes and an opaque stream pointer.
 *
 * Part 1 exports the kernel entry points that VortexSTARK's Rust FFI block
 * binds (src/cuda/ffi.rs). Names, argument order, argument meaning and the
 * host-pointer conventions are identical, so that block links against this
 * library unchanged. Each declaration cites the ffi.rs line it replaces.
 *
 * Part 2 replaces the CUDA runtime calls that ffi.rs/device/buffer.rs make
 * (cudaMalloc, cudaMemcpy, streams, ...). These are vx_* functions, not
 * cuda* look-alikes: the Rust side binds them under their own names.
 *
 * Part 3 is the fused device-resident prover (no per-layer host round trips).
 * It produces the same KRAK v1 proof bytes as the reference's prove() +
 * serialize_proof (src/prover.rs:1721-1978, src/bin/stark_cli.rs:907-975).
 *
 * Part 4 is the batched wide commitment: many M31 columns of mixed sizes
 * through the stwo-basis LDE, one lifted Blake2s tree and fused FRI
 * fold+commit layers.
 *
 * Part 5 is a proof server: continuous batching of proof jobs of mixed sizes
 * under an HBM budget, with per-job latency.
 *
 * Conventions shared by every entry point:
 *  - device buffers are allocated by the caller; kernels never allocate,
 *    except the vx_prover_* and vx_wide_* objects, which own their workspace;
 *  - "(host)" arguments are read before the call returns;
 *  - launches go to the calling thread's current stream (vx_set_stream,
 *    default: the null stream) unless a stream argument is given;
 *  - functions the reference ends with a device synchronisation (the
 *    cuda_circle_ntt_* family) are complete when they return; the rest are
 *    stream ordered;
 *  - errors never abort: the first failure is kept per thread and read
 *    with vx_get_last_error() / vx_get_last_error_message().
 */
#ifndef VORTEX_HIP_H
#define VORTEX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ===================== Part 1: reference kernel ABI ====================== */

/* Elementwise M31 ops (ffi.rs:212-213, cuda/fields.cu:32-43). */
void cuda_m31_add(const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t n);
void cuda_m31_mul(const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t n);
/* Device-wide synchronisation (ffi.rs:214). */
void cuda_device_sync(void);

/* Native circle NTT (ffi.rs:243-293, cuda/circle_ntt.cu:181-334).
 * h_layer_offsets / h_layer_sizes are (host) arrays of n_line_layers words. */
void cuda_circle_ntt_evaluate(uint32_t* d_data, const uint32_t* d_twiddles, const uint32_t* d_circle_twids,
                              const uint32_t* h_layer_offsets, const uint32_t* h_layer_sizes,
                              uint32_t n_line_layers, uint32_t n);                                 /* ffi.rs:243 */
void cuda_circle_ntt_interpolate(uint32_t* d_data, const uint32_t* d_itwiddles, const uint32_t* d_circle_itwids,
                                 const uint32_t* h_layer_offsets, const uint32_t* h_layer_sizes,
                                 uint32_t n_line_layers, uint32_t n);                              /* ffi.rs:253 */
void cuda_circle_ntt_evaluate_batch(uint32_t** d_columns, const uint32_t* d_twiddles,
                                    const uint32_t* d_circle_twids, const uint32_t* h_layer_offsets,
                                    const uint32_t* h_layer_sizes, uint32_t n_line_layers, uint32_t n,
                                    uint32_t n_cols);                                              /* ffi.rs:263 */
void cuda_circle_ntt_interpolate_batch(uint32_t** d_columns, const uint32_t* d_itwiddles,
                                       const uint32_t* d_circle_itwids, const uint32_t* h_layer_offsets,
                                       const uint32_t* h_layer_sizes, uint32_t n_line_layers, uint32_t n,
                                       uint32_t n_cols);                                           /* ffi.rs:274 */
void cuda_circle_ntt_layer(uint32_t* d_data, const uint32_t* d_twiddles, uint32_t layer_idx, uint32_t n,
                           int32_t forward);                                                       /* ffi.rs:285 */
void cuda_bit_reverse_m31(uint32_t* data, uint32_t log_n);                                         /* ffi.rs:293 */

/* FRI folds, SoA QM31 columns (ffi.rs:307-323, cuda/fri.cu:82-121).
 * alpha / alpha_sq are (host) arrays of 4 words. */
void cuda_fold_line_soa(const uint32_t* in0, const uint32_t* in1, const uint32_t* in2, const uint32_t* in3,
                        const uint32_t* twiddles, uint32_t* out0, uint32_t* out1, uint32_t* out2,
                        uint32_t* out3, const uint32_t* alpha, uint32_t half_n);                   /* ffi.rs:307 */
void cuda_fold_circle_into_line_soa(uint32_t* dst0, uint32_t* dst1, uint32_t* dst2, uint32_t* dst3,
                                    const uint32_t* src0, const uint32_t* src1, const uint32_t* src2,
                                    const uint32_t* src3, const uint32_t* twiddles, const uint32_t* alpha,
                                    const uint32_t* alpha_sq, uint32_t half_n);                    /* ffi.rs:315 */

/* Twiddles (ffi.rs:327-363, cuda/twiddles.cu, cuda/batch_inverse.cu). */
void cuda_compute_fold_twiddle_sources(uint32_t initial_x, uint32_t initial_y, uint32_t step_x, uint32_t step_y,
                                       uint32_t* output, uint32_t n, uint32_t log_n,
                                       int32_t extract_y);                                         /* ffi.rs:327 */
void cuda_batch_inverse_m31(const uint32_t* input, uint32_t* output, uint32_t n);                  /* ffi.rs:335 */
void cuda_compute_fold_twiddle_sources_stream(uint32_t initial_x, uint32_t initial_y, uint32_t step_x,
                                              uint32_t step_y, uint32_t* output, uint32_t n, uint32_t log_n,
                                              int32_t extract_y, void* stream);                    /* ffi.rs:337 */
void cuda_batch_inverse_m31_stream(const uint32_t* input, uint32_t* output, uint32_t n, void* stream); /* :346 */
void cuda_compute_coset_points(uint32_t initial_x, uint32_t initial_y, uint32_t step_x, uint32_t step_y,
                               uint32_t* output_x, uint32_t* output_y, uint32_t n);                /* ffi.rs:348 */
void cuda_squash_x(const uint32_t* input, uint32_t* output, uint32_t out_n);                        /* ffi.rs:355 */
void cuda_extract_and_squash(const uint32_t* input, uint32_t* twiddle_out, uint32_t* squash_out,
                             uint32_t half_n);                                                     /* ffi.rs:357 */

/* Fibonacci quotient and zero pad (ffi.rs:367-396, cuda/constraint.cu:94-153).
 * alpha is a (host) array of 4 words. */
void cuda_fibonacci_quotient(const uint32_t* trace, uint32_t* out0, uint32_t* out1, uint32_t* out2,
                             uint32_t* out3, const uint32_t* alpha, uint32_t n);                   /* ffi.rs:367 */
void cuda_zero_pad(const uint32_t* src, uint32_t* dst, uint32_t src_n, uint32_t dst_n);             /* ffi.rs:374 */
void cuda_fibonacci_quotient_chunk(const uint32_t* trace, uint32_t* out0, uint32_t* out1, uint32_t* out2,
                                   uint32_t* out3, const uint32_t* alpha, uint32_t offset, uint32_t chunk_n,
                                   uint32_t global_n);                                             /* ffi.rs:381 */
void cuda_fibonacci_quotient_chunk_stream(const uint32_t* trace, uint32_t* out0, uint32_t* out1,
                                          uint32_t* out2, uint32_t* out3, const uint32_t* alpha,
                                          uint32_t offset, uint32_t chunk_n, uint32_t global_n,
                                          void* stream);                                           /* ffi.rs:390 */

/* Blake2s Merkle (ffi.rs:643-697, cuda/blake2s.cu:522-632). Hashes are 8 words. */
void cuda_merkle_hash_leaves(const uint32_t* const* columns, uint32_t* hashes, uint32_t n_cols,
                             uint32_t n_leaves);                                                   /* ffi.rs:643 */
void cuda_merkle_hash_nodes(const uint32_t* children, uint32_t* parents, uint32_t n_parents);      /* ffi.rs:650 */
void cuda_merkle_reduce_to_root(const uint32_t* nodes, uint32_t* root_out, uint32_t n_nodes);      /* ffi.rs:656 */
void cuda_merkle_commit_small_soa4(const uint32_t* col0, const uint32_t* col1, const uint32_t* col2,
                                   const uint32_t* col3, uint32_t* root_out, uint32_t n_leaves);   /* ffi.rs:662 */
void cuda_merkle_hash_leaves_merge_soa4(const uint32_t* col0, const uint32_t* col1, const uint32_t* col2,
                                        const uint32_t* col3, uint32_t* parents, uint32_t n_leaves); /* :669 */
void cuda_merkle_tiled_soa4(const uint32_t* col0, const uint32_t* col1, const uint32_t* col2,
                            const uint32_t* col3, uint32_t* subtree_roots, uint32_t n_leaves);     /* ffi.rs:676 */
void cuda_merkle_tiled_soa4_stream(const uint32_t* col0, const uint32_t* col1, const uint32_t* col2,
                                   const uint32_t* col3, uint32_t* subtree_roots, uint32_t n_leaves,
                                   void* stream);                                                  /* ffi.rs:683 */
void cuda_merkle_tiled_generic(const uint32_t* const* columns, uint32_t* subtree_roots, uint32_t n_cols,
                               uint32_t n_leaves);                                                 /* ffi.rs:691 */

/* stwo-basis circle NTT over the flat twiddle buffer of StwoNttCache::new
 * (ffi.rs:234-239, cuda/circle_ntt_stwo.cu:268-402; twiddles ntt.rs:262-324):
 * evaluate maps coefficients to values on the canonic domain in bit-reversed
 * order, interpolate inverts it including the 1/n scale. Complete on return.
 * n must be a power of two >= 4 (n = 2 would read past the 1-word buffer). */
void cuda_stwo_ntt_evaluate(uint32_t* d_data, const uint32_t* d_twiddles, uint32_t n);           /* ffi.rs:236 */
void cuda_stwo_ntt_interpolate(uint32_t* d_data, const uint32_t* d_itwiddles, uint32_t n);       /* ffi.rs:238 */

/* Lifted multi-column Blake2s leaves (ffi.rs:769-776, cuda/merkle_leaves_lifted.cu):
 * col_ptrs is a device array of column pointers, schedule a device array of
 * n_chunks LeafHashChunk {u32 col_indices[16]; u32 n_cols; u32 log_size} (72 bytes
 * each, at most 64 chunks). Leaf i hashes the columns in schedule order, each at
 * its lifted row. output_hashes receives 8 words per leaf. */
void cuda_build_leaves_lifted(const uint32_t* const* col_ptrs, const void* schedule, uint32_t n_chunks,
                              uint32_t lifting_log_size, uint32_t* output_hashes, uint32_t n_leaves); /* :769 */

/* QM31 bit reversal out of place (4 words per element) and the 4 -> 64 column
 * leaf packing of stwo's PackLeavesOps (ffi.rs:956-965, cuda/bit_reverse_wide.cu,
 * cuda/pack_leaves.cu): output_cols[c*16 + p][i] = input_cols[c][16 i + p]. */
void cuda_bit_reverse_qm31(const uint32_t* inp, uint32_t* out, uint32_t n, uint32_t log_n);       /* ffi.rs:956 */
void cuda_pack_leaves(const uint32_t* const* input_col_ptrs, uint32_t* const* output_col_ptrs,
                      uint32_t n);                                                                 /* ffi.rs:960 */

/* stwo DEEP quotient (ffi.rs:840-897, 957-958; cuda/quotient.cu, cuda/accumulate_lift.cu;
 * callers quotient_ops.rs:35-223, accumulation_ops.rs:46-88). All pointers are device
 * pointers; col_ptrs / numer_ptrs* are device arrays of device pointers; QM31 values
 * are 4 M31 words {a.re, a.im, b.re, b.im} of a + b u.
 * accumulate_numerators: out*[row] = sum_i (c_i * col_ptrs[col_indices[i]][row] - b_i)
 *   (SoA QM31 out, n_rows each).
 * compute_quotients_combine: for each row of the 2^lifting_log_size domain (points
 *   (domain_xs[row], domain_ys[row]), bit-reversed canonic order),
 *   out*[row] = sum_j (numer_j[lift_j(row)] - first_linear_acc_j * y) / den_j, with
 *   den_j = (Re(px_j) - x) Im(py_j) - (Re(py_j) - y) Im(px_j) in CM31 (a zero den
 *   contributes 0) and lift_j(row) = ((row >> (lifting_log_size - acc_log_sizes[j] + 1)) << 1)
 *   + (row & 1).
 * accumulate_lift: col[i] += curr[((i >> (log_ratio + 1)) << 1) | (i & 1)] for i < col_n,
 *   col and curr arrays of QM31 elements (4 consecutive words each).
 * Asynchronous on the stream set with vx_set_stream. */
void cuda_accumulate_numerators(const uint32_t* const* col_ptrs, const uint32_t* col_indices,
                                const uint32_t* b_coeffs, const uint32_t* c_coeffs, uint32_t n_batch_cols,
                                uint32_t n_rows, uint32_t* out0, uint32_t* out1, uint32_t* out2,
                                uint32_t* out3);                                                   /* ffi.rs:850 */
void cuda_compute_quotients_combine(const uint32_t* sample_points_x, const uint32_t* sample_points_y,
                                    const uint32_t* first_linear_acc, const uint32_t* const* numer_ptrs0,
                                    const uint32_t* const* numer_ptrs1, const uint32_t* const* numer_ptrs2,
                                    const uint32_t* const* numer_ptrs3, const uint32_t* acc_log_sizes,
                                    uint32_t n_accs, const uint32_t* domain_xs, const uint32_t* domain_ys,
                                    uint32_t lifting_log_size, uint32_t n_rows, uint32_t* out0,
                                    uint32_t* out1, uint32_t* out2, uint32_t* out3);               /* ffi.rs:878 */
void cuda_accumulate_lift(uint32_t* col, const uint32_t* curr, uint32_t col_n, uint32_t log_ratio); /* ffi.rs:958 */

/* Fused DEEP quotient (no reference counterpart: QuotientOps::accumulate_numerators
 * for every (point, size) batch followed by compute_quotients_and_combine,
 * quotient_ops.rs:35-223, in one launch; the same words as that pipeline). cols:
 * host array of n_cols device pointers to bit-reversed M31 evaluations of
 * 2^col_log_sizes[c] rows (1 <= log <= lifting_log_size); points: host,
 * n_points (1..4) x 8 words (x then y, QM31); line_coeffs: host,
 * n_points x n_cols x 12 words (a, b, c of column c at point p; all zero for a
 * column not sampled there); domain_xs/ys: device, the lifting domain's points
 * (vx_circle_domain_points). out*: device SoA QM31, 2^lifting_log_size rows.
 *   out[row] = sum_p (sum_c (c_pc col_c[lift_c(row)] - b_pc) - (sum_c a_pc) y) / den_p(row)
 * Asynchronous on stream. A vx_deep stages each call's metadata in one of two
 * pinned/device slots in turn: a call waits on the host only for the upload of
 * the call two back, and its stream for that call's kernel. */
typedef struct vx_deep vx_deep;
vx_deep* vx_deep_create(uint32_t max_cols, uint32_t max_points);
int32_t vx_deep_quotient(vx_deep* d, const uint32_t* const* cols, const uint32_t* col_log_sizes, uint32_t n_cols,
                         uint32_t n_points, const uint32_t* points, const uint32_t* line_coeffs,
                         uint32_t lifting_log_size, const uint32_t* domain_xs, const uint32_t* domain_ys,
                         uint32_t* out0, uint32_t* out1, uint32_t* out2, uint32_t* out3, void* stream);
void vx_deep_destroy(vx_deep* d);
/* The same quotient with no allocation at all (no reference counterpart; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops). Same arguments and words as vx_deep_quotient without
 * the vx_deep object: the launch metadata (16 n_points + 4 n_cols +
 * 4 n_cols n_points words, at most 8192: e.g. 680 columns at 2 points) is
 * staged through kernel arguments into a fixed 32 KiB device array of the
 * library, one per GPU; larger requests are rejected (-1, error -2). Calls on
 * one GPU share that array, so calls that overlap on one GPU (different
 * streams) race; calls on one stream are ordered. Asynchronous on stream. */
int32_t vx_deep_quotient_noalloc(const uint32_t* const* cols, const uint32_t* col_log_sizes, uint32_t n_cols,
                                 uint32_t n_points, const uint32_t* points, const uint32_t* line_coeffs,
                                 uint32_t lifting_log_size, const uint32_t* domain_xs, const uint32_t* domain_ys,
                                 uint32_t* out0, uint32_t* out1, uint32_t* out2, uint32_t* out3, void* stream);
/* Points of CanonicCoset(log_size).circle_domain() at bit_reverse_index(row) for
 * every row, computed on the device (what quotient_ops.rs:140-147 builds on the host). */
int32_t vx_circle_domain_points(uint32_t log_size, uint32_t* xs, uint32_t* ys, void* stream);

/* Blake2s proof-of-work nonce search (ffi.rs:627-639, 965-976; cuda/grind.cu,
 * cuda/grind_m31_output.cu). Tries the nonces [batch_offset, batch_offset +
 * n_threads) against prefixed_digest (device, 8 words: Blake2s(0x12345678 LE ||
 * 12 zero bytes || channel state || pow_bits LE), reduced word-wise mod p for the
 * M31 channel) and lowers *result (device, set to 2^64 - 1 by the caller before
 * the first batch) to the smallest nonce whose Blake2s(prefixed_digest || nonce
 * LE) has >= pow_bits trailing zeros in its first 16 bytes (u128 LE; the M31
 * variant reduces those four words mod p first). Nonces not below *result are
 * skipped, so a batch after a success does no work. Asynchronous. */
void cuda_grind_pow(const uint32_t* prefixed_digest, uint64_t* result, uint32_t pow_bits, uint64_t batch_offset,
                    uint32_t n_threads);                                                           /* ffi.rs:632 */
void cuda_grind_pow_m31_output(const uint32_t* prefixed_digest, uint64_t* result, uint32_t pow_bits,
                               uint64_t batch_offset, uint32_t n_threads);                         /* ffi.rs:970 */

/* ===================== Part 2: runtime (replaces ffi.rs:6-76 cuda* calls) ==== */

int32_t vx_get_last_error(void);                 /* 0 = none; reading clears it */
const char* vx_get_last_error_message(void);     /* message of the last recorded error */
int32_t vx_device_count(void);
int32_t vx_set_device(int32_t device);
int32_t vx_malloc(void** dev_ptr, size_t bytes);
int32_t vx_free(void* dev_ptr);
int32_t vx_malloc_host(void** host_ptr, size_t bytes); /* pinned */
int32_t vx_free_host(void* host_ptr);
int32_t vx_memcpy_h2d(void* dst, const void* src, size_t bytes);
int32_t vx_memcpy_d2h(void* dst, const void* src, size_t bytes);
int32_t vx_memcpy_d2d(void* dst, const void* src, size_t bytes);
int32_t vx_memset(void* dev_ptr, int32_t value, size_t bytes);
int32_t vx_device_synchronize(void);
int32_t vx_mem_get_info(size_t* free_bytes, size_t* total_bytes);
int32_t vx_stream_create(void** stream);
int32_t vx_stream_synchronize(void* stream);
int32_t vx_stream_destroy(void* stream);
void vx_set_stream(void* stream);                /* per-thread stream for Part 1 launches */
/* HIP events for timing launches on a stream (null: the null stream); elapsed
 * waits for end. Replace the reference's use of host timers around syncs. */
int32_t vx_event_create(void** event);
int32_t vx_event_record(void* event, void* stream);
int32_t vx_event_elapsed_ms(void* start, void* end, float* ms);
int32_t vx_event_destroy(void* event);
void* vx_get_stream(void);

/* Pass plans of the circle NTT (native and stwo basis). A transform of 2^log_n
 * points runs as passes over LDS tiles: a contiguous pass of m[0] layers, then
 * strided passes of m[i] layers moving 2^c[i] contiguous words per row
 * (c in {4, 5, 6}). vx_ntt_autotune times every candidate split of each size on
 * the device and installs the fastest for this process; plan_set installs a
 * given one (n_passes = 0: the default split); plan_get reports the one in use
 * (returns the number of passes). Any plan gives bit-identical results.
 * Evaluations whose input fills only the first quarter (the blowup-4 LDE of
 * the provers) have a table of their own, tuned by the same call on such
 * inputs: its plans may end in a replicating top pass (r[i] = 1), whose two
 * top layers only copy the quarter into all four, so the pass holds 2^(m-2)
 * rows in LDS and transforms them once per quarter; lde_plan_set with
 * n_passes = 0 falls back to the general table, plan_reset clears both. No
 * reference counterpart (tuning of this library's pass structure). */
int32_t vx_ntt_autotune(uint32_t log_min, uint32_t log_max, float* best_ms);
/* Measured int32 VALU issue rate of the current device in T lane-ops/s (a
 * grid of v_add3 / v_xor / v_alignbit chains, the instructions of Blake2s). */
int32_t vx_measure_valu_peak(double* tops, double* ms);
/* Measured latency of one Merkle node hash (a Blake2s compression of two
 * children) in a chain of dependent nodes on a single wave, in ns: four lanes
 * per compression as the narrow tree levels and the transcript run it
 * (*ns_quad), or one lane per compression (*ns_lane). No reference
 * counterpart (a device calibration). Returns 0 or -1. */
int32_t vx_measure_hash_latency(double* ns_quad, double* ns_lane);
int32_t vx_ntt_plan_set(uint32_t log_n, uint32_t n_passes, const uint32_t* m, const uint32_t* c);
uint32_t vx_ntt_plan_get(uint32_t log_n, uint32_t* m, uint32_t* c);
void vx_ntt_plan_reset(void);
int32_t vx_ntt_lde_plan_set(uint32_t log_n, uint32_t n_passes, const uint32_t* m, const uint32_t* c,
                            const uint32_t* r);
uint32_t vx_ntt_lde_plan_get(uint32_t log_n, uint32_t* m, uint32_t* c, uint32_t* r);
/* The provers' LDE runs the interpolation's top pass and the evaluation's
 * replicating top pass as one bridge pass when the two cover the same tile
 * (the quarter's coefficients then never go to memory); enable = 0 runs them
 * as two passes (same results). On by default. No reference counterpart. */
void vx_ntt_set_lde_bridge(int32_t enable);

/* Chunk depth of the Merkle stage kernels (no reference counterpart): a wave
 * hashes 2^c chunks of 128 items before its 64 -> 1 reduction. The depth is
 * picked per launch by the items it covers (all proofs of a batch): a tuned
 * value when one is set, else the largest c leaving 3072 waves. vx_merkle_autotune
 * times whole tree commitments of 2^log items (12 <= log <= 27) at every
 * depth on the device and keeps the fastest; plan returns the tuned depth of
 * 2^log_items items (-1: default rule); plan_set sets one (-1 clears it). Any
 * depth gives the same roots. */
int32_t vx_merkle_autotune(uint32_t log_min, uint32_t log_max, float* best_ms);
int32_t vx_merkle_stage_plan(uint32_t log_items);
int32_t vx_merkle_stage_plan_set(uint32_t log_items, int32_t logc);
void vx_merkle_stage_plan_reset(void);

/* Whole PoW step on the device (GrindOps::grind, grind_ops.rs:125-180, then
 * Channel::mix_u64, channel.rs:117-128): channel is a device transcript whose
 * first 8 words are the channel state (a vx DevChannel, 64 bytes, qualifies).
 * The prefixed digest is formed on the device, *result (device) is set to the
 * smallest valid nonce of [start, start + count) or 2^64 - 1, and with mix != 0
 * a found nonce is mixed into the state (n_draws reset to 0). Asynchronous on
 * stream (null: the default stream); one launch for any count, no host round
 * trip. pow_bits above 128 is rejected (no nonce can meet it). Returns 0, or -1
 * with the error recorded. */
int32_t vx_grind_blake2s(void* channel, uint32_t pow_bits, int32_t m31_output, uint64_t start, uint64_t count,
                         uint64_t* result, int32_t mix, void* stream);

/* ===================== Part 3: fused device-resident prover ================= */

typedef struct vx_prover vx_prover;

/* Builds a prover for traces of 2^log_n rows (blowup 4, eval domain 2^(log_n+2)):
 * twiddle tables (the reference's ProverCache, prover.rs:143-189) and all
 * workspace are allocated here once. n_queries / log_last_layer default to the
 * reference protocol (80, 3) when 0 is passed. 2 <= log_n <= 29; at 29 the
 * evaluation and first fold domain is the whole circle group subgroup(31)
 * (prover.rs:626-630, 754-757) and one prover holds about 91 GB. Returns NULL
 * on failure. */
vx_prover* vx_prover_create(uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer);
/* Same, with room for `capacity` proofs of this size per launch sequence: one
 * launch of each kernel serves every proof of a batch (each proof owns an
 * equal-size arena of device memory; vx_prover_device_bytes reports the total). */
vx_prover* vx_prover_create_batch(uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer, uint32_t capacity);
uint32_t vx_prover_capacity(const vx_prover* p);
/* Device bytes the prover allocated (arenas, twiddles, descriptors). */
uint64_t vx_prover_device_bytes(const vx_prover* p);
/* Device bytes vx_prover_create_batch would allocate for these parameters (0 if
 * unsupported); host arithmetic only, no device call. For callers admitting
 * work against a memory budget; the reference has no counterpart. */
uint64_t vx_prover_estimate_bytes(uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer, uint32_t capacity);
void vx_prover_destroy(vx_prover* p);
/* Stream the prover's launches go to (default: its own stream). */
void* vx_prover_stream(vx_prover* p);
/* Size in bytes of the KRAK v1 proof this prover emits. */
uint64_t vx_prover_proof_size(const vx_prover* p);

/* Enqueues a full proof of the Fibonacci trace at d_trace (2^log_n M31 words in
 * device memory, t[0] = a, t[1] = b) on the prover's stream: LDE, trace and
 * quotient commitments, FRI folds and commitments, Fiat-Shamir transcript,
 * query sampling and decommitment. Nothing waits for the host; the KRAK bytes
 * are left in device memory (vx_prover_proof_device). Returns 0 when enqueued. */
int32_t vx_prover_prove_device(vx_prover* p, const uint32_t* d_trace, uint32_t a, uint32_t b);
/* Batched form: proofs 0 .. count-1 of the traces h_traces[i] (a (host) array of
 * device pointers) with h_ab[2i], h_ab[2i+1] = (a, b) of proof i; count <= capacity. */
int32_t vx_prover_prove_batch_device(vx_prover* p, const uint32_t* const* h_traces, const uint32_t* h_ab,
                                     uint32_t count);
/* Batched form for traces in host memory (the reference's flow: CPU trace in a
 * pinned buffer, H2D, prove; air.rs:136-189, buffer.rs:414-532, prover.rs:62-79):
 * h_traces[i] are host pointers of 2^log_n words. They are copied into one of two
 * device slots on the prover's upload stream while the previous call is still
 * proving, then proved as vx_prover_prove_batch_device. Memory from
 * vx_malloc_host is read asynchronously by DMA (keep it unchanged until the
 * call's proofs are complete); other host memory is copied before this returns.
 * The first call allocates the two slots (2 x capacity x 2^log_n words). */
int32_t vx_prover_prove_batch_host(vx_prover* p, const uint32_t* const* h_traces, const uint32_t* h_ab,
                                   uint32_t count);
/* Batched Fibonacci proofs with the traces generated on the device from (a, b). */
int32_t vx_prove_fibonacci_batch(vx_prover* p, const uint32_t* h_ab, uint32_t count);
/* Replay the launch sequence of a batch as one HIP graph (captured on first use
 * of each batch count, up to 32 counts per prover; further counts are launched
 * directly) instead of individual launches. */
int32_t vx_prover_set_graph(vx_prover* p, int32_t enable);
/* Device address of the KRAK v1 bytes of the last proof (valid after the stream completes). */
const void* vx_prover_proof_device(const vx_prover* p);
const void* vx_prover_proof_device_at(const vx_prover* p, uint32_t index);
/* Waits for the stream and copies the proof to host memory. Returns bytes written or 0. */
uint64_t vx_prover_proof_copy(vx_prover* p, void* host_out, uint64_t cap);
uint64_t vx_prover_proof_copy_at(vx_prover* p, uint32_t index, void* host_out, uint64_t cap);
/* Enqueues on `stream` (NULL: the prover's) copies of proofs 0 .. count-1 to
 * host_proofs (proof_size bytes each, back to back) and, when host_failed is
 * not NULL, each proof's transcript failure flag (non-zero: draw_felt gave up).
 * The copies run on the DMA engines unless vx_prover_set_copy_kernel chose a
 * copy kernel. */
int32_t vx_prover_enqueue_copy(vx_prover* p, uint32_t count, void* host_proofs, uint32_t* host_failed, void* stream);
/* enable = 1: vx_prover_enqueue_copy writes pinned (device-visible) host memory
 * with a kernel on the stream instead of the DMA engines (0, the default); the
 * bytes are the same. No reference counterpart. Returns 0 or -1. */
int32_t vx_prover_set_copy_kernel(vx_prover* p, int32_t enable);
/* Size in bytes of the KRAK v1 proof of a 2^log_n trace (no device needed). */
uint64_t vx_krak_proof_size(uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer);

/* Convenience: generate the Fibonacci trace on device (air.rs:30-45 semantics) and
 * prove, returning the proof bytes on the host (prove(a, b, log_n) + serialize_proof). */
uint64_t vx_prove_fibonacci(vx_prover* p, uint32_t a, uint32_t b, void* host_out, uint64_t cap);

/* Fibonacci trace t0 = a, t1 = b, t(i+2) = t(i) + t(i+1) mod p generated on the device
 * (chunked by 2x2 matrix powers, air.rs:85-189). */
int32_t vx_fibonacci_trace(uint32_t a, uint32_t b, uint32_t log_n, uint32_t* d_out, void* stream);

/* Launch timeline: when enabled, the prover records a HIP event on its stream
 * after every launch of vx_prover_prove_device; entry i > 0 is the time between
 * marks i-1 and i (labels name the phase and kernel). */
int32_t vx_prover_set_timing(vx_prover* p, int32_t enable);
uint32_t vx_prover_timeline_count(const vx_prover* p);
const char* vx_prover_timeline_label(const vx_prover* p, uint32_t i);
float vx_prover_timeline_ms(vx_prover* p, uint32_t i);

/* Phase hooks for tests: device addresses of the prover's last LDE (2^(log_n+2)
 * words) and of FRI layer j (4 SoA columns of 2^(log_n+1-j) words each). */
const uint32_t* vx_prover_eval_device(const vx_prover* p);
const uint32_t* vx_prover_fri_layer_device(const vx_prover* p, uint32_t j, uint32_t column);

/* ===================== Part 4: batched wide commitment ===================== */

/* Builds the stwo flat twiddle buffer of a domain of size 2^log_n on the device
 * (2^(log_n-1) words; ntt.rs:262-324) and, when d_inv is not NULL, its inverse. */
int32_t vx_stwo_twiddles(uint32_t log_n, uint32_t* d_fwd, uint32_t* d_inv, void* stream);

typedef struct vx_wide vx_wide;

/* A wide commitment context for n_cols M31 trace columns of 2^log_sizes[c] rows
 * (ascending, as stwo orders tree columns, merkle_ops.rs:49), LDE blowup
 * 2^log_blowup, FRI down to 2^log_last_layer. Per commitment, on the
 * context's stream:
 *   1. every column: stwo interpolate on its canonic domain, evaluate on the
 *      2^log_blowup larger one (poly_ops.rs:242-260, 496-535);
 *   2. one lifted Blake2s tree over all LDE columns, S = max log size +
 *      log_blowup leaves (merkle_leaves_lifted.cu:26-98); fresh transcript:
 *      mix_digest(root), beta = draw_felt, alpha_0 = draw_felt;
 *   3. f = sum_c beta^c * lifted_c on the 2^S canonic domain, circle-folded
 *      with alpha_0 and 1/y of its points (fri_ops.rs:40-78), committed as SoA4;
 *   4. line folds on half_odds domains (fri.rs:143-224), each committed, each
 *      root mixed and the next alpha drawn, down to 2^log_last_layer values. */
vx_wide* vx_wide_create(uint32_t n_cols, const uint32_t* log_sizes, uint32_t log_blowup, uint32_t log_last_layer);
void vx_wide_destroy(vx_wide* w);
void* vx_wide_stream(vx_wide* w);
uint32_t vx_wide_n_fri_layers(const vx_wide* w);      /* S - log_last_layer */
uint32_t vx_wide_lifting_log_size(const vx_wide* w);  /* S */
/* Enqueues one commitment; h_cols is a (host) array of n_cols device pointers. */
int32_t vx_wide_commit_device(vx_wide* w, const uint32_t* const* h_cols);
/* Waits and copies results: lifted root (8 words), FRI roots (8 per layer), the
 * last layer (4 words per value, interleaved) and beta (4 words); any may be NULL. */
int32_t vx_wide_results(vx_wide* w, uint32_t* root8, uint32_t* fri_roots, uint32_t* last, uint32_t* beta4);
/* Device addresses of LDE column c and of FRI layer j (4 SoA columns). */
const uint32_t* vx_wide_lde_device(const vx_wide* w, uint32_t c);
const uint32_t* vx_wide_fri_layer_device(const vx_wide* w, uint32_t j, uint32_t column);
int32_t vx_wide_set_timing(vx_wide* w, int32_t enable);
uint32_t vx_wide_timeline_count(const vx_wide* w);
const char* vx_wide_timeline_label(const vx_wide* w, uint32_t i);
float vx_wide_timeline_ms(vx_wide* w, uint32_t i);

/* ===================== Part 5: proof server ================================ */

typedef struct vx_server vx_server;

/* A server for Fibonacci proof jobs (n_queries / log_last_layer as for the
 * prover; 0 selects the reference protocol). budget_bytes caps the device
 * memory its batched provers hold (0: 90% of the free memory; a larger budget
 * is clipped to that). Finished proofs wait in a pinned host ring of
 * result_bytes (0: 1 GiB) until their job is released. Not thread-safe: one
 * thread drives a server. */
vx_server* vx_server_create(uint64_t budget_bytes, uint64_t result_bytes, uint32_t n_queries,
                            uint32_t log_last_layer);
void vx_server_destroy(vx_server* s);
/* Queues a job of n_cols columns of 2^log_n rows; column c is the Fibonacci trace
 * with (a, b) = (ab[2c], ab[2c+1]) (host). Returns the job id or -1. */
int64_t vx_server_submit(vx_server* s, uint32_t log_n, uint32_t n_cols, const uint32_t* ab);
/* One scheduling step: retire finished batches, launch batches on idle provers,
 * add provers within the budget. Returns the number of actions taken. */
int32_t vx_server_poll(vx_server* s);
/* Creates up to n_provers (at most 2) provers of size class log_n ahead of
 * traffic, within the HBM budget, so the first batches do not pay for twiddle
 * set-up (the reference's prewarm_eval_pool, prover.rs:62-79). Returns the
 * number created, or -1. */
int32_t vx_server_prewarm(vx_server* s, uint32_t log_n, uint32_t n_provers);
/* Batches replay captured HIP graphs (one per prover and batch count). */
int32_t vx_server_set_graphs(vx_server* s, int32_t enable);
/* Order in which a scheduling step launches the size classes: smallest proofs
 * first (1, the default: short, latency-bound batches are dispatched ahead of
 * the long ones that share their hardware queue) or largest first (0). */
int32_t vx_server_set_order(vx_server* s, int32_t smallest_first);
/* Batches of at least 2^log_rows trace rows (proofs x rows per proof, default
 * 2^23) are big: they run at most max_in_flight at a time (0, the default: no
 * limit) and are staggered (vx_server_set_stagger); smaller batches are not
 * limited. Returns 0 or -1. */
int32_t vx_server_set_big_batches(vx_server* s, uint32_t max_in_flight, uint32_t log_rows);
/* Staggered big batches (1): a big batch's stream waits on the device until
 * the big batch launched before it has finished its heavy part (LDE and the
 * trace, quotient and circle-fold trees), so big batches overlap only in their
 * latency- and PCIe-bound phases (FRI line folds, queries, decommitment, proof
 * copy); this suits a job list submitted at once. 0 (the default): big batches
 * start as soon as launched. No reference counterpart (scheduling of this
 * server). Returns 0 or -1. */
int32_t vx_server_set_stagger(vx_server* s, int32_t enable);
/* Proofs reach the result ring through a copy kernel on each batch's stream
 * (1) or the DMA engines (0, the default); see vx_prover_set_copy_kernel. No
 * reference counterpart. Returns 0 or -1. */
int32_t vx_server_set_copy_kernel(vx_server* s, int32_t enable);
/* Polls until every job has finished, successfully or not; returns 0, or -1
 * when a job failed since the previous wait (the others still complete; the
 * first such failure is the thread's recorded error). */
int32_t vx_server_wait_all(vx_server* s);
/* Fault injection for failure-isolation tests (no reference counterpart):
 * proof `col` of a queued job is reported failed by its batch, through the
 * device-side transcript failure flag a real failure raises (status -3); the
 * other proofs of that batch and every other job are unaffected. The proof
 * must not be on the device yet. Returns 0 or -1. */
int32_t vx_server_inject_fault(vx_server* s, int64_t job, uint32_t col);
uint64_t vx_server_jobs(const vx_server* s);
uint64_t vx_server_completed(const vx_server* s);
/* 1 done, 0 pending, negative: error (-3 transcript, -4 does not fit the budget,
 * -5 result ring full while nothing runs: release finished jobs, -6 no prover
 * of the size class could be created while nothing else held memory). */
int32_t vx_server_job_status(const vx_server* s, int64_t job);
double vx_server_job_latency_ms(const vx_server* s, int64_t job);
uint64_t vx_server_proof_copy(const vx_server* s, int64_t job, uint32_t col, void* out, uint64_t cap);
void vx_server_release_job(vx_server* s, int64_t job);  /* frees a finished job's ring space */
/* Replays arrivals in real time (job i at arrival_s[i] seconds, ascending; its
 * n_cols[i] (a, b) pairs follow each other in ab) and returns when all are done
 * (0, or -1 if a job failed or could not be submitted); latency_ms (optional,
 * n_jobs entries) gets arrival-to-completion times. */
int32_t vx_server_run_trace(vx_server* s, uint32_t n_jobs, const double* arrival_s, const uint32_t* log_n,
                            const uint32_t* n_cols, const uint32_t* ab, double* latency_ms);
/* [0] batches, [1] proofs, [2] jobs done, [3] device bytes held, [4] peak, [5] budget */
void vx_server_stats(const vx_server* s, uint64_t out[6]);

#ifdef __cplusplus
}
#endif
#endif /* VORTEX_HIP_H */
