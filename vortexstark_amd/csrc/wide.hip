// stwo-basis pieces of the commit path and the batched wide commitment.
//
// Reference-ABI entry points (Part 1 of the header):
//   cuda_stwo_ntt_{evaluate,interpolate}  cuda/circle_ntt_stwo.cu:268-402
//   cuda_build_leaves_lifted              cuda/merkle_leaves_lifted.cu:33-120
//   cuda_pack_leaves                      cuda/pack_leaves.cu:22-52
//   cuda_bit_reverse_qm31                 cuda/bit_reverse_wide.cu:21-50
//
// The wide commitment (vx_wide_*) takes many M31 columns of mixed power-of-two
// sizes, all resident in HBM, through:
//
//   stwo interpolate (trace domain) -> stwo evaluate (blowup domain), one
//   launch per pass for all columns of a size -> one lifted Blake2s tree over
//   every column -> beta, alpha_0 from the transcript -> circle fold of
//   f = sum_c beta^c * lifted_c (f formed on the fly from the LDE columns)
//   fused with its commitment -> line folds fused with their commitments.
//
// As in the prover, every root is mixed into the transcript by the kernel that
// computes it and the next challenge is drawn there, so one commitment is a
// single stream of launches with no host round trip.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "commit.hpp"
#include "common.hpp"
#include "fri_tail.hpp"
#include "kernels.hpp"

namespace vx {

// ---------------------------------------------------------------- stwo twiddles
// Flat buffer of src/ntt.rs:262-324 for a domain of size 2^log_n
// (tc = 2^(log_n-1) words): level j holds x(half_odds(k).at(bitrev_{k-1}(r)))
// for r < 2^(k-1), k = log_n - 1 - j, at offset tc - 2^k; the last word is 1.
// half_odds(k).at(i) = G^((4i + 1) * 2^(29-k)) (circle.rs:132-143, k <= 29).
__global__ void stwo_twiddles_kernel(uint32_t* out, uint64_t tc) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= tc) return;
    if (g == tc - 1) {
        out[g] = 1;
        return;
    }
    const uint64_t rem = tc - g;  // in (2^(k-1), 2^k]
    uint32_t k = 0;
    while ((1ull << k) < rem) k++;
    const uint32_t r = (uint32_t)(g - (tc - (1ull << k)));
    const uint32_t idx = brev_bits(r, k - 1);
    const uint32_t e = (uint32_t)(((4ull * idx + 1) << (29 - k)) & 0x7FFFFFFFull);
    out[g] = cp_pow_mul(CPt{1, 0}, CPt{2u, 1268011823u}, e).x;
}

// coordinate of initial * step^bitrev_bits(i), i < n
__global__ void brt_coord_kernel(CPt initial, CPt step, uint32_t* out, uint64_t n, uint32_t bits, int use_y) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const CPt p = cp_pow_mul(initial, step, brev_bits((uint32_t)i, bits));
    out[i] = use_y ? p.y : p.x;
}

void build_stwo_twiddles(uint32_t log_n, uint32_t* fwd, uint32_t* inv, hipStream_t s) {
    const uint64_t tc = (1ull << log_n) / 2;
    if (!tc) return;
    stwo_twiddles_kernel<<<ceil_div(tc, 256), 256, 0, s>>>(fwd, tc);
    check_launch("stwo_twiddles_kernel");
    if (inv) batch_inverse(fwd, inv, tc, s);  // ntt.rs:307-320 (no zero entries)
}

// ---------------------------------------------------------------- lifted leaves (reference ABI)
// cuda/merkle_leaves_lifted.cu:19-24: the schedule entry the Rust side uploads.
struct LeafHashChunk {
    uint32_t col_indices[16];
    uint32_t n_cols;
    uint32_t log_size;
};
constexpr uint32_t LIFT_ABI_MAX_CHUNKS = 64;  // 1024 columns

struct LiftAbiCol {
    const uint32_t* p;
    uint32_t log;
    uint32_t pad;
};

// The block flattens the schedule into LDS (column pointer + log size per
// message word), then each thread hashes one leaf block by block.
__global__ __launch_bounds__(256) void lifted_leaves_kernel(const uint32_t* const* col_ptrs,
                                                           const LeafHashChunk* schedule, uint32_t n_chunks,
                                                           uint32_t S, uint32_t* out, uint64_t n_leaves) {
    __shared__ LiftAbiCol tab[16 * LIFT_ABI_MAX_CHUNKS];
    __shared__ uint32_t total_s;
    const uint32_t tid = threadIdx.x;
    if (tid < n_chunks) {
        uint32_t start = 0;
        for (uint32_t c = 0; c < tid; c++) start += min(schedule[c].n_cols, 16u);
        const LeafHashChunk ch = schedule[tid];
        const uint32_t nc = min(ch.n_cols, 16u);
        for (uint32_t k = 0; k < nc; k++) tab[start + k] = {col_ptrs[ch.col_indices[k]], ch.log_size, 0};
        if (tid == n_chunks - 1) total_s = start + nc;
    }
    if (tid == 0 && n_chunks == 0) total_s = 0;
    __syncthreads();
    const uint32_t total = total_s;
    const uint64_t leaf = (uint64_t)blockIdx.x * blockDim.x + tid;
    if (leaf >= n_leaves) return;
    uint32_t h[8];
    b2s_init(h);
    for (uint32_t c0 = 0;; c0 += 16) {
        const uint32_t nw = total - c0 < 16 ? total - c0 : 16;
        uint32_t m[16];
#pragma unroll
        for (uint32_t k = 0; k < 16; k++) {
            uint32_t v = 0;
            if (k < nw) {
                const LiftAbiCol c = tab[c0 + k];
                uint64_t row = leaf;
                if (c.log != S) row = ((leaf >> (S - c.log + 1)) << 1) | (leaf & 1);
                // a 1-row column would be read at row 1 by the formula; stay inside it
                if (row >> c.log) row = (1ull << c.log) - 1;
                v = c.p[row];
            }
            m[k] = v;
        }
        const bool last = c0 + 16 >= total;
        b2s_compress(h, m, 4 * (c0 + nw), last ? 0xFFFFFFFFu : 0u);
        if (last) break;
    }
    store8(out + 8 * leaf, h);
}

// ---------------------------------------------------------------- pack / bit reverse
// out[c * 16 + p][i] = in[c][16 i + p]: one thread per (c, i) writes one word
// to each of 16 output columns, so the writes of a wave are coalesced.
__global__ void pack_leaves_kernel(const uint32_t* const* in, uint32_t* const* out, uint64_t n) {
    const uint64_t rows = (n + 15) / 16;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 4 * rows) return;
    const uint32_t c = (uint32_t)(g / rows);
    const uint64_t i = g % rows;
    const uint32_t* src = in[c];
#pragma unroll
    for (uint32_t p = 0; p < 16; p++) {
        const uint64_t j = 16 * i + p;
        if (j < n) out[c * 16 + p][i] = src[j];
    }
}

template <bool VEC>
__global__ void bit_reverse_qm31_kernel(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t log_n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t j = brev_bits((uint32_t)i, log_n);
    if (VEC) {
        ((uint4*)out)[j] = ((const uint4*)in)[i];
    } else {
        for (int k = 0; k < 4; k++) out[4 * j + k] = in[4 * i + k];
    }
}

// ---------------------------------------------------------------- wide commitment
// beta^c for c < n (one thread; n is a column count)
__global__ void beta_powers_kernel(const uint32_t* beta, uint32_t n, uint32_t* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const QM31 b = load_alpha(beta);
    QM31 acc = {{1, 0, 0, 0}};
    for (uint32_t c = 0; c < n; c++) {
        for (int k = 0; k < 4; k++) out[4 * c + k] = acc.v[k];
        acc = qmul(acc, b);
    }
}

}  // namespace vx

using namespace vx;

struct vx_wide {
    uint32_t n_cols = 0, blowup = 0, log_last = 0, S = 0, J = 0;
    std::vector<uint32_t> logs;  // trace log sizes, ascending
    struct Group {
        uint32_t log, first, count;
    };
    std::vector<Group> groups;
    hipStream_t stream = nullptr;
    // The LDE of each size group after the first runs on its own side stream,
    // forked from and joined back into `stream` with events, so the small
    // groups' latency-bound passes overlap the largest group's.
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> side_done;
    hipEvent_t fork = nullptr;
    std::vector<void*> allocs;
    uint32_t* tw_fwd = nullptr;  // stwo flat buffer of the 2^S domain
    uint32_t* tw_inv = nullptr;
    uint32_t* tw_circle = nullptr;
    std::vector<uint32_t*> tw_line;  // [j] for line fold j >= 1
    uint32_t* lde = nullptr;
    std::vector<uint64_t> lde_off;
    uint32_t** d_lde_ptrs = nullptr;
    const uint32_t** d_src_ptrs = nullptr;
    const uint32_t** h_src_ptrs = nullptr;  // pinned staging of the caller's column pointers
    std::vector<const uint32_t*> last_src;
    bool src_aligned = false;  // every caller column pointer is 16-byte aligned
    LiftCol* d_lift = nullptr;
    std::vector<uint32_t*> layer;  // J layers, 4 columns each contiguous
    Scratch scratch = {nullptr, nullptr};
    DevChannel* ch = nullptr;
    uint32_t* chal = nullptr;       // [beta, alpha_0, alpha_1, ...], 4 words each
    uint32_t* beta_pows = nullptr;  // 4 words per column
    uint32_t* roots = nullptr;      // lifted root, then J FRI roots
    Timeline tl;
    void mark(const std::string& l) { tl.mark(l, stream); }

    uint64_t layer_size(uint32_t j) const { return 1ull << (S - 1 - j); }
    uint32_t* col(uint32_t j, uint32_t c) const { return layer[j] + (uint64_t)c * layer_size(j); }
    bool alloc(void** p, uint64_t bytes) {
        if (!check(hipMalloc(p, bytes ? bytes : 4), "vx_wide: hipMalloc")) return false;
        allocs.push_back(*p);
        return true;
    }
    ~vx_wide() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipStream_t x : side) (void)hipStreamSynchronize(x);
        for (hipStream_t x : side) (void)hipStreamDestroy(x);
        for (hipEvent_t e : side_done) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        for (void* p : allocs) (void)hipFree(p);
        if (h_src_ptrs) (void)hipHostFree((void*)h_src_ptrs);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
CPt cpt(HostPt h) { return {h.x, h.y}; }
}  // namespace

extern "C" {

void cuda_stwo_ntt_evaluate(uint32_t* d_data, const uint32_t* d_twiddles, uint32_t n) {
    if (!is_pow2(n) || n < 4) {
        set_error(-2, "cuda_stwo_ntt_evaluate: n must be a power of two >= 4");
        return;
    }
    hipStream_t s = current_stream();
    ntt_forward(d_data, d_data, n, stwo_tables(d_twiddles, ilog2(n)), s);
    check(hipStreamSynchronize(s), "cuda_stwo_ntt_evaluate");
}
void cuda_stwo_ntt_interpolate(uint32_t* d_data, const uint32_t* d_itwiddles, uint32_t n) {
    if (!is_pow2(n) || n < 4) {
        set_error(-2, "cuda_stwo_ntt_interpolate: n must be a power of two >= 4");
        return;
    }
    hipStream_t s = current_stream();
    ntt_inverse(d_data, d_data, stwo_tables(d_itwiddles, ilog2(n)), s);
    check(hipStreamSynchronize(s), "cuda_stwo_ntt_interpolate");
}

void cuda_build_leaves_lifted(const uint32_t* const* col_ptrs, const void* schedule, uint32_t n_chunks,
                              uint32_t lifting_log_size, uint32_t* output_hashes, uint32_t n_leaves) {
    if (n_chunks > LIFT_ABI_MAX_CHUNKS) {
        set_error(-2, "cuda_build_leaves_lifted: more than 64 chunks (1024 columns)");
        return;
    }
    if (lifting_log_size > 31 || n_leaves > (1ull << lifting_log_size)) {
        set_error(-2, "cuda_build_leaves_lifted: n_leaves exceeds 2^lifting_log_size");
        return;
    }
    if (!n_leaves) return;
    lifted_leaves_kernel<<<ceil_div(n_leaves, 256), 256, 0, current_stream()>>>(
        col_ptrs, (const LeafHashChunk*)schedule, n_chunks, lifting_log_size, output_hashes, n_leaves);
    check_launch("cuda_build_leaves_lifted");
}

void cuda_pack_leaves(const uint32_t* const* input_col_ptrs, uint32_t* const* output_col_ptrs, uint32_t n) {
    if (!n) return;
    const uint64_t threads = 4 * ((uint64_t)(n + 15) / 16);
    pack_leaves_kernel<<<ceil_div(threads, 256), 256, 0, current_stream()>>>(input_col_ptrs, output_col_ptrs, n);
    check_launch("cuda_pack_leaves");
}

void cuda_bit_reverse_qm31(const uint32_t* inp, uint32_t* out, uint32_t n, uint32_t log_n) {
    if (!n) return;
    if (log_n > 31 || (uint64_t)n > (1ull << log_n)) {
        set_error(-2, "cuda_bit_reverse_qm31: n exceeds 2^log_n");
        return;
    }
    const bool vec = ((uintptr_t)inp % 16 == 0) && ((uintptr_t)out % 16 == 0);
    if (vec)
        bit_reverse_qm31_kernel<true><<<ceil_div(n, 256), 256, 0, current_stream()>>>(inp, out, n, log_n);
    else
        bit_reverse_qm31_kernel<false><<<ceil_div(n, 256), 256, 0, current_stream()>>>(inp, out, n, log_n);
    check_launch("cuda_bit_reverse_qm31");
}

int32_t vx_stwo_twiddles(uint32_t log_n, uint32_t* d_fwd, uint32_t* d_inv, void* stream) {
    if (log_n < 1 || log_n > 29 || !d_fwd) {
        set_error(-2, "vx_stwo_twiddles: log_n must be in [1, 29]");
        return -1;
    }
    build_stwo_twiddles(log_n, d_fwd, d_inv, (hipStream_t)stream);
    return peek_error() == 0 ? 0 : -1;
}

vx_wide* vx_wide_create(uint32_t n_cols, const uint32_t* log_sizes, uint32_t log_blowup, uint32_t log_last_layer) {
    if (!n_cols || !log_sizes || n_cols > 4096) {
        set_error(-2, "vx_wide_create: need 1..4096 columns");
        return nullptr;
    }
    uint32_t mx = 0;
    for (uint32_t c = 0; c < n_cols; c++) {
        if (log_sizes[c] < 2 || (c && log_sizes[c] < log_sizes[c - 1])) {
            set_error(-2, "vx_wide_create: log sizes must be >= 2 and ascending");
            return nullptr;
        }
        mx = std::max(mx, log_sizes[c]);
    }
    const uint32_t S = mx + log_blowup;
    if (log_blowup < 1 || S > 28 || log_last_layer < 1 || log_last_layer >= S || log_last_layer > 10) {
        set_error(-2, "vx_wide_create: unsupported blowup / last layer");
        return nullptr;
    }
    vx_wide* w = new vx_wide();
    w->n_cols = n_cols;
    w->blowup = log_blowup;
    w->log_last = log_last_layer;
    w->S = S;
    w->J = S - log_last_layer;
    w->logs.assign(log_sizes, log_sizes + n_cols);
    for (uint32_t c = 0; c < n_cols; c++) {
        if (w->groups.empty() || w->groups.back().log != log_sizes[c]) w->groups.push_back({log_sizes[c], c, 0});
        w->groups.back().count++;
    }
    bool ok = check(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking), "hipStreamCreate");
    // side streams for the LDE of the smaller size groups (at most three)
    const size_t n_side = std::min<size_t>(3, w->groups.size() - 1);
    ok = ok && (n_side == 0 || check(hipEventCreateWithFlags(&w->fork, hipEventDisableTiming), "hipEventCreate"));
    for (size_t k = 0; ok && k < n_side; k++) {
        hipStream_t x = nullptr;
        hipEvent_t e = nullptr;
        ok = check(hipStreamCreateWithFlags(&x, hipStreamNonBlocking), "hipStreamCreate");
        if (ok) w->side.push_back(x);
        ok = ok && check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        if (ok) w->side_done.push_back(e);
    }
    hipStream_t s = w->stream;
    const uint64_t half = 1ull << (S - 1);
    ok = ok && w->alloc((void**)&w->tw_fwd, 4 * half) && w->alloc((void**)&w->tw_inv, 4 * half) &&
         w->alloc((void**)&w->tw_circle, 4 * half);
    uint32_t* tw_scratch = nullptr;
    ok = ok && w->alloc((void**)&tw_scratch, 4 * half);
    if (ok) {
        build_stwo_twiddles(S, w->tw_fwd, w->tw_inv, s);
        // circle fold on the canonic domain: 1 / y(half_odds(S-1).at(bitrev_{S-1}(i)))
        HostPt hi, hs;
        host_coset(2, S - 1, hi, hs);
        brt_coord_kernel<<<ceil_div(half, 256), 256, 0, s>>>(cpt(hi), cpt(hs), tw_scratch, half, S - 1, 1);
        check_launch("brt_coord_kernel");
        batch_inverse(tw_scratch, w->tw_circle, half, s);
        // line fold j reads layer j-1 on half_odds(S - j) (fri.rs:143-160)
        w->tw_line.assign(w->J, nullptr);
        for (uint32_t j = 1; ok && j < w->J; j++) {
            const uint32_t k = S - j;
            ok = w->alloc((void**)&w->tw_line[j], 4 * ((1ull << k) / 2));
            if (!ok) break;
            HostPt di, ds;
            host_coset(2, k, di, ds);
            build_fold_twiddles(cpt(di), cpt(ds), k, false, w->tw_line[j], tw_scratch, s);
        }
    }
    // LDE columns, contiguous, each 2^(log + blowup) words
    uint64_t total = 0;
    for (uint32_t c = 0; c < n_cols; c++) {
        w->lde_off.push_back(total);
        total += 1ull << (log_sizes[c] + log_blowup);
    }
    ok = ok && w->alloc((void**)&w->lde, 4 * total) && w->alloc((void**)&w->d_lde_ptrs, 8ull * n_cols) &&
         w->alloc((void**)&w->d_src_ptrs, 8ull * n_cols) && w->alloc((void**)&w->d_lift, sizeof(LiftCol) * n_cols);
    ok = ok && check(hipHostMalloc((void**)&w->h_src_ptrs, 8ull * n_cols), "hipHostMalloc");
    w->layer.assign(w->J, nullptr);
    for (uint32_t j = 0; ok && j < w->J; j++) ok = w->alloc((void**)&w->layer[j], 16 * w->layer_size(j));
    const uint64_t mxn = std::max(max_stage_nodes(1ull << S), max_line_quad_nodes(1ull << (S + 1)));
    ok = ok && w->alloc((void**)&w->scratch.a, 32 * (mxn ? mxn : 1)) &&
         w->alloc((void**)&w->scratch.b, 32 * (mxn ? mxn : 1));
    ok = ok && w->alloc((void**)&w->ch, sizeof(DevChannel)) && w->alloc((void**)&w->chal, 16ull * (w->J + 1)) &&
         w->alloc((void**)&w->beta_pows, 16ull * n_cols) && w->alloc((void**)&w->roots, 32ull * (w->J + 1));
    if (ok) {
        std::vector<uint32_t*> ptrs(n_cols);
        std::vector<LiftCol> lift(n_cols);
        for (uint32_t c = 0; c < n_cols; c++) {
            ptrs[c] = w->lde + w->lde_off[c];
            const uint32_t L = log_sizes[c] + log_blowup;
            lift[c] = {ptrs[c], L == S ? 0u : S - L + 1, 0u};
        }
        ok = check(hipMemcpyAsync(w->d_lde_ptrs, ptrs.data(), 8ull * n_cols, hipMemcpyHostToDevice, s), "H2D") &&
             check(hipMemcpyAsync(w->d_lift, lift.data(), sizeof(LiftCol) * n_cols, hipMemcpyHostToDevice, s),
                   "H2D") &&
             check(hipStreamSynchronize(s), "vx_wide_create");
    }
    if (!ok) {
        delete w;
        return nullptr;
    }
    return w;
}

void vx_wide_destroy(vx_wide* w) { delete w; }
void* vx_wide_stream(vx_wide* w) { return w ? (void*)w->stream : nullptr; }
uint32_t vx_wide_n_fri_layers(const vx_wide* w) { return w ? w->J : 0; }
uint32_t vx_wide_lifting_log_size(const vx_wide* w) { return w ? w->S : 0; }
const uint32_t* vx_wide_lde_device(const vx_wide* w, uint32_t c) {
    return (w && c < w->n_cols) ? w->lde + w->lde_off[c] : nullptr;
}
const uint32_t* vx_wide_fri_layer_device(const vx_wide* w, uint32_t j, uint32_t column) {
    return (w && j < w->J && column < 4) ? w->col(j, column) : nullptr;
}

int32_t vx_wide_commit_device(vx_wide* w, const uint32_t* const* h_cols) {
    if (!w || !h_cols) {
        set_error(-2, "vx_wide_commit_device: null argument");
        return -1;
    }
    hipStream_t s = w->stream;
    const uint32_t S = w->S, B = w->blowup;
    w->tl.n = 0;
    w->mark("start");
    if (w->last_src.size() != w->n_cols || !std::equal(w->last_src.begin(), w->last_src.end(), h_cols)) {
        // the staging buffer may still feed an earlier copy: drain first
        if (!check(hipStreamSynchronize(s), "vx_wide_commit_device")) return -1;
        w->src_aligned = true;
        for (uint32_t c = 0; c < w->n_cols; c++) {
            w->h_src_ptrs[c] = h_cols[c];
            if ((uintptr_t)h_cols[c] & 15) w->src_aligned = false;
        }
        if (!check(hipMemcpyAsync(w->d_src_ptrs, w->h_src_ptrs, 8ull * w->n_cols, hipMemcpyHostToDevice, s), "H2D"))
            return -1;
        w->last_src.assign(h_cols, h_cols + w->n_cols);
    }
    const uint64_t tc = 1ull << (S - 1);
    // LDE of every size group: interpolate, then evaluate on the 2^B-times
    // larger domain (a suffix of the 2^S twiddle buffer is the buffer of every
    // smaller domain). The largest group runs on the main stream, every other
    // group on a side stream of its own, forked after the column pointers are
    // uploaded and joined before the commitment.
    const size_t ng = w->groups.size(), ns = std::min(w->side.size(), ng - 1);
    if (ns && !check(hipEventRecord(w->fork, s), "hipEventRecord")) return -1;
    for (size_t k = 0; k < ns; k++)
        if (!check(hipStreamWaitEvent(w->side[k], w->fork, 0), "hipStreamWaitEvent")) return -1;
    for (size_t gi = 0; gi < ng; gi++) {
        const auto& g = w->groups[ng - 1 - gi];  // groups ascend in size: the largest first
        hipStream_t gs = gi == 0 || !ns ? s : w->side[(gi - 1) % ns];
        const NttTables ti = stwo_tables(w->tw_inv + (tc - (1ull << (g.log - 1))), g.log);
        const NttTables te = stwo_tables(w->tw_fwd + (tc - (1ull << (g.log + B - 1))), g.log + B);
        if (B == 2) {
            lde_batch(w->d_lde_ptrs + g.first, g.count, ti, te, gs, w->d_src_ptrs + g.first, w->src_aligned);
        } else {
            ntt_inverse_batch(w->d_lde_ptrs + g.first, g.count, ti, gs, w->d_src_ptrs + g.first, w->src_aligned);
            ntt_forward_batch(w->d_lde_ptrs + g.first, g.count, te, gs, 1ull << g.log, true);
        }
    }
    for (size_t k = 0; k < ns; k++)
        if (!check(hipEventRecord(w->side_done[k], w->side[k]), "hipEventRecord") ||
            !check(hipStreamWaitEvent(s, w->side_done[k], 0), "hipStreamWaitEvent"))
            return -1;
    w->mark("lde");
    std::string phase = "lifted_commit";
    auto mk = [&](const char* what) { w->mark(phase + ":" + what); };
    const LevelStore none = {nullptr, 0, 0};
    TailOut t0 = {w->roots, w->ch, w->chal, 1};
    t0.n_felts = 2;  // beta, then alpha_0
    commit_tree(SrcLifted{w->d_lift, w->n_cols}, 1ull << S, none, t0, w->scratch, s, mk);
    beta_powers_kernel<<<1, 64, 0, s>>>(w->chal, w->n_cols, w->beta_pows);
    check_launch("beta_powers_kernel");
    w->mark("beta_powers");
    phase = "fri0_circle_fold_commit";
    {
        SrcWideFold src = {w->d_lift, w->n_cols, w->beta_pows, w->chal + 4, w->tw_circle,
                           w->col(0, 0), w->col(0, 1), w->col(0, 2), w->col(0, 3)};
        uint32_t* nxt = w->J > 1 ? w->chal + 8 : nullptr;
        commit_tree(src, w->layer_size(0), none, TailOut{w->roots + 8, w->ch, nxt, 0}, w->scratch, s, mk);
    }
    uint32_t js = 1;  // as in the plain prover: larger layers get a stage launch each
    while (js < w->J && w->layer_size(js) > FRI_TAIL_MAX_ITEMS) js++;
    for (uint32_t j = 1; j < js; j++) {
        phase = "fri" + std::to_string(j) + "_line_fold_commit";
        SrcLineFold src = {w->col(j - 1, 0), w->col(j - 1, 1), w->col(j - 1, 2), w->col(j - 1, 3), w->chal + 4 * (j + 1),
                           w->tw_line[j],    w->col(j, 0),     w->col(j, 1),     w->col(j, 2),     w->col(j, 3)};
        uint32_t* nxt = j + 1 < w->J ? w->chal + 4 * (j + 2) : nullptr;
        const TailOut to{w->roots + 8 * (j + 1), w->ch, nxt, 0};
        if (w->layer_size(j) <= LINE_QUAD_MAX)
            commit_line_fold_quad(src, w->layer_size(j), none, to, w->scratch, s, mk);
        else
            commit_tree(src, w->layer_size(j), none, to, w->scratch, s, mk, 1, 0, FRI_TAIL_MAX_ITEMS);
    }
    if (js < w->J) {
        auto layer = [&](uint32_t j) {
            FriTailLayer L = {};
            for (int c = 0; c < 4; c++) {
                L.in[c] = w->col(j - 1, c);
                L.out[c] = w->col(j, c);
            }
            L.tw = w->tw_line[j];
            L.root_dst = w->roots + 8 * (j + 1);
            L.alpha_dst = j + 1 < w->J ? w->chal + 4 * (j + 2) : nullptr;
            L.st = none;
            L.n = (uint32_t)w->layer_size(j);
            return L;
        };
        launch_fri_tail(js, w->J, layer, [&](uint32_t j) { return (const uint32_t*)(w->chal + 4 * (j + 1)); }, w->ch,
                        1, 0, s);
        w->mark("fri_small_layers");
    }
    return peek_error() == 0 ? 0 : -1;
}

int32_t vx_wide_results(vx_wide* w, uint32_t* root8, uint32_t* fri_roots, uint32_t* last, uint32_t* beta4) {
    if (!w) return -1;
    hipStream_t s = w->stream;
    bool ok = true;
    if (root8) ok = ok && check(hipMemcpyAsync(root8, w->roots, 32, hipMemcpyDeviceToHost, s), "D2H");
    if (fri_roots) ok = ok && check(hipMemcpyAsync(fri_roots, w->roots + 8, 32ull * w->J, hipMemcpyDeviceToHost, s), "D2H");
    if (beta4) ok = ok && check(hipMemcpyAsync(beta4, w->chal, 16, hipMemcpyDeviceToHost, s), "D2H");
    std::vector<uint32_t> soa;
    const uint64_t nl = w->layer_size(w->J - 1);
    if (last) {
        soa.resize(4 * nl);
        ok = ok && check(hipMemcpyAsync(soa.data(), w->layer[w->J - 1], 16 * nl, hipMemcpyDeviceToHost, s), "D2H");
    }
    DevChannel ch;
    ok = ok && check(hipMemcpyAsync(&ch, w->ch, sizeof(ch), hipMemcpyDeviceToHost, s), "D2H");
    ok = ok && check(hipStreamSynchronize(s), "vx_wide_results");
    if (!ok) return -1;
    if (last)
        for (uint64_t i = 0; i < nl; i++)
            for (int c = 0; c < 4; c++) last[4 * i + c] = soa[c * nl + i];
    if (ch.failed) {
        set_error(-3, "transcript: draw_felt exhausted its retry budget");
        return -1;
    }
    return 0;
}

int32_t vx_wide_set_timing(vx_wide* w, int32_t enable) {
    if (!w) return -1;
    w->tl.reset(enable != 0);
    return 0;
}
uint32_t vx_wide_timeline_count(const vx_wide* w) { return w ? (uint32_t)w->tl.n : 0; }
const char* vx_wide_timeline_label(const vx_wide* w, uint32_t i) { return w ? w->tl.label(i) : ""; }
float vx_wide_timeline_ms(vx_wide* w, uint32_t i) { return w ? w->tl.ms(i) : 0.f; }

}  // extern "C"
