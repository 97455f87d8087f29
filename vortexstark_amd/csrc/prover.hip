// Fused, device-resident Fibonacci prover (Part 3 of the C ABI).
//
// Same protocol and bytes as the reference's prove() + serialize_proof
// (src/prover.rs:1721-1978, src/bin/stark_cli.rs:907-975), organised for one
// MI355X instead of around host round trips:
//
//   INTT (trace domain) -> NTT (eval domain, zero pad folded into the first
//   pass) -> trace tree -> quotient tree (quotient recomputed on the fly) ->
//   circle fold fused with the first FRI tree -> line fold fused with each
//   next FRI tree -> queries -> decommitment
//
// Every Merkle root is mixed into the Fiat-Shamir transcript by the kernel
// that computes it, and that kernel also draws the next challenge into device
// memory, so the whole proof is one stream of launches with no host wait.
// The KRAK v1 byte image is assembled in device memory: a template with the
// static words (magic, counts, path depths) is written once when the prover is
// created and every proof overwrites all the data words.
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "commit.hpp"
#include "fri_tail.hpp"
#include "merkle_engine.hpp"

namespace vx {

// ---------------------------------------------------------------- trace generation
// air.rs:85-131: (F(k), F(k+1)) from (a, b) by powers of [[1,1],[1,0]].
__device__ void fib_at(uint32_t a, uint32_t b, uint64_t k, uint32_t& fk, uint32_t& fk1) {
    uint32_t m00 = 1, m01 = 0, m10 = 0, m11 = 1;
    uint32_t b00 = 1, b01 = 1, b10 = 1, b11 = 0;
    while (k) {
        if (k & 1) {
            uint32_t t00 = madd(mmul(m00, b00), mmul(m01, b10)), t01 = madd(mmul(m00, b01), mmul(m01, b11));
            uint32_t t10 = madd(mmul(m10, b00), mmul(m11, b10)), t11 = madd(mmul(m10, b01), mmul(m11, b11));
            m00 = t00;
            m01 = t01;
            m10 = t10;
            m11 = t11;
        }
        uint32_t t00 = madd(mmul(b00, b00), mmul(b01, b10)), t01 = madd(mmul(b00, b01), mmul(b01, b11));
        uint32_t t10 = madd(mmul(b10, b00), mmul(b11, b10)), t11 = madd(mmul(b10, b01), mmul(b11, b11));
        b00 = t00;
        b01 = t01;
        b10 = t10;
        b11 = t11;
        k >>= 1;
    }
    fk = madd(mmul(m10, b), mmul(m11, a));
    fk1 = madd(mmul(m00, b), mmul(m01, a));
}
__global__ void fib_trace_kernel(uint32_t a, uint32_t b, uint64_t n, uint32_t chunk, uint32_t* out) {
    uint64_t start = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * chunk;
    if (start >= n) return;
    uint32_t f0, f1;
    fib_at(a, b, start, f0, f1);
    uint64_t end = start + chunk < n ? start + chunk : n;
    for (uint64_t i = start; i < end; i++) {
        out[i] = f0;
        uint32_t f2 = madd(f0, f1);
        f0 = f1;
        f1 = f2;
    }
}

// ---------------------------------------------------------------- queries
// Last layer into the proof, mix_felts(last layer), then n_q draw_number(E)
// (prover.rs:1922-1927). After mix_felts the draws only differ in n_draws, so
// every query is squeezed by its own lane. One block per proof of a batch
// (buffers bs words apart); ab holds (a, b) per proof.
constexpr uint32_t MAX_LAST_WORDS = 4096;
__global__ __launch_bounds__(256) void queries_kernel(DevChannel* ch0, const uint32_t* l0, const uint32_t* l1,
                                                     const uint32_t* l2, const uint32_t* l3, uint32_t n_last,
                                                     uint32_t* proof0, uint64_t last_off, uint64_t q_off,
                                                     uint32_t n_q, uint64_t E, uint32_t* queries0,
                                                     const uint32_t* ab, uint64_t bs) {
    __shared__ uint32_t words[MAX_LAST_WORDS];
    __shared__ uint32_t st[8];
    __shared__ uint32_t msg[16];
    const uint32_t tid = threadIdx.x;
    const uint64_t off = (uint64_t)blockIdx.x * bs;
    DevChannel* ch = (DevChannel*)((uint32_t*)ch0 + off);
    uint32_t* proof = proof0 + off;
    uint32_t* queries = queries0 + off;
    for (uint32_t i = tid; i < n_last; i += blockDim.x) {
        uint32_t v[4] = {l0[off + i], l1[off + i], l2[off + i], l3[off + i]};
        for (int c = 0; c < 4; c++) {
            words[4 * i + c] = v[c];
            proof[last_off + 4 * i + c] = v[c];
        }
    }
    __syncthreads();
    if (tid < 64) {
        // mix_felts (channel.rs:48-61: Blake2s over state || felts in 64-byte
        // blocks) by wave 0 quad-cooperatively: every quad hashes the same
        // blocks, quad 0's chaining value is kept
        const uint32_t l = tid & 3;
        const uint32_t total_words = 8 + 4 * n_last, total_bytes = 4 * total_words;
        const uint32_t nb = (total_bytes + 63) / 64;
        uint32_t hl, hh;
        b2s_quad_init(hl, hh);
        for (uint32_t b = 0; b < nb; b++) {
            if (tid < 16) {
                const uint32_t gi = 16 * b + tid;
                msg[tid] = gi < 8 ? ch->state[gi] : (gi < total_words ? words[gi - 8] : 0u);
            }
            VX_WAVE_SYNC();
            const bool last = b + 1 == nb;
            b2s_compress_quad(msg, last ? total_bytes : 64u * (b + 1), last ? 0xFFFFFFFFu : 0u, hl, hh);
            VX_WAVE_SYNC();  // every lane has read the block before the next is written
        }
        if (tid < 4) {
            st[l] = hl;
            st[4 + l] = hh;
        }
        VX_WAVE_SYNC();
        if (tid < 8) ch->state[tid] = st[tid];
        if (tid == 0) {
            ch->n_draws = n_q;
            proof[3] = ab[2 * blockIdx.x];
            proof[4] = ab[2 * blockIdx.x + 1];
        }
    }
    __syncthreads();
    uint32_t s[8];
    for (int c = 0; c < 8; c++) s[c] = st[c];
    for (uint32_t i = tid; i < n_q; i += blockDim.x) {
        uint32_t w[8];
        ch_squeeze(s, i, w);
        uint64_t raw = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
        uint32_t qi = (uint32_t)(raw % E);
        queries[i] = qi;
        proof[q_off + i] = qi;
    }
}

// ---------------------------------------------------------------- decommitment
// prover.rs:1184-1247 and 2171-2234: value at the query index and at its
// sibling (index ^ 1), with both authentication paths (siblings from the leaf
// level up, root excluded). Grid: (query group, tree, proof); tree
// descriptors address proof 0 and move by blockIdx.z * bs words.
struct TreeDesc {
    uint32_t kind;   // 0: one M31 column, 1: Fibonacci quotient of the LDE, 2: SoA4
    uint32_t depth;  // log2(leaves)
    uint32_t shift;  // leaf index = query >> shift
    uint32_t from;   // level store covers levels >= from
    const uint32_t* c[4];
    uint64_t mask;
    const uint32_t* alpha;
    uint32_t* store;
    uint64_t val_off, sib_off, path_off, sib_path_off;
};

__device__ void tree_values(const TreeDesc& t, uint64_t off, uint64_t i, uint32_t v[4]) {
    if (t.kind == 0) {
        SrcCol1{t.c[0]}.at(off).values(i, v);
    } else if (t.kind == 1) {
        SrcQuotient{t.c[0], t.mask, t.alpha}.at(off).values(i, v);
    } else {
        SrcSoa4{t.c[0], t.c[1], t.c[2], t.c[3]}.at(off).values(i, v);
    }
}
__device__ void tree_leaf(const TreeDesc& t, uint64_t off, uint64_t i, uint32_t h[8]) {
    if (t.kind == 0) {
        SrcCol1{t.c[0]}.at(off).leaf(i, h);
    } else if (t.kind == 1) {
        SrcQuotient{t.c[0], t.mask, t.alpha}.at(off).leaf(i, h);
    } else {
        SrcSoa4{t.c[0], t.c[1], t.c[2], t.c[3]}.at(off).leaf(i, h);
    }
}

constexpr uint32_t DECOMMIT_SUB = 5;                                    // recomputed subtree: up to 32 leaves
constexpr uint32_t DECOMMIT_GROUP = 8;                                  // queries of one tree per workgroup
constexpr uint32_t DECOMMIT_THREADS = DECOMMIT_GROUP << DECOMMIT_SUB;  // one thread per subtree leaf
constexpr uint32_t DECOMMIT_MAX_DEPTH = 32;

// Lowest level a tree of depth d keeps in its level store. Trees of up to 2^16
// leaves keep every level from 1 up (a query's path is read, only its two
// leaves are hashed); deeper trees keep 2^16 nodes and more until the lowest
// stored level reaches DECOMMIT_SUB, where a query rebuilds a 32-leaf subtree.
constexpr uint32_t STORE_FULL_DEPTH = 16;
inline uint32_t store_from(uint32_t d) {
    const uint32_t f = d > STORE_FULL_DEPTH ? d - STORE_FULL_DEPTH + 1 : 1;
    return std::min(std::min(f, DECOMMIT_SUB), d);
}

// Workgroup = DECOMMIT_GROUP queries of one tree of one proof. The subtrees
// below the level store are rebuilt together: every level's nodes of all the
// group's queries are packed into the lowest threads, so the waves that hash
// are full and the others wait at the barrier (a single query per wave would
// leave 31 of 63 compressions on partly idle waves).
__global__ __launch_bounds__(DECOMMIT_THREADS) void decommit_kernel(const TreeDesc* trees, const uint32_t* queries0,
                                                                     uint32_t* proof0, uint32_t n_q, uint64_t bs) {
    constexpr uint32_t S = 1u << DECOMMIT_SUB;
    constexpr uint32_t NW = DECOMMIT_THREADS;
    __shared__ uint32_t lv[8 * NW];  // word w of node k of query g's subtree level at w * NW + g * S + k
    __shared__ uint32_t path[DECOMMIT_GROUP][8 * DECOMMIT_MAX_DEPTH];  // sibling of the ancestor per level
    __shared__ uint32_t leaf_self[DECOMMIT_GROUP][8];                   // leaf hash of the query index itself
    __shared__ uint64_t li_s[DECOMMIT_GROUP];
    const TreeDesc t = trees[blockIdx.y];
    const uint64_t off = (uint64_t)blockIdx.z * bs;
    const uint32_t* queries = queries0 + off;
    uint32_t* proof = proof0 + off;
    const uint32_t tid = threadIdx.x;
    const uint32_t D = t.depth;
    const uint32_t lam = t.from < D ? t.from : D;
    const uint32_t lmask = (1u << lam) - 1;
    const uint32_t width = t.kind == 0 ? 1 : 4;
    const uint32_t q0 = blockIdx.x * DECOMMIT_GROUP;
    const uint32_t nq = n_q - q0 < DECOMMIT_GROUP ? n_q - q0 : DECOMMIT_GROUP;

    if (tid < nq) {
        const uint32_t q = q0 + tid;
        const uint64_t li = (uint64_t)queries[q] >> t.shift;
        li_s[tid] = li;
        uint32_t v[4], sv[4];
        tree_values(t, off, li, v);
        tree_values(t, off, li ^ 1, sv);
        for (uint32_t k = 0; k < width; k++) {
            proof[t.val_off + (uint64_t)q * width + k] = v[k];
            proof[t.sib_off + (uint64_t)q * width + k] = sv[k];
        }
    }
    __syncthreads();
    {
        const uint32_t g = tid >> lam, k = tid & lmask;
        if (g < nq) {
            uint32_t h[8];
            tree_leaf(t, off, (li_s[g] & ~(uint64_t)lmask) + k, h);
#pragma unroll
            for (int w = 0; w < 8; w++) lv[w * NW + g * S + k] = h[w];
        }
    }
    __syncthreads();
    if (tid < 8 * nq) {
        const uint32_t g = tid >> 3, w = tid & 7;
        leaf_self[g][w] = lv[w * NW + g * S + ((uint32_t)li_s[g] & lmask)];
    }
    for (uint32_t L = 0; L < lam; L++) {
        if (tid < 8 * nq) {
            const uint32_t g = tid >> 3, w = tid & 7;
            const uint32_t pos = ((uint32_t)li_s[g] & lmask) >> L;
            path[g][L * 8 + w] = lv[w * NW + g * S + (pos ^ 1)];
        }
        if (L + 1 == lam) break;  // level lam comes from the store
        const uint32_t hl = lam - L - 1;  // log2 of each query's nodes on level L + 1
        const uint32_t g = tid >> hl, k = tid & ((1u << hl) - 1);
        const bool act = g < nq;
        uint32_t l[8], r[8];
        if (act) {
#pragma unroll
            for (int w = 0; w < 8; w++) {
                l[w] = lv[w * NW + g * S + 2 * k];
                r[w] = lv[w * NW + g * S + 2 * k + 1];
            }
        }
        __syncthreads();
        if (act) {
            uint32_t h[8];
            b2s_node(l, r, h);
#pragma unroll
            for (int w = 0; w < 8; w++) lv[w * NW + g * S + k] = h[w];
        }
        __syncthreads();
    }
    const LevelStore st = LevelStore{t.store, D, lam}.shifted(off);
    const uint32_t up = D - lam;
    for (uint32_t idx = tid; idx < nq * up * 8; idx += DECOMMIT_THREADS) {
        const uint32_t w = idx & 7, L = lam + (idx >> 3) % up, g = (idx >> 3) / up;
        path[g][L * 8 + w] = st.at(L, (li_s[g] >> L) ^ 1)[w];
    }
    __syncthreads();
    const uint64_t rec = 1 + 8ull * D;
    for (uint32_t idx = tid; idx < nq * (1 + 8 * D); idx += DECOMMIT_THREADS) {
        const uint32_t g = idx / (1 + 8 * D), k = idx % (1 + 8 * D);
        uint32_t* p0 = proof + t.path_off + (uint64_t)(q0 + g) * rec;
        uint32_t* p1 = proof + t.sib_path_off + (uint64_t)(q0 + g) * rec;
        if (k == 0) {
            p0[0] = D;
            p1[0] = D;
        } else {
            p0[k] = path[g][k - 1];
            p1[k] = k - 1 < 8 ? leaf_self[g][k - 1] : path[g][k - 1];
        }
    }
}

// Fibonacci traces of a batch: proof blockIdx.y gets (a, b) = ab[2y], ab[2y+1]
// and writes its trace bs words past proof 0's.
__global__ void fib_trace_batch_kernel(const uint32_t* ab, uint64_t n, uint32_t chunk, uint32_t* out0, uint64_t bs) {
    const uint32_t y = blockIdx.y;
    uint64_t start = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * chunk;
    if (start >= n) return;
    uint32_t* out = out0 + (uint64_t)y * bs;
    uint32_t f0, f1;
    fib_at(ab[2 * y], ab[2 * y + 1], start, f0, f1);
    uint64_t end = start + chunk < n ? start + chunk : n;
    for (uint64_t i = start; i < end; i++) {
        out[i] = f0;
        uint32_t f2 = madd(f0, f1);
        f0 = f1;
        f1 = f2;
    }
}

// Copies words of src to count - 1 places stride words apart (dst0 = src + stride).
__global__ void replicate_kernel(const uint32_t* src, uint64_t words, uint64_t stride) {
    uint32_t* d = const_cast<uint32_t*>(src) + (uint64_t)(blockIdx.y + 1) * stride;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
        d[i] = src[i];
}

// Proofs 0 .. count-1 (words each, R words apart) and their transcript failure
// flags, written by the device straight into pinned host memory (opt-in, see
// vx_prover_set_copy_kernel): the copy stays on the prover's compute queue,
// with no copy-engine hand-off. The grid is small (PCIe, not the CUs, sets the
// pace), but its waves still hold CU slots for the whole transfer; with many
// batches in flight the copy engines (the default) finish the mix sooner.
constexpr uint32_t P2H_BLOCKS = 64;
__global__ void proofs_to_host_kernel(const uint32_t* src, uint64_t R, uint64_t words, uint32_t count, uint32_t* dst,
                                      const uint32_t* failed_src, uint32_t* failed_dst) {
    const uint64_t nth = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t y = 0; y < count; y++) {
        const uint32_t* s = src + (uint64_t)y * R;
        uint32_t* d = dst + (uint64_t)y * words;
        for (uint64_t i = tid; i < words; i += nth) d[i] = s[i];
    }
    if (failed_dst)
        for (uint64_t y = tid; y < count; y += nth) failed_dst[y] = failed_src[y * R];
    __threadfence_system();
}

__global__ void raise_flag_kernel(uint32_t* flag) { *flag = 1; }

}  // namespace vx

// ---------------------------------------------------------------- prover object
using namespace vx;

// A prover for traces of 2^log_n rows with room for `capacity` proofs per
// launch sequence. Every proof owns one arena of R words (trace slot, LDE, FRI
// layers, level stores, Merkle scratch, transcript, challenges, queries, KRAK
// image); proof p's arena starts p * R words after proof 0's, so a single
// stride moves every per-proof pointer and one launch serves the whole batch.
struct vx_prover {
    uint32_t log_n = 0, log_e = 0, n_q = 0, log_last = 0, J = 0, capacity = 1;
    uint64_t N = 0, E = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    uint64_t alloc_bytes = 0;  // device bytes of allocs
    NttTables trace_tab, eval_tab;
    std::vector<uint32_t*> fri_tw;
    // arena layout (word offsets inside one proof's arena)
    uint64_t R = 0;
    uint64_t o_trace = 0, o_eval = 0, o_scr_a = 0, o_scr_b = 0, o_ch = 0, o_alpha = 0, o_queries = 0, o_proof = 0;
    std::vector<uint64_t> o_layer, o_store;
    std::vector<LevelStore> stores;  // proof 0
    uint32_t* arena = nullptr;
    uint64_t proof_words = 0;
    uint64_t off_fri_roots = 0, off_last = 0, off_queries = 0;
    TreeDesc* d_trees = nullptr;
    uint32_t** d_eval_ptrs = nullptr;   // capacity pointers: LDE of proof p
    uint32_t** d_slot_ptrs = nullptr;   // capacity pointers: trace slot of proof p
    // call parameters: trace pointers (capacity) then (a, b) per proof
    uint64_t* d_call = nullptr;
    uint64_t* h_call = nullptr;  // pinned staging
    uint8_t* h_setup = nullptr;  // pinned staging of the set-up copies (kept until destroy)
    hipEvent_t call_copied = nullptr;  // recorded after the last copy out of h_call
    std::vector<uint64_t> last_call;
    bool traces_aligned = false;  // every trace pointer of the current call is 16-byte aligned
    // graph replay of the launch sequence: one instantiated graph per batch
    // count and trace alignment, at most GRAPH_CACHE of them
    static constexpr size_t GRAPH_CACHE = 32;
    // The sequence is captured as two graphs, its heavy part (LDE and the
    // trees over the LDE and the first FRI layer, whose stage kernels fill
    // the device) and its light part (the FRI line folds, queries and
    // decommitment), with heavy_done recorded between them.
    struct Captured {
        uint32_t count;
        bool aligned;  // the captured launches move trace tiles as uint4
        hipGraph_t graph[2];
        hipGraphExec_t exec[2];
    };
    bool use_graph = false;
    std::vector<Captured> graphs;  // most recently used last
    // Recorded on the stream after the heavy part of every launch sequence.
    hipEvent_t heavy_done = nullptr;
    // vx_prover_enqueue_copy into device-visible host memory uses
    // proofs_to_host_kernel instead of the copy engines.
    bool copy_kernel = false;
    // Host-trace calls (vx_prover_prove_batch_host): two device slots of
    // `capacity` traces each, filled over PCIe on their own stream, so one
    // call's upload overlaps the proving of the call before it; a slot is
    // refilled only after the call that read it has finished.
    hipStream_t up_stream = nullptr;
    uint32_t* up_slot[2] = {nullptr, nullptr};
    hipEvent_t up_done[2] = {nullptr, nullptr}, slot_free[2] = {nullptr, nullptr};
    bool slot_used[2] = {false, false};
    uint32_t up_next = 0;
    Timeline tl;
    void mark(const std::string& label) { tl.mark(label, stream); }

    uint32_t* at(uint64_t o) const { return arena + o; }
    uint32_t* col(uint32_t j, uint32_t c) const { return arena + o_layer[j] + (uint64_t)c * (E >> (j + 1)); }
    uint32_t* alpha_q() const { return arena + o_alpha; }
    uint32_t* alpha_f(uint32_t j) const { return arena + o_alpha + 4 * (1 + j); }
    DevChannel* ch() const { return (DevChannel*)(arena + o_ch); }
    uint32_t* proof(uint32_t p = 0) const { return arena + (uint64_t)p * R + o_proof; }
    const uint32_t* ab_dev() const { return (const uint32_t*)(d_call + capacity); }

    // Device memory is carved from one pool sized by the prover's exact
    // estimate (one hipMalloc instead of dozens); regions are 256-byte aligned.
    uint8_t* pool = nullptr;
    uint64_t pool_size = 0, pool_used = 0;
    bool alloc(void** p, uint64_t bytes) {
        const uint64_t need = (bytes + 255) & ~255ull;
        if (pool && pool_used + need <= pool_size) {
            *p = pool + pool_used;
            pool_used += need;
        } else {
            if (!check(hipMalloc(p, bytes ? bytes : 4), "vx_prover: hipMalloc")) return false;
            allocs.push_back(*p);
        }
        alloc_bytes += bytes;
        return true;
    }
    static void destroy(Captured& c) {
        for (int k = 0; k < 2; k++) {
            if (c.exec[k]) (void)hipGraphExecDestroy(c.exec[k]);
            if (c.graph[k]) (void)hipGraphDestroy(c.graph[k]);
        }
    }
    void drop_graphs() {
        for (Captured& c : graphs) destroy(c);
        graphs.clear();
    }
    ~vx_prover() {
        if (up_stream) (void)hipStreamSynchronize(up_stream);
        if (stream) (void)hipStreamSynchronize(stream);
        for (int k = 0; k < 2; k++) {
            if (up_done[k]) (void)hipEventDestroy(up_done[k]);
            if (slot_free[k]) (void)hipEventDestroy(slot_free[k]);
        }
        if (up_stream) (void)hipStreamDestroy(up_stream);
        drop_graphs();
        for (void* p : allocs) (void)hipFree(p);
        if (call_copied) (void)hipEventDestroy(call_copied);
        if (heavy_done) (void)hipEventDestroy(heavy_done);
        if (h_call) (void)hipHostFree(h_call);
        if (h_setup) (void)hipHostFree(h_setup);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

struct Layout {
    std::vector<uint32_t> words;  // static template
    uint64_t fri_roots, last, queries;
    std::vector<TreeDesc> trees;  // offsets only
};

// KRAK v1 word layout (stark_cli.rs:948-975).
void build_layout(vx_prover* p, Layout& L) {
    std::vector<uint32_t>& w = L.words;
    const uint32_t nq = p->n_q;
    w = {0x4B52414Bu, 1u, p->log_n, 0, 0};
    w.resize(5 + 16, 0);  // trace root, quotient root
    w.push_back(p->J);
    L.fri_roots = w.size();
    w.resize(w.size() + 8ull * p->J, 0);
    w.push_back(1u << p->log_last);
    L.last = w.size();
    w.resize(w.size() + (4ull << p->log_last), 0);
    w.push_back(nq);
    L.queries = w.size();
    w.resize(w.size() + nq, 0);
    auto decom = [&](uint32_t width, uint32_t depth, TreeDesc& t) {
        w.push_back(nq);
        t.val_off = w.size();
        w.resize(w.size() + (uint64_t)nq * width, 0);
        t.sib_off = w.size();
        w.resize(w.size() + (uint64_t)nq * width, 0);
        t.path_off = w.size();
        for (uint32_t q = 0; q < nq; q++) {
            w.push_back(depth);
            w.resize(w.size() + 8ull * depth, 0);
        }
        t.sib_path_off = w.size();
        for (uint32_t q = 0; q < nq; q++) {
            w.push_back(depth);
            w.resize(w.size() + 8ull * depth, 0);
        }
    };
    L.trees.assign(2 + p->J, TreeDesc{});
    decom(1, p->log_e, L.trees[0]);
    decom(4, p->log_e, L.trees[1]);
    w.push_back(p->J);
    for (uint32_t j = 0; j < p->J; j++) decom(4, p->log_e - 1 - j, L.trees[2 + j]);
}

// Words of the KRAK v1 image build_layout makes, counted without building it.
uint64_t layout_words(const vx_prover* p) {
    const uint64_t nq = p->n_q;
    auto decom = [&](uint64_t width, uint64_t depth) { return 1 + 2 * nq * width + 2 * nq * (1 + 8 * depth); };
    uint64_t w = 5 + 16 + 1 + 8ull * p->J + 1 + (4ull << p->log_last) + 1 + nq;
    w += decom(1, p->log_e) + decom(4, p->log_e) + 1;
    for (uint32_t j = 0; j < p->J; j++) w += decom(4, p->log_e - 1 - j);
    return w;
}

CPt to_cpt(HostPt h) { return {h.x, h.y}; }

// Word offsets of one arena (each region 256-byte aligned).
void plan_arena(vx_prover* p) {
    uint64_t o = 0;
    auto take = [&](uint64_t words) {
        uint64_t at = o;
        o += (words + 63) & ~63ull;
        return at;
    };
    p->o_trace = take(p->N);
    p->o_eval = take(p->E);
    p->o_layer.assign(p->J, 0);
    for (uint32_t j = 0; j < p->J; j++) p->o_layer[j] = take(4 * (p->E >> (j + 1)));
    p->o_store.clear();
    p->stores.clear();
    auto depth_of = [&](uint32_t t) { return t < 2 ? p->log_e : p->log_e - 1 - (t - 2); };
    for (uint32_t t = 0; t < 2 + p->J; t++) {
        const uint32_t d = depth_of(t);
        p->o_store.push_back(take(level_store_words(d, store_from(d))));
    }
    const uint64_t mx = std::max(max_stage_nodes(p->E), max_line_quad_nodes(p->E));
    p->o_scr_a = take(8 * (mx ? mx : 1));
    p->o_scr_b = take(8 * (mx ? mx : 1));
    p->o_ch = take(sizeof(DevChannel) / 4);
    p->o_alpha = take(4ull * (1 + p->J));
    p->o_queries = take(p->n_q);
    p->o_proof = take(p->proof_words);
    p->R = o;
}

// Enqueues the heavy part of the proof sequence for proofs 0 .. count-1 (trace
// pointers and (a, b) already in d_call): the LDE, the trace and quotient
// trees and the circle fold with its tree.
void enqueue_heavy(vx_prover* p, uint32_t count) {
    hipStream_t s = p->stream;
    const uint64_t E = p->E, R = p->R;
    p->mark("start");
    std::string phase;
    auto mk = [&](const char* what) { p->mark(phase + ":" + what); };
    // LDE: interpolate on half_coset(log_n) into the first N words of the LDE
    // buffer, then evaluate on half_coset(log_n + 2) in place (zero pad implicit)
    lde_batch(p->d_eval_ptrs, count, p->trace_tab, p->eval_tab, s, (const uint32_t* const*)p->d_call,
              p->traces_aligned, [&](const char* what) { p->mark(std::string("lde:") + what); });
    Scratch sc = {p->at(p->o_scr_a), p->at(p->o_scr_b)};
    uint32_t* pr = p->proof();
    uint32_t* ev = p->at(p->o_eval);
    phase = "trace_commit";
    commit_tree(SrcCol1{ev}, E, p->stores[0], TailOut{pr + 5, p->ch(), p->alpha_q(), 1}, sc, s, mk, count, R);
    phase = "quotient_commit";
    commit_tree(SrcQuotient{ev, E - 1, p->alpha_q()}, E, p->stores[1], TailOut{pr + 13, p->ch(), p->alpha_f(0), 0},
                sc, s, mk, count, R);
    {
        phase = "fri0_circle_fold_commit";
        SrcCircleFold src = {ev,           E - 1,        p->alpha_q(), p->alpha_f(0), p->fri_tw[0],
                             p->col(0, 0), p->col(0, 1), p->col(0, 2), p->col(0, 3)};
        uint32_t* nxt = p->J > 1 ? p->alpha_f(1) : nullptr;
        commit_tree(src, E / 2, p->stores[2], TailOut{pr + p->off_fri_roots, p->ch(), nxt, 0}, sc, s, mk, count, R);
    }
}

// Enqueues the light part: the FRI line folds with their trees, the queries
// and the decommitment.
void enqueue_light(vx_prover* p, uint32_t count) {
    hipStream_t s = p->stream;
    const uint64_t E = p->E, R = p->R;
    std::string phase;
    auto mk = [&](const char* what) { p->mark(phase + ":" + what); };
    Scratch sc = {p->at(p->o_scr_a), p->at(p->o_scr_b)};
    uint32_t* pr = p->proof();
    // layers of at most FRI_TAIL_MAX_ITEMS values all go to one fri_tail_kernel
    // launch; the larger ones take a stage launch even where the tree would fit
    // one tail block (direct_max), so that their leaves spread over many CUs
    uint32_t js = 1;
    while (js < p->J && (E >> (js + 1)) > FRI_TAIL_MAX_ITEMS) js++;
    for (uint32_t j = 1; j < js; j++) {
        phase = "fri" + std::to_string(j) + "_line_fold_commit";
        SrcLineFold src = {p->col(j - 1, 0), p->col(j - 1, 1), p->col(j - 1, 2), p->col(j - 1, 3), p->alpha_f(j),
                           p->fri_tw[j],     p->col(j, 0),     p->col(j, 1),     p->col(j, 2),     p->col(j, 3)};
        uint32_t* nxt = j + 1 < p->J ? p->alpha_f(j + 1) : nullptr;
        const TailOut to{pr + p->off_fri_roots + 8ull * j, p->ch(), nxt, 0};
        if (count == 1 && (E >> (j + 1)) <= LINE_QUAD_MAX)
            commit_line_fold_quad(src, E >> (j + 1), p->stores[2 + j], to, sc, s, mk, count, R);
        else
            commit_tree(src, E >> (j + 1), p->stores[2 + j], to, sc, s, mk, count, R, FRI_TAIL_MAX_ITEMS);
    }
    if (js < p->J) {
        auto layer = [&](uint32_t j) {
            FriTailLayer L = {};
            for (int c = 0; c < 4; c++) {
                L.in[c] = p->col(j - 1, c);
                L.out[c] = p->col(j, c);
            }
            L.tw = p->fri_tw[j];
            L.root_dst = pr + p->off_fri_roots + 8ull * j;
            L.alpha_dst = j + 1 < p->J ? p->alpha_f(j + 1) : nullptr;
            L.st = p->stores[2 + j];
            L.n = (uint32_t)(E >> (j + 1));
            return L;
        };
        launch_fri_tail(js, p->J, layer, [&](uint32_t j) { return (const uint32_t*)p->alpha_f(j); }, p->ch(), count,
                        R, s);
        p->mark("fri_small_layers");
    }
    const uint32_t jl = p->J - 1;
    queries_kernel<<<count, 256, 0, s>>>(p->ch(), p->col(jl, 0), p->col(jl, 1), p->col(jl, 2), p->col(jl, 3),
                                         1u << p->log_last, pr, p->off_last, p->off_queries, p->n_q, E,
                                         p->at(p->o_queries), p->ab_dev(), R);
    check_launch("queries_kernel");
    p->mark("queries");
    decommit_kernel<<<dim3((p->n_q + DECOMMIT_GROUP - 1) / DECOMMIT_GROUP, 2 + p->J, count), DECOMMIT_THREADS, 0, s>>>(
        p->d_trees, p->at(p->o_queries), pr, p->n_q, R);
    check_launch("decommit_kernel");
    p->mark("decommit");
}

// Uploads trace pointers and (a, b) of a call; waits for the stream first when
// the pinned staging buffer must change (an earlier copy may still read it).
bool upload_call(vx_prover* p, const uint32_t* const* traces, const uint32_t* ab, uint32_t count) {
    std::vector<uint64_t> want(2ull * p->capacity, 0);
    p->traces_aligned = true;
    for (uint32_t i = 0; i < count; i++) {
        want[i] = (uint64_t)(uintptr_t)traces[i];
        if (want[i] & 15) p->traces_aligned = false;
        ((uint32_t*)(want.data() + p->capacity))[2 * i] = ab[2 * i] % P;
        ((uint32_t*)(want.data() + p->capacity))[2 * i + 1] = ab[2 * i + 1] % P;
    }
    if (want == p->last_call) return true;
    // the previous copy out of the pinned staging must have been read first
    if (p->call_copied && !check(hipEventSynchronize(p->call_copied), "vx_prover: upload")) return false;
    if (!p->call_copied && !check(hipEventCreateWithFlags(&p->call_copied, hipEventDisableTiming), "hipEventCreate"))
        return false;
    std::memcpy(p->h_call, want.data(), 8 * want.size());
    if (!check(hipMemcpyAsync(p->d_call, p->h_call, 8 * want.size(), hipMemcpyHostToDevice, p->stream), "H2D") ||
        !check(hipEventRecord(p->call_copied, p->stream), "hipEventRecord"))
        return false;
    p->last_call = want;
    return true;
}

int32_t run_batch(vx_prover* p, uint32_t count) {
    auto& G = p->graphs;
    auto it = std::find_if(G.begin(), G.end(), [&](const vx_prover::Captured& c) {
        return c.count == count && c.aligned == p->traces_aligned;
    });
    // Graphs are kept until the prover is destroyed (destroying one waits for
    // the whole device); once GRAPH_CACHE counts are cached, other counts are
    // launched directly.
    if (!p->heavy_done && !check(hipEventCreateWithFlags(&p->heavy_done, hipEventDisableTiming), "hipEventCreate"))
        return -1;
    if (!p->use_graph || (it == G.end() && G.size() == vx_prover::GRAPH_CACHE)) {
        p->tl.n = 0;
        enqueue_heavy(p, count);
        if (!check(hipEventRecord(p->heavy_done, p->stream), "hipEventRecord")) return -1;
        enqueue_light(p, count);
        return peek_error() == 0 ? 0 : -1;
    }
    if (it != G.end()) {
        std::rotate(it, it + 1, G.end());  // most recently used last
    } else {
        const bool timing = p->tl.on;
        p->tl.on = false;  // no event records inside a captured graph
        vx_prover::Captured c = {count, p->traces_aligned, {nullptr, nullptr}, {nullptr, nullptr}};
        bool ok = true;
        for (int k = 0; ok && k < 2; k++) {
            ok = check(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            if (!ok) break;
            if (k == 0)
                enqueue_heavy(p, count);
            else
                enqueue_light(p, count);
            ok = check(hipStreamEndCapture(p->stream, &c.graph[k]), "hipStreamEndCapture") && !peek_error() &&
                 check(hipGraphInstantiate(&c.exec[k], c.graph[k], nullptr, nullptr, 0), "hipGraphInstantiate");
        }
        p->tl.on = timing;
        if (!ok) {
            vx_prover::destroy(c);
            return -1;
        }
        G.push_back(c);
    }
    const vx_prover::Captured& c = G.back();
    return check(hipGraphLaunch(c.exec[0], p->stream), "hipGraphLaunch") &&
                   check(hipEventRecord(p->heavy_done, p->stream), "hipEventRecord") &&
                   check(hipGraphLaunch(c.exec[1], p->stream), "hipGraphLaunch")
               ? 0
               : -1;
}

// Applies the reference defaults (80 queries, last layer 2^3) for zero
// arguments and checks the parameters a prover supports.
bool prover_params(uint32_t log_n, uint32_t& n_queries, uint32_t& log_last_layer, uint32_t capacity) {
    if (n_queries == 0) n_queries = 80;           // prover.rs:90
    if (log_last_layer == 0) log_last_layer = 3;  // fri.rs:265
    // log_n + 2 <= 31: the largest evaluation domain is the whole circle group
    // (circle.rs:102 asserts log_size <= 31)
    return !(log_n < 2 || log_n + 2 > 31 || log_last_layer > log_n + 1 || log_last_layer > 10 || n_queries > 4096 ||
             capacity < 1 || capacity > 65535);
}

void set_sizes(vx_prover* p, uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer, uint32_t capacity) {
    p->log_n = log_n;
    p->log_e = log_n + 2;  // BLOWUP_BITS = 2 (prover.rs:86)
    p->n_q = n_queries;
    p->log_last = log_last_layer;
    p->capacity = capacity;
    p->N = 1ull << log_n;
    p->E = 1ull << p->log_e;
    p->J = p->log_e - log_last_layer;
}

// Device bytes vx_prover_create_batch allocates besides the arenas: NTT
// twiddles of the trace and evaluation domains, fold twiddles and their
// scratch, tree descriptors and per-proof call slots.
uint64_t shared_bytes(const vx_prover* p) {
    const uint64_t N = p->N, E = p->E;
    uint64_t b = 2 * 4 * (N - 1) + 2 * 4 * (N / 2) + 4 * (E - 1) + 4 * (E / 2) + 4 * (E / 2);
    for (uint32_t j = 0; j < p->J; j++) b += 4 * ((1ull << (p->log_e - j)) / 2);
    return b + sizeof(TreeDesc) * (2 + p->J) + (8 + 8 + 16) * (uint64_t)p->capacity;
}

// Device address of host memory the device may write directly (allocated by
// hipHostMalloc), or null for any other memory.
void* device_view_of_host(void* p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (a.type != hipMemoryTypeHost || !a.devicePointer || !a.hostPointer) return nullptr;
    return (char*)a.devicePointer + ((char*)p - (char*)a.hostPointer);
}

}  // namespace

namespace vx {
// Used by the proof server to report one proof of a batch failed through the
// same flag a transcript failure raises (its fault-injection hook).
bool prover_flag_failed(vx_prover* p, uint32_t index) {
    if (!p || index >= p->capacity) return false;
    uint32_t* flag = (uint32_t*)p->ch() + (uint64_t)index * p->R + offsetof(DevChannel, failed) / 4;
    raise_flag_kernel<<<1, 1, 0, p->stream>>>(flag);
    return check_launch("raise_flag_kernel");
}
// The event recorded on the prover's stream after the heavy part of its latest
// launch sequence (the LDE and the trees of the trace, the quotient and the
// circle fold); null before the first.
hipEvent_t prover_heavy_done(vx_prover* p) { return p ? p->heavy_done : nullptr; }
}  // namespace vx

extern "C" {

vx_prover* vx_prover_create_batch(uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer, uint32_t capacity) {
    if (!prover_params(log_n, n_queries, log_last_layer, capacity)) {
        set_error(-2, "vx_prover_create: unsupported parameters");
        return nullptr;
    }
    vx_prover* p = new vx_prover();
    set_sizes(p, log_n, n_queries, log_last_layer, capacity);
    const uint64_t N = p->N, E = p->E;
    p->pool_size = vx_prover_estimate_bytes(log_n, n_queries, log_last_layer, capacity) + 256ull * (64 + p->J);
    if (!check(hipMalloc((void**)&p->pool, p->pool_size), "vx_prover: hipMalloc")) {
        delete p;
        return nullptr;
    }
    p->allocs.push_back(p->pool);
    bool ok = check(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreate");
    hipStream_t s = p->stream;

    // twiddle tables (ProverCache, prover.rs:159-176), shared by the batch
    uint32_t *t_line, *t_circle, *t_iline, *t_icircle, *e_line, *e_circle;
    ok = ok && p->alloc((void**)&t_line, 4 * (N - 1)) && p->alloc((void**)&t_circle, 4 * (N / 2)) &&
         p->alloc((void**)&t_iline, 4 * (N - 1)) && p->alloc((void**)&t_icircle, 4 * (N / 2)) &&
         p->alloc((void**)&e_line, 4 * (E - 1)) && p->alloc((void**)&e_circle, 4 * (E / 2));
    if (!ok) {
        delete p;
        return nullptr;
    }
    // evaluation domain half_coset(log_e), or the whole group subgroup(31) when
    // log_e = 31, where no half coset exists (prover.rs:626-630)
    HostPt ti, ts, ei, es;
    host_coset(1, log_n, ti, ts);
    host_coset(p->log_e <= 30 ? 1 : 0, p->log_e, ei, es);
    build_ntt_tables(to_cpt(ti), to_cpt(ts), log_n, t_line, t_circle, t_iline, t_icircle, p->trace_tab.offs, s);
    build_ntt_tables(to_cpt(ei), to_cpt(es), p->log_e, e_line, e_circle, nullptr, nullptr, p->eval_tab.offs, s);
    p->trace_tab.line = t_iline;
    p->trace_tab.circle = t_icircle;
    p->trace_tab.log_n = log_n;
    p->eval_tab.line = e_line;
    p->eval_tab.circle = e_circle;
    p->eval_tab.log_n = p->log_e;

    // FRI fold twiddles (fri.rs:143-160): [0] circle on half_coset(log_e), y
    // (subgroup(31) when log_e = 31, prover.rs:754-757: its points (1, 0) and
    // (-1, 0) have y = 0, whose inverse the batch inverse takes as 0);
    // [j] line on half_odds(log_e - j), x.
    uint32_t* tw_scratch;
    ok = p->alloc((void**)&tw_scratch, 4 * (E / 2));
    p->fri_tw.assign(p->J, nullptr);
    for (uint32_t j = 0; ok && j < p->J; j++) {
        uint32_t k = p->log_e - j;
        ok = p->alloc((void**)&p->fri_tw[j], 4 * ((1ull << k) / 2));
        if (!ok) break;
        HostPt di, dstp;
        host_coset(j == 0 ? (k <= 30 ? 1 : 0) : 2, k, di, dstp);
        build_fold_twiddles(to_cpt(di), to_cpt(dstp), k, j == 0, p->fri_tw[j], tw_scratch, s);
    }

    Layout L;
    build_layout(p, L);
    p->proof_words = L.words.size();
    if (p->proof_words != layout_words(p)) {  // the counted and the built layout must agree
        set_error(-1, "vx_prover_create: proof layout size mismatch");
        delete p;
        return nullptr;
    }
    p->off_fri_roots = L.fri_roots;
    p->off_last = L.last;
    p->off_queries = L.queries;
    plan_arena(p);
    ok = ok && p->alloc((void**)&p->arena, 4 * p->R * capacity);
    ok = ok && p->alloc((void**)&p->d_trees, sizeof(TreeDesc) * L.trees.size()) &&
         p->alloc((void**)&p->d_eval_ptrs, 8ull * capacity) && p->alloc((void**)&p->d_slot_ptrs, 8ull * capacity) &&
         p->alloc((void**)&p->d_call, 16ull * capacity);
    ok = ok && check(hipHostMalloc((void**)&p->h_call, 16ull * capacity), "hipHostMalloc");
    if (ok) {
        for (uint32_t t = 0; t < 2 + p->J; t++) {
            const uint32_t d = t < 2 ? p->log_e : p->log_e - 1 - (t - 2);
            p->stores.push_back(LevelStore{p->at(p->o_store[t]), d, store_from(d)});
        }
        std::vector<TreeDesc>& T = L.trees;
        uint32_t* ev = p->at(p->o_eval);
        T[0].kind = 0;
        T[0].depth = p->log_e;
        T[0].shift = 0;
        T[0].c[0] = ev;
        T[1].kind = 1;
        T[1].depth = p->log_e;
        T[1].shift = 0;
        T[1].c[0] = ev;
        T[1].mask = E - 1;
        T[1].alpha = p->alpha_q();
        for (uint32_t j = 0; j < p->J; j++) {
            TreeDesc& t = T[2 + j];
            t.kind = 2;
            t.depth = p->log_e - 1 - j;
            t.shift = j + 1;
            for (int c = 0; c < 4; c++) t.c[c] = p->col(j, c);
        }
        for (size_t i = 0; i < T.size(); i++) {
            T[i].from = p->stores[i].from;
            T[i].store = p->stores[i].base;
        }
        // Set-up copies go through pinned staging owned by the prover, so they are
        // asynchronous and nothing here waits for the device (a new prover must not
        // stall the caller behind work of other streams); the prover's first
        // launch is ordered after them on its stream.
        const uint64_t tb = sizeof(TreeDesc) * T.size(), wb = 4 * p->proof_words, cb = 8ull * capacity;
        ok = ok && check(hipHostMalloc((void**)&p->h_setup, tb + wb + 2 * cb), "hipHostMalloc");
        if (ok) {
            uint8_t* h = p->h_setup;
            std::memcpy(h, T.data(), tb);
            std::memcpy(h + tb, L.words.data(), wb);
            uint64_t* evp = (uint64_t*)(h + tb + wb);
            uint64_t* slots = evp + capacity;
            for (uint32_t i = 0; i < capacity; i++) {
                evp[i] = (uint64_t)(uintptr_t)(p->arena + (uint64_t)i * p->R + p->o_eval);
                slots[i] = (uint64_t)(uintptr_t)(p->arena + (uint64_t)i * p->R + p->o_trace);
            }
            ok = check(hipMemcpyAsync(p->d_trees, h, tb, hipMemcpyHostToDevice, s), "H2D") &&
                 check(hipMemcpyAsync(p->proof(0), h + tb, wb, hipMemcpyHostToDevice, s), "H2D") &&
                 check(hipMemcpyAsync(p->d_eval_ptrs, evp, cb, hipMemcpyHostToDevice, s), "H2D") &&
                 check(hipMemcpyAsync(p->d_slot_ptrs, slots, cb, hipMemcpyHostToDevice, s), "H2D");
        }
        // KRAK template into every other proof slot
        if (ok && capacity > 1) {
            const uint32_t gx = (uint32_t)std::min<uint64_t>(64, (p->proof_words + 255) / 256);
            replicate_kernel<<<dim3(gx, capacity - 1), 256, 0, s>>>(p->proof(0), p->proof_words, p->R);
            ok = check_launch("replicate_kernel");
        }
    }
    if (!ok) {
        delete p;
        return nullptr;
    }
    return p;
}

vx_prover* vx_prover_create(uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer) {
    return vx_prover_create_batch(log_n, n_queries, log_last_layer, 1);
}

void vx_prover_destroy(vx_prover* p) { delete p; }
void* vx_prover_stream(vx_prover* p) { return p ? (void*)p->stream : nullptr; }
uint32_t vx_prover_capacity(const vx_prover* p) { return p ? p->capacity : 0; }
uint64_t vx_prover_proof_size(const vx_prover* p) { return p ? 4 * p->proof_words : 0; }
uint64_t vx_prover_device_bytes(const vx_prover* p) { return p ? p->alloc_bytes : 0; }
uint64_t vx_prover_estimate_bytes(uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer, uint32_t capacity) {
    if (!prover_params(log_n, n_queries, log_last_layer, capacity)) return 0;
    vx_prover p;
    set_sizes(&p, log_n, n_queries, log_last_layer, capacity);
    p.proof_words = layout_words(&p);
    plan_arena(&p);
    return 4 * p.R * capacity + shared_bytes(&p);
}
const void* vx_prover_proof_device(const vx_prover* p) { return p ? p->proof(0) : nullptr; }
const void* vx_prover_proof_device_at(const vx_prover* p, uint32_t index) {
    return (p && index < p->capacity) ? p->proof(index) : nullptr;
}
const uint32_t* vx_prover_eval_device(const vx_prover* p) { return p ? p->at(p->o_eval) : nullptr; }
const uint32_t* vx_prover_fri_layer_device(const vx_prover* p, uint32_t j, uint32_t column) {
    if (!p || j >= p->J || column >= 4) return nullptr;
    return p->col(j, column);
}

int32_t vx_prover_prove_batch_device(vx_prover* p, const uint32_t* const* h_traces, const uint32_t* h_ab,
                                     uint32_t count) {
    if (!p || !h_traces || !h_ab || count == 0 || count > p->capacity) {
        set_error(-2, "vx_prover_prove_batch_device: bad arguments");
        return -1;
    }
    for (uint32_t i = 0; i < count; i++)
        if (!h_traces[i]) {
            set_error(-2, "vx_prover_prove_batch_device: null trace");
            return -1;
        }
    if (!upload_call(p, h_traces, h_ab, count)) return -1;
    return run_batch(p, count);
}

int32_t vx_prover_prove_batch_host(vx_prover* p, const uint32_t* const* h_traces, const uint32_t* h_ab,
                                   uint32_t count) {
    if (!p || !h_traces || !h_ab || count == 0 || count > p->capacity) {
        set_error(-2, "vx_prover_prove_batch_host: bad arguments");
        return -1;
    }
    for (uint32_t i = 0; i < count; i++)
        if (!h_traces[i]) {
            set_error(-2, "vx_prover_prove_batch_host: null trace");
            return -1;
        }
    const uint64_t tb = 4 * p->N;
    if (!p->up_stream) {  // first host-trace call: two slots, events, then the upload stream
        bool ok = true;
        for (int k = 0; ok && k < 2; k++) {
            if (!p->up_slot[k]) ok = p->alloc((void**)&p->up_slot[k], tb * p->capacity);
            if (ok && !p->up_done[k])
                ok = check(hipEventCreateWithFlags(&p->up_done[k], hipEventDisableTiming), "hipEventCreate");
            if (ok && !p->slot_free[k])
                ok = check(hipEventCreateWithFlags(&p->slot_free[k], hipEventDisableTiming), "hipEventCreate");
        }
        // the stream is made last: its presence means everything above exists (a
        // failed first call leaves what it made for the next call to complete)
        if (!ok || !check(hipStreamCreateWithFlags(&p->up_stream, hipStreamNonBlocking), "hipStreamCreate")) {
            p->up_stream = nullptr;
            return -1;
        }
    }
    const uint32_t k = p->up_next;
    hipStream_t u = p->up_stream;
    if (p->slot_used[k] && !check(hipStreamWaitEvent(u, p->slot_free[k], 0), "hipStreamWaitEvent")) return -1;
    std::vector<const uint32_t*> slots(count);
    bool pageable = false;
    for (uint32_t i = 0; i < count; i++) {
        slots[i] = p->up_slot[k] + (uint64_t)i * p->N;
        pageable = pageable || !device_view_of_host((void*)h_traces[i]);
        if (!check(hipMemcpyAsync((void*)slots[i], h_traces[i], tb, hipMemcpyHostToDevice, u), "H2D trace"))
            return -1;
    }
    if (!check(hipEventRecord(p->up_done[k], u), "hipEventRecord") ||
        !check(hipStreamWaitEvent(p->stream, p->up_done[k], 0), "hipStreamWaitEvent"))
        return -1;
    // memory that is not pinned may be freed or reused as soon as this returns:
    // wait for its copies (the previous call keeps proving meanwhile)
    if (pageable && !check(hipEventSynchronize(p->up_done[k]), "hipEventSynchronize")) return -1;
    if (!upload_call(p, slots.data(), h_ab, count) || run_batch(p, count)) return -1;
    if (!check(hipEventRecord(p->slot_free[k], p->stream), "hipEventRecord")) return -1;
    p->slot_used[k] = true;
    p->up_next ^= 1;
    return 0;
}

int32_t vx_prover_prove_device(vx_prover* p, const uint32_t* d_trace, uint32_t a, uint32_t b) {
    if (!p || !d_trace) {
        set_error(-2, "vx_prover_prove_device: null argument");
        return -1;
    }
    const uint32_t ab[2] = {a, b};
    return vx_prover_prove_batch_device(p, &d_trace, ab, 1);
}

int32_t vx_prove_fibonacci_batch(vx_prover* p, const uint32_t* h_ab, uint32_t count) {
    if (!p || !h_ab || count == 0 || count > p->capacity) {
        set_error(-2, "vx_prove_fibonacci_batch: bad arguments");
        return -1;
    }
    std::vector<const uint32_t*> slots(count);
    for (uint32_t i = 0; i < count; i++) slots[i] = p->arena + (uint64_t)i * p->R + p->o_trace;
    if (!upload_call(p, slots.data(), h_ab, count)) return -1;
    const uint32_t chunk = 256;
    const uint64_t threads = (p->N + chunk - 1) / chunk;
    fib_trace_batch_kernel<<<dim3(ceil_div(threads, 256), count), 256, 0, p->stream>>>(
        p->ab_dev(), p->N, chunk, p->arena + p->o_trace, p->R);
    if (!check_launch("fib_trace_batch_kernel")) return -1;
    return run_batch(p, count);
}

uint64_t vx_prover_proof_copy_at(vx_prover* p, uint32_t index, void* host_out, uint64_t cap) {
    if (!p || index >= p->capacity) return 0;
    uint64_t bytes = 4 * p->proof_words;
    if (!host_out || cap < bytes) {
        set_error(-2, "vx_prover_proof_copy: buffer too small");
        return 0;
    }
    DevChannel ch;
    const DevChannel* dch = (const DevChannel*)((const uint32_t*)p->ch() + (uint64_t)index * p->R);
    if (!check(hipMemcpyAsync(host_out, p->proof(index), bytes, hipMemcpyDeviceToHost, p->stream), "D2H")) return 0;
    if (!check(hipMemcpyAsync(&ch, dch, sizeof(ch), hipMemcpyDeviceToHost, p->stream), "D2H")) return 0;
    if (!check(hipStreamSynchronize(p->stream), "vx_prover_proof_copy")) return 0;
    if (ch.failed) {
        set_error(-3, "transcript: draw_felt exhausted its retry budget");
        return 0;
    }
    return bytes;
}
uint64_t vx_prover_proof_copy(vx_prover* p, void* host_out, uint64_t cap) {
    return vx_prover_proof_copy_at(p, 0, host_out, cap);
}

int32_t vx_prover_enqueue_copy(vx_prover* p, uint32_t count, void* host_proofs, uint32_t* host_failed, void* stream) {
    if (!p || count == 0 || count > p->capacity || !host_proofs) {
        set_error(-2, "vx_prover_enqueue_copy: bad arguments");
        return -1;
    }
    hipStream_t s = stream ? (hipStream_t)stream : p->stream;
    const uint64_t pb = 4 * p->proof_words, pitch = 4 * p->R;
    uint32_t* dev_proofs = (uint32_t*)device_view_of_host(host_proofs);
    uint32_t* dev_failed = host_failed ? (uint32_t*)device_view_of_host(host_failed) : nullptr;
    if (p->copy_kernel && dev_proofs && ((uintptr_t)dev_proofs & 3) == 0 &&
        (!host_failed || (dev_failed && ((uintptr_t)dev_failed & 3) == 0))) {
        proofs_to_host_kernel<<<P2H_BLOCKS, 256, 0, s>>>(p->proof(0), p->R, p->proof_words, count, dev_proofs,
                                                         &p->ch()->failed, dev_failed);
        return check_launch("proofs_to_host_kernel") ? 0 : -1;
    }
    // the default, and any host memory: one strided DMA copy per field (proof i
    // sits i * R words after proof 0) on the copy engines, which leaves the CUs
    // to the kernels of other batches
    if (!check(hipMemcpy2DAsync(host_proofs, pb, p->proof(0), pitch, pb, count, hipMemcpyDeviceToHost, s), "D2H"))
        return -1;
    if (host_failed &&
        !check(hipMemcpy2DAsync(host_failed, 4, &p->ch()->failed, pitch, 4, count, hipMemcpyDeviceToHost, s), "D2H"))
        return -1;
    return 0;
}

int32_t vx_merkle_autotune(uint32_t log_min, uint32_t log_max, float* best_ms) {
    if (log_min < 12 || log_max < log_min || log_max > 27) {
        set_error(-2, "vx_merkle_autotune: sizes must satisfy 12 <= log_min <= log_max <= 27");
        return -1;
    }
    const uint64_t N = 1ull << log_max, mx = max_stage_nodes(N);
    hipStream_t st = nullptr;
    uint32_t *data = nullptr, *sa = nullptr, *sb = nullptr, *root = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate") &&
              check(hipMalloc((void**)&data, 4 * N), "hipMalloc") && check(hipMalloc((void**)&sa, 32 * mx), "hipMalloc") &&
              check(hipMalloc((void**)&sb, 32 * mx), "hipMalloc") && check(hipMalloc((void**)&root, 32), "hipMalloc") &&
              check(hipEventCreate(&e0), "hipEvent") && check(hipEventCreate(&e1), "hipEvent") &&
              check(hipMemsetAsync(data, 0x11, 4 * N, st), "hipMemsetAsync");  // canonical M31 words
    const LevelStore none = {nullptr, 0, 0};
    const TailOut out = {root, nullptr, nullptr, 0};
    auto no_mark = [](const char*) {};
    // smaller sizes first: a tree's node stages use the depths already tuned for their sizes
    for (uint32_t lt = log_min; ok && lt <= log_max; lt++) {
        const uint64_t n = 1ull << lt;
        float best = 1e30f;
        uint32_t bc = 0;
        for (uint32_t c = 0; ok && c <= MERKLE_LOGC_MAX && (n >> (7 + c)) >= 2; c++) {
            stage_plan_table()[lt].store((uint8_t)(c + 1));
            commit_tree(SrcCol1{data}, n, none, out, Scratch{sa, sb}, st, no_mark);  // warm-up
            const int reps = 3;
            ok = check(hipEventRecord(e0, st), "hipEventRecord");
            for (int r = 0; r < reps; r++) commit_tree(SrcCol1{data}, n, none, out, Scratch{sa, sb}, st, no_mark);
            ok = ok && check(hipEventRecord(e1, st), "hipEventRecord") && check(hipEventSynchronize(e1), "sync");
            float ms = 0.f;
            ok = ok && check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime") && peek_error() == 0;
            if (ok && ms / reps < best) {
                best = ms / reps;
                bc = c;
            }
        }
        stage_plan_table()[lt].store(ok ? (uint8_t)(bc + 1) : (uint8_t)0);
        if (best_ms) best_ms[lt - log_min] = best;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamSynchronize(st);
    for (uint32_t* b : {data, sa, sb, root})
        if (b) (void)hipFree(b);
    if (st) (void)hipStreamDestroy(st);
    return ok ? 0 : -1;
}

int32_t vx_merkle_stage_plan(uint32_t log_items) {
    return log_items < STAGE_PLAN_LOGS ? (int32_t)stage_plan_table()[log_items].load() - 1 : -1;
}

int32_t vx_merkle_stage_plan_set(uint32_t log_items, int32_t logc) {
    if (log_items >= STAGE_PLAN_LOGS || logc < -1 || logc > (int32_t)MERKLE_LOGC_MAX) {
        set_error(-2, "vx_merkle_stage_plan_set: log_items < 40 and -1 <= logc <= 5");
        return -1;
    }
    stage_plan_table()[log_items].store((uint8_t)(logc + 1));
    return 0;
}

void vx_merkle_stage_plan_reset(void) {
    for (uint32_t i = 0; i < STAGE_PLAN_LOGS; i++) stage_plan_table()[i].store(0);
}

uint64_t vx_krak_proof_size(uint32_t log_n, uint32_t n_queries, uint32_t log_last_layer) {
    if (!prover_params(log_n, n_queries, log_last_layer, 1)) return 0;
    vx_prover shape;  // sizes only: no device resources
    set_sizes(&shape, log_n, n_queries, log_last_layer, 1);
    return 4ull * layout_words(&shape);
}

int32_t vx_prover_set_graph(vx_prover* p, int32_t enable) {
    if (!p) return -1;
    if (!check(hipStreamSynchronize(p->stream), "vx_prover_set_graph")) return -1;
    p->use_graph = enable != 0;
    if (!p->use_graph) p->drop_graphs();
    return 0;
}

int32_t vx_prover_set_copy_kernel(vx_prover* p, int32_t enable) {
    if (!p) {
        set_error(-2, "vx_prover_set_copy_kernel: null prover");
        return -1;
    }
    p->copy_kernel = enable != 0;
    return 0;
}

int32_t vx_prover_set_timing(vx_prover* p, int32_t enable) {
    if (!p) return -1;
    p->tl.reset(enable != 0);
    return 0;
}
uint32_t vx_prover_timeline_count(const vx_prover* p) { return p ? (uint32_t)p->tl.n : 0; }
const char* vx_prover_timeline_label(const vx_prover* p, uint32_t i) { return p ? p->tl.label(i) : ""; }
float vx_prover_timeline_ms(vx_prover* p, uint32_t i) { return p ? p->tl.ms(i) : 0.f; }

int32_t vx_fibonacci_trace(uint32_t a, uint32_t b, uint32_t log_n, uint32_t* d_out, void* stream) {
    uint64_t n = 1ull << log_n;
    const uint32_t chunk = 256;
    uint64_t threads = (n + chunk - 1) / chunk;
    fib_trace_kernel<<<ceil_div(threads, 256), 256, 0, (hipStream_t)stream>>>(a, b, n, chunk, d_out);
    return check_launch("vx_fibonacci_trace") ? 0 : -1;
}

uint64_t vx_prove_fibonacci(vx_prover* p, uint32_t a, uint32_t b, void* host_out, uint64_t cap) {
    if (!p) return 0;
    const uint32_t ab[2] = {a, b};
    if (vx_prove_fibonacci_batch(p, ab, 1)) return 0;
    return vx_prover_proof_copy(p, host_out, cap);
}

}  // extern "C"
