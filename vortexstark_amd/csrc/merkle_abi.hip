// Reference-ABI Blake2s Merkle entry points (cuda/blake2s.cu:522-632) on top of
// the wave64 engine. Hash values and tree shapes are the reference's; the
// tiled forms return the root of every 1024-leaf tile.
#include "common.hpp"
#include "merkle_engine.hpp"

namespace vx {

// Up to 16 columns behind a device pointer array; leaf uses the first
// min(n_cols, 16) columns and t = 4 * min(n_cols, 16) (cuda/blake2s.cu:25-60).
struct SrcGeneric {
    const uint32_t* const* cols;
    uint32_t n_cols;
    __device__ __forceinline__ SrcGeneric at(uint64_t) const { return *this; }
    __device__ __forceinline__ void leaf(uint64_t i, uint32_t h[8]) const {
        uint32_t eff = n_cols > 16 ? 16 : n_cols;
        uint32_t m[16];
#pragma unroll
        for (int k = 0; k < 16; k++) m[k] = (uint32_t)k < eff ? cols[k][i] : 0u;
        b2s_leaf(m, eff, h);
    }
    __device__ __forceinline__ void pair(uint64_t i, uint32_t h[8]) const {
        uint32_t l[8], r[8];
        leaf(i, l);
        leaf(i + 1, r);
        b2s_node(l, r, h);
    }
};

__global__ void hash_leaves_kernel(SrcGeneric src, uint32_t* out, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t h[8];
    src.leaf(i, h);
    store8(out + 8ull * i, h);
}
__global__ void hash_nodes_kernel(const uint32_t* ch, uint32_t* par, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t h[8];
    SrcNodes{ch}.pair(2ull * i, h);
    store8(par + 8ull * i, h);
}
__global__ void leaves_merge_soa4_kernel(SrcSoa4 src, uint32_t* par, uint32_t n_pairs) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pairs) return;
    uint32_t h[8];
    src.pair(2ull * i, h);
    store8(par + 8ull * i, h);
}
__global__ void copy8_kernel(const uint32_t* s, uint32_t* d) {
    if (threadIdx.x < 8) d[threadIdx.x] = s[threadIdx.x];
}

template <class Src>
static void tail_blocks(Src src, uint32_t n_items, uint32_t blocks, uint32_t* roots, hipStream_t st,
                        const char* what) {
    LevelStore none = {nullptr, 0, 0};
    TailOut out = {roots, nullptr, nullptr, 0};
    merkle_tail_kernel<Src><<<blocks, TAIL_THREADS, 0, st>>>(src, 0, n_items, none, out, 0);
    check_launch(what);
}

}  // namespace vx

using namespace vx;

extern "C" {

void cuda_merkle_hash_leaves(const uint32_t* const* columns, uint32_t* hashes, uint32_t n_cols, uint32_t n_leaves) {
    if (!n_leaves) return;
    hash_leaves_kernel<<<ceil_div(n_leaves, 256), 256, 0, current_stream()>>>(SrcGeneric{columns, n_cols}, hashes,
                                                                              n_leaves);
    check_launch("cuda_merkle_hash_leaves");
}

void cuda_merkle_hash_nodes(const uint32_t* children, uint32_t* parents, uint32_t n_parents) {
    if (!n_parents) return;
    hash_nodes_kernel<<<ceil_div(n_parents, 256), 256, 0, current_stream()>>>(children, parents, n_parents);
    check_launch("cuda_merkle_hash_nodes");
}

void cuda_merkle_reduce_to_root(const uint32_t* nodes, uint32_t* root_out, uint32_t n_nodes) {
    if (n_nodes == 0) return;
    if (n_nodes == 1) {
        copy8_kernel<<<1, 64, 0, current_stream()>>>(nodes, root_out);
        check_launch("cuda_merkle_reduce_to_root");
        return;
    }
    if (!is_pow2(n_nodes) || n_nodes > MERKLE_TAIL_MAX) {
        set_error(-2, "cuda_merkle_reduce_to_root: n_nodes must be a power of two <= 4096");
        return;
    }
    tail_blocks(SrcNodes{nodes}, n_nodes, 1, root_out, current_stream(), "cuda_merkle_reduce_to_root");
}

void cuda_merkle_commit_small_soa4(const uint32_t* c0, const uint32_t* c1, const uint32_t* c2, const uint32_t* c3,
                                   uint32_t* root_out, uint32_t n_leaves) {
    if (n_leaves < 2 || !is_pow2(n_leaves) || n_leaves > MERKLE_TAIL_MAX) {
        set_error(-2, "cuda_merkle_commit_small_soa4: n_leaves must be a power of two in [2, 4096]");
        return;
    }
    tail_blocks(SrcSoa4{c0, c1, c2, c3}, n_leaves, 1, root_out, current_stream(), "cuda_merkle_commit_small_soa4");
}

void cuda_merkle_hash_leaves_merge_soa4(const uint32_t* c0, const uint32_t* c1, const uint32_t* c2,
                                        const uint32_t* c3, uint32_t* parents, uint32_t n_leaves) {
    uint32_t n_pairs = n_leaves / 2;
    if (!n_pairs) return;
    leaves_merge_soa4_kernel<<<ceil_div(n_pairs, 256), 256, 0, current_stream()>>>(SrcSoa4{c0, c1, c2, c3}, parents,
                                                                                  n_pairs);
    check_launch("cuda_merkle_hash_leaves_merge_soa4");
}

void cuda_merkle_tiled_soa4(const uint32_t* c0, const uint32_t* c1, const uint32_t* c2, const uint32_t* c3,
                            uint32_t* subtree_roots, uint32_t n_leaves) {
    cuda_merkle_tiled_soa4_stream(c0, c1, c2, c3, subtree_roots, n_leaves, current_stream());
}

void cuda_merkle_tiled_soa4_stream(const uint32_t* c0, const uint32_t* c1, const uint32_t* c2, const uint32_t* c3,
                                   uint32_t* subtree_roots, uint32_t n_leaves, void* stream) {
    uint32_t tiles = n_leaves / 1024;
    if (!tiles) return;
    tail_blocks(SrcSoa4{c0, c1, c2, c3}, 1024, tiles, subtree_roots, (hipStream_t)stream, "cuda_merkle_tiled_soa4");
}

void cuda_merkle_tiled_generic(const uint32_t* const* columns, uint32_t* subtree_roots, uint32_t n_cols,
                               uint32_t n_leaves) {
    uint32_t tiles = n_leaves / 1024;
    if (!tiles) return;
    tail_blocks(SrcGeneric{columns, n_cols}, 1024, tiles, subtree_roots, current_stream(),
                "cuda_merkle_tiled_generic");
}

}  // extern "C"
