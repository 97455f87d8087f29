// The small FRI layers of a proof in one launch.
//
// Up to 1024 values a layer's fold, leaf hashes, tree and transcript step fit
// one workgroup, and launching kernels per layer costs more than the layer
// itself. fri_tail_kernel runs every remaining line fold of a proof in one
// workgroup of 1024 threads: fold (cuda/fri.cu:10-39) into the next layer,
// hash it as an SoA4 tree (level-store nodes written as by the stage/tail
// kernels; one leaf per thread; levels wider than the block's 256 quads with
// one compression per thread, the rest quad-cooperatively across the whole
// block), reduce to the root, mix the root into the transcript and draw the
// next challenge, which stays in LDS for the next layer (and is also written
// to device memory). A 2048-value layer is left to a stage launch and a tree
// tail: its 4095 compressions in one workgroup are bound by one CU's issue
// rate (about 34 us, against about 20 us for a stage over 16 CUs and a tail).
// blockIdx.y selects the proof of a batch (buffers bs words apart).
#pragma once
#include "merkle_engine.hpp"

namespace vx {

constexpr uint32_t FRI_TAIL_MAX_LAYERS = 12;
constexpr uint32_t FRI_TAIL_MAX_ITEMS = 1024;  // values of the largest layer handled here
constexpr uint32_t FRI_TAIL_THREADS = 1024;
static_assert(FRI_TAIL_MAX_ITEMS <= FRI_TAIL_THREADS, "one leaf per thread");

struct FriTailLayer {
    const uint32_t* in[4];   // previous layer, 2n values per column
    uint32_t* out[4];        // this layer, n values per column
    const uint32_t* tw;      // n fold twiddles (shared by a batch)
    uint32_t* root_dst;      // 8 words
    uint32_t* alpha_dst;     // next layer's challenge (null after the last layer)
    LevelStore st;
    uint32_t n;              // values of this layer: a power of two in [2, FRI_TAIL_MAX_ITEMS]
};
struct FriTailArgs {
    FriTailLayer L[FRI_TAIL_MAX_LAYERS];
    uint32_t n_layers;
    const uint32_t* alpha0;  // challenge of the first layer
    DevChannel* channel;
};

static __global__ __launch_bounds__(FRI_TAIL_THREADS) void fri_tail_kernel(FriTailArgs args, uint64_t bs) {
    constexpr uint32_t QUADS = FRI_TAIL_THREADS / 4;
    __shared__ uint4 bufa[2 * FRI_TAIL_MAX_ITEMS];  // node-major: node j in bufa[2j], bufa[2j + 1]
    __shared__ uint4 bufb[FRI_TAIL_MAX_ITEMS];      // half of that
    __shared__ uint32_t msg[16];
    __shared__ uint32_t al[4];
    const uint32_t tid = threadIdx.x;
    const uint64_t boff = (uint64_t)blockIdx.y * bs;
    DevChannel* ch = (DevChannel*)((uint32_t*)args.channel + boff);
    if (tid < 4) al[tid] = args.alpha0[boff + tid];
    __syncthreads();
    for (uint32_t li = 0; li < args.n_layers; li++) {
        const FriTailLayer& L = args.L[li];
        const LevelStore st = L.st.shifted(boff);
        const QM31 alpha = {{al[0], al[1], al[2], al[3]}};
        const uint32_t* i0 = L.in[0] + boff;
        const uint32_t* i1 = L.in[1] + boff;
        const uint32_t* i2 = L.in[2] + boff;
        const uint32_t* i3 = L.in[3] + boff;
        uint32_t* cur = (uint32_t*)bufa;
        uint32_t* nxt = (uint32_t*)bufb;
        // one leaf per thread, each a fold of two values; their parents are the
        // first of the parallel levels below (a thread hashing two leaves and
        // their parent in turn would leave most lanes idle)
        const uint32_t k = tid;
        if (k < L.n) {
            const uint32_t x = 2 * k;
            const uint2 a = *(const uint2*)(i0 + x), b = *(const uint2*)(i1 + x);
            const uint2 c = *(const uint2*)(i2 + x), d = *(const uint2*)(i3 + x);
            const QM31 f = fold_pair({{a.x, b.x, c.x, d.x}}, {{a.y, b.y, c.y, d.y}}, alpha, L.tw[k]);
            L.out[0][boff + k] = f.v[0];
            L.out[1][boff + k] = f.v[1];
            L.out[2][boff + k] = f.v[2];
            L.out[3][boff + k] = f.v[3];
            uint32_t h[8];
            b2s_leaf4(f.v[0], f.v[1], f.v[2], f.v[3], h);
            bufa[2 * k] = make_uint4(h[0], h[1], h[2], h[3]);
            bufa[2 * k + 1] = make_uint4(h[4], h[5], h[6], h[7]);
        }
        uint32_t m = L.n, lvl = 0;
        __syncthreads();
        // wide levels (more compressions than quads): one per thread, all lanes busy
        while (m > 2 * QUADS) {
            const uint32_t hm = m / 2;
            for (uint32_t j = tid; j < hm; j += FRI_TAIL_THREADS) {
                const uint4* c4 = (const uint4*)cur + 4 * j;
                const uint4 p0 = c4[0], p1 = c4[1], p2 = c4[2], p3 = c4[3];
                const uint32_t lh[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
                const uint32_t rh[8] = {p2.x, p2.y, p2.z, p2.w, p3.x, p3.y, p3.z, p3.w};
                uint32_t h[8];
                b2s_node(lh, rh, h);
                if (st.want(lvl + 1)) store8(st.at(lvl + 1, j), h);
                uint4* n4 = (uint4*)nxt + 2 * j;
                n4[0] = make_uint4(h[0], h[1], h[2], h[3]);
                n4[1] = make_uint4(h[4], h[5], h[6], h[7]);
            }
            __syncthreads();
            uint32_t* t = cur;
            cur = nxt;
            nxt = t;
            m = hm;
            lvl++;
        }
        const uint32_t* root = block_quad_reduce<FRI_TAIL_THREADS>(cur, nxt, m, lvl, 0, st);
        if (tid < 64) {  // one wave: root and transcript
            if (tid < 8) L.root_dst[boff + tid] = root[tid];
            uint32_t* adst = L.alpha_dst ? L.alpha_dst + boff : nullptr;
            quad_transcript(ch, root, msg, false, adst);
            if (adst && tid < 4) al[tid] = adst[tid];  // written by this wave
        }
        __syncthreads();
    }
}

// Launches fri_tail_kernel over layers [j0, J) (in chunks of FRI_TAIL_MAX_LAYERS);
// layer(j) fills the descriptor of layer j, alpha(j) is the device address of
// its challenge (alpha(j0) already drawn).
template <class LayerFn, class AlphaFn>
void launch_fri_tail(uint32_t j0, uint32_t J, LayerFn&& layer, AlphaFn&& alpha, DevChannel* ch, uint32_t batch,
                     uint64_t bs, hipStream_t s) {
    for (uint32_t j = j0; j < J; j += FRI_TAIL_MAX_LAYERS) {
        FriTailArgs a = {};
        a.n_layers = J - j < FRI_TAIL_MAX_LAYERS ? J - j : FRI_TAIL_MAX_LAYERS;
        for (uint32_t i = 0; i < a.n_layers; i++) {
            a.L[i] = layer(j + i);
            if (a.L[i].n < 2 || a.L[i].n > FRI_TAIL_MAX_ITEMS || (a.L[i].n & (a.L[i].n - 1))) {
                set_error(-2, "fri_tail: layer size out of range");
                return;
            }
        }
        a.alpha0 = alpha(j);
        a.channel = ch;
        fri_tail_kernel<<<dim3(1, batch), FRI_TAIL_THREADS, 0, s>>>(a, bs);
        check_launch("fri_tail_kernel");
    }
}

}  // namespace vx
