// Device calibration: the int32 VALU issue rate the Blake2s rooflines are
// quoted against, measured instead of assumed, and the latency of one Merkle
// node hash, which bounds every tree level too narrow to fill the device.
//
// Every lane runs CHAINS independent dependency chains of the three
// instructions a Blake2s G step is made of (v_add3_u32, v_xor_b32,
// v_alignbit_b32); a grid of many waves per SIMD keeps every SIMD issuing.
// The loop body is exactly 3 * CHAINS * UNROLL VALU instructions (checked in
// the gfx950 ISA of this file), so lane-ops = lanes * iterations * that count.
#include "common.hpp"
#include "merkle_engine.hpp"

namespace vx {

constexpr int VALU_CHAINS = 8;
constexpr int VALU_UNROLL = 8;

__global__ __launch_bounds__(256) void valu_peak_kernel(uint32_t iters, uint32_t seed, uint32_t* out) {
    uint32_t x[VALU_CHAINS], y[VALU_CHAINS];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int c = 0; c < VALU_CHAINS; c++) {
        x[c] = t * 2654435761u + c + seed;
        y[c] = (t ^ (c * 40503u)) + seed;
    }
    for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
        for (int u = 0; u < VALU_UNROLL; u++) {
#pragma unroll
            for (int c = 0; c < VALU_CHAINS; c++) {
                x[c] = x[c] + y[c] + seed;                            // v_add3_u32
                y[c] = y[c] ^ x[c];                                   // v_xor_b32
                y[c] = __builtin_amdgcn_alignbit(y[c], y[c], 7 + u);  // v_alignbit_b32 (rotate)
            }
        }
    }
    uint32_t acc = 0;
#pragma unroll
    for (int c = 0; c < VALU_CHAINS; c++) acc ^= x[c] ^ y[c];
    if (acc == 0x9E3779B9u) out[t] = acc;  // keeps the chains live; practically never stores
}

// One wave hashes n Merkle nodes in sequence, each node's children being the
// previous node (node = H(prev || prev), as a tree level of one node would):
// QUAD, every quad of the wave runs its own chain with b2s_compress_quad, the
// children passed through LDS as a narrow tree level does; otherwise every
// lane runs its own chain with b2s_node. Launched as a single wave, so the
// time per node is the latency of one compression with its level's overhead.
template <bool QUAD>
__global__ __launch_bounds__(64) void hash_latency_kernel(uint32_t n, uint32_t* out) {
    const uint32_t lane = threadIdx.x;
    if constexpr (QUAD) {
        __shared__ uint32_t msg[16 * 16];
        uint32_t* qm = msg + 16 * (lane >> 2);
        const uint32_t l = lane & 3;
        uint32_t hl, hh;
        b2s_quad_init(hl, hh);
        hl ^= lane;
        for (uint32_t i = 0; i < n; i++) {
            qm[l] = hl;
            qm[4 + l] = hh;
            qm[8 + l] = hl;
            qm[12 + l] = hh;
            VX_WAVE_SYNC();
            uint32_t nl, nh;
            b2s_quad_init(nl, nh);
            b2s_compress_quad(qm, 64u, 0xFFFFFFFFu, nl, nh);
            VX_WAVE_SYNC();
            hl = nl;
            hh = nh;
        }
        out[2 * lane] = hl;
        out[2 * lane + 1] = hh;
    } else {
        uint32_t h[8];
#pragma unroll
        for (int k = 0; k < 8; k++) h[k] = lane * 8 + k;
        for (uint32_t i = 0; i < n; i++) {
            uint32_t o[8];
            b2s_node(h, h, o);
#pragma unroll
            for (int k = 0; k < 8; k++) h[k] = o[k];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) out[8 * lane + k] = h[k];
    }
}

}  // namespace vx

using namespace vx;

extern "C" {

int32_t vx_measure_valu_peak(double* tops, double* ms_out) {
    if (!tops) return -1;
    int dev = 0, cus = 0;
    if (!check(hipGetDevice(&dev), "hipGetDevice") ||
        !check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute"))
        return -1;
    uint32_t* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const uint32_t blocks = (uint32_t)cus * 16, threads = 256, iters = 2048;
    bool ok = check(hipMalloc((void**)&out, 4ull * blocks * threads), "hipMalloc") &&
              check(hipEventCreate(&e0), "hipEventCreate") && check(hipEventCreate(&e1), "hipEventCreate");
    float ms = 0.f;
    if (ok) {
        valu_peak_kernel<<<blocks, threads>>>(iters, 1u, out);  // warm-up
        ok = check(hipEventRecord(e0, nullptr), "hipEventRecord");
        valu_peak_kernel<<<blocks, threads>>>(iters, 2u, out);
        ok = ok && check_launch("valu_peak_kernel") && check(hipEventRecord(e1, nullptr), "hipEventRecord") &&
             check(hipEventSynchronize(e1), "hipEventSynchronize") &&
             check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    }
    if (ok && ms > 0) {
        const double ops = (double)blocks * threads * iters * 3.0 * VALU_CHAINS * VALU_UNROLL;
        *tops = ops / (ms * 1e-3) / 1e12;
        if (ms_out) *ms_out = ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    return ok ? 0 : -1;
}

int32_t vx_measure_hash_latency(double* ns_quad, double* ns_lane) {
    if (!ns_quad || !ns_lane) {
        set_error(-2, "vx_measure_hash_latency: null output");
        return -1;
    }
    uint32_t* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = check(hipMalloc((void**)&out, 4 * 512), "hipMalloc") && check(hipEventCreate(&e0), "hipEventCreate") &&
              check(hipEventCreate(&e1), "hipEventCreate");
    // the difference of a long and a short chain cancels the launch overhead
    constexpr uint32_t N_LONG = 4096, N_SHORT = 1024;
    double* res[2] = {ns_quad, ns_lane};
    for (int q = 0; ok && q < 2; q++) {
        float ms[2] = {0.f, 0.f};
        for (int k = 0; ok && k < 2; k++) {
            const uint32_t n = k ? N_LONG : N_SHORT;
            auto launch = [&] {
                if (q == 0)
                    hash_latency_kernel<true><<<1, 64>>>(n, out);
                else
                    hash_latency_kernel<false><<<1, 64>>>(n, out);
            };
            launch();  // warm-up
            ok = check(hipEventRecord(e0, nullptr), "hipEventRecord");
            launch();
            ok = ok && check_launch("hash_latency_kernel") && check(hipEventRecord(e1, nullptr), "hipEventRecord") &&
                 check(hipEventSynchronize(e1), "hipEventSynchronize") &&
                 check(hipEventElapsedTime(&ms[k], e0, e1), "hipEventElapsedTime");
        }
        if (ok) *res[q] = (ms[1] - ms[0]) * 1e6 / (N_LONG - N_SHORT);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    return ok ? 0 : -1;
}

}  // extern "C"
