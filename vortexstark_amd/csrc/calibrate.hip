// Device calibration: the int32 VALU issue rate the Blake2s rooflines are
// quoted against, measured instead of assumed.
//
// Every lane runs CHAINS independent dependency chains of the three
// instructions a Blake2s G step is made of (v_add3_u32, v_xor_b32,
// v_alignbit_b32); a grid of many waves per SIMD keeps every SIMD issuing.
// The loop body is exactly 3 * CHAINS * UNROLL VALU instructions (checked in
// the gfx950 ISA of this file), so lane-ops = lanes * iterations * that count.
#include "common.hpp"

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

}  // extern "C"
