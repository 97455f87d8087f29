// Reference-ABI FRI fold, Fibonacci quotient and zero-pad kernels
// (cuda/fri.cu:10-121, cuda/constraint.cu:10-153). These are the per-call
// entry points the reference prover binds; the fused prover path computes the
// same values inside its commit kernels instead of materialising them.
#include "common.hpp"
#include "field.hpp"

namespace vx {

__global__ void fold_line_soa_kernel(const uint32_t* __restrict__ in0, const uint32_t* __restrict__ in1,
                                     const uint32_t* __restrict__ in2, const uint32_t* __restrict__ in3,
                                     const uint32_t* __restrict__ tw, uint32_t* __restrict__ out0,
                                     uint32_t* __restrict__ out1, uint32_t* __restrict__ out2,
                                     uint32_t* __restrict__ out3, QM31 alpha, uint32_t half_n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= half_n) return;
    QM31 f0 = {{in0[2 * i], in1[2 * i], in2[2 * i], in3[2 * i]}};
    QM31 f1 = {{in0[2 * i + 1], in1[2 * i + 1], in2[2 * i + 1], in3[2 * i + 1]}};
    QM31 r = fold_pair(f0, f1, alpha, tw[i]);
    out0[i] = r.v[0];
    out1[i] = r.v[1];
    out2[i] = r.v[2];
    out3[i] = r.v[3];
}

__global__ void fold_circle_soa_kernel(uint32_t* __restrict__ d0, uint32_t* __restrict__ d1,
                                       uint32_t* __restrict__ d2, uint32_t* __restrict__ d3,
                                       const uint32_t* __restrict__ s0, const uint32_t* __restrict__ s1,
                                       const uint32_t* __restrict__ s2, const uint32_t* __restrict__ s3,
                                       const uint32_t* __restrict__ tw, QM31 alpha, QM31 alpha_sq,
                                       uint32_t half_n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= half_n) return;
    QM31 f0 = {{s0[2 * i], s1[2 * i], s2[2 * i], s3[2 * i]}};
    QM31 f1 = {{s0[2 * i + 1], s1[2 * i + 1], s2[2 * i + 1], s3[2 * i + 1]}};
    QM31 fp = fold_pair(f0, f1, alpha, tw[i]);
    QM31 prev = {{d0[i], d1[i], d2[i], d3[i]}};
    QM31 r = qadd(qmul(prev, alpha_sq), fp);
    d0[i] = r.v[0];
    d1[i] = r.v[1];
    d2[i] = r.v[2];
    d3[i] = r.v[3];
}

// q[i] = alpha * (t[(i+2) mod n] - (t[(i+1) mod n] + t[i]))  (constraint.cu:10-36)
__global__ void quotient_kernel(const uint32_t* __restrict__ t, uint32_t* __restrict__ o0,
                                uint32_t* __restrict__ o1, uint32_t* __restrict__ o2, uint32_t* __restrict__ o3,
                                QM31 alpha, uint32_t offset, uint32_t chunk_n, uint32_t n) {
    uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= chunk_n) return;
    uint32_t i = offset + li;
    // a full remainder, as the reference takes, so every n >= 1 wraps correctly
    uint32_t i1 = (i + 1) % n;
    uint32_t i2 = (i + 2) % n;
    uint32_t c = msub(t[i2], madd(t[i1], t[i]));
    QM31 r = qmul_m31(alpha, c);
    o0[li] = r.v[0];
    o1[li] = r.v[1];
    o2[li] = r.v[2];
    o3[li] = r.v[3];
}

__global__ void zero_pad_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, uint32_t src_n,
                                uint32_t dst_n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < dst_n) dst[i] = i < src_n ? src[i] : 0u;
}

static QM31 q_from_host(const uint32_t* a) { return {{a[0], a[1], a[2], a[3]}}; }

}  // namespace vx

using namespace vx;

extern "C" {

void cuda_fold_line_soa(const uint32_t* in0, const uint32_t* in1, const uint32_t* in2, const uint32_t* in3,
                        const uint32_t* twiddles, uint32_t* out0, uint32_t* out1, uint32_t* out2, uint32_t* out3,
                        const uint32_t* alpha, uint32_t half_n) {
    if (!half_n) return;
    fold_line_soa_kernel<<<ceil_div(half_n, 256), 256, 0, current_stream()>>>(
        in0, in1, in2, in3, twiddles, out0, out1, out2, out3, q_from_host(alpha), half_n);
    check_launch("cuda_fold_line_soa");
}

void cuda_fold_circle_into_line_soa(uint32_t* dst0, uint32_t* dst1, uint32_t* dst2, uint32_t* dst3,
                                    const uint32_t* src0, const uint32_t* src1, const uint32_t* src2,
                                    const uint32_t* src3, const uint32_t* twiddles, const uint32_t* alpha,
                                    const uint32_t* alpha_sq, uint32_t half_n) {
    if (!half_n) return;
    fold_circle_soa_kernel<<<ceil_div(half_n, 256), 256, 0, current_stream()>>>(
        dst0, dst1, dst2, dst3, src0, src1, src2, src3, twiddles, q_from_host(alpha), q_from_host(alpha_sq), half_n);
    check_launch("cuda_fold_circle_into_line_soa");
}

void cuda_fibonacci_quotient(const uint32_t* trace, uint32_t* out0, uint32_t* out1, uint32_t* out2, uint32_t* out3,
                             const uint32_t* alpha, uint32_t n) {
    cuda_fibonacci_quotient_chunk_stream(trace, out0, out1, out2, out3, alpha, 0, n, n, current_stream());
}
void cuda_fibonacci_quotient_chunk(const uint32_t* trace, uint32_t* out0, uint32_t* out1, uint32_t* out2,
                                   uint32_t* out3, const uint32_t* alpha, uint32_t offset, uint32_t chunk_n,
                                   uint32_t global_n) {
    cuda_fibonacci_quotient_chunk_stream(trace, out0, out1, out2, out3, alpha, offset, chunk_n, global_n,
                                         current_stream());
}
void cuda_fibonacci_quotient_chunk_stream(const uint32_t* trace, uint32_t* out0, uint32_t* out1, uint32_t* out2,
                                          uint32_t* out3, const uint32_t* alpha, uint32_t offset, uint32_t chunk_n,
                                          uint32_t global_n, void* stream) {
    if (!chunk_n) return;
    if (!global_n || (uint64_t)offset + chunk_n > global_n) {
        set_error(-2, "cuda_fibonacci_quotient: need offset + chunk_n <= global_n");
        return;
    }
    quotient_kernel<<<ceil_div(chunk_n, 256), 256, 0, (hipStream_t)stream>>>(
        trace, out0, out1, out2, out3, q_from_host(alpha), offset, chunk_n, global_n);
    check_launch("cuda_fibonacci_quotient");
}
void cuda_zero_pad(const uint32_t* src, uint32_t* dst, uint32_t src_n, uint32_t dst_n) {
    if (!dst_n) return;
    zero_pad_kernel<<<ceil_div(dst_n, 256), 256, 0, current_stream()>>>(src, dst, src_n, dst_n);
    check_launch("cuda_zero_pad");
}

}  // extern "C"
