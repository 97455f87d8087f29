// stwo DEEP quotient on the device: the three kernels stwo's QuotientOps and
// AccumulationOps call (cuda/quotient.cu:25-253, cuda/accumulate_lift.cu:13-51;
// callers crates/cuda-backend/src/quotient_ops.rs:35-223,
// accumulation_ops.rs:46-88).
//
//  accumulate_numerators   out[row] = sum_i (c_i * col[idx_i][row] - b_i)
//  quotients_combine       out[row] = sum_j (numer_j[lift_j(row)] - a_j * y) / den_j(row)
//  accumulate_lift         col[i] += curr[lift(i)]   (QM31, 4 words per element)
//
// accumulate_numerators is bound by HBM (4 bytes of column a term): the sum of
// products c_i * f is kept in 64 bits (one v_mad_u64_u32 per coordinate and
// column) and reduced once every three columns, and sum_i b_i is formed once
// per thread. quotients_combine is bound by the VALU: den_j(row) is the CM31
//   (Re(px) - x) * Im(py) - (Re(py) - y) * Im(px)
//     = (Re(px) Im(py) - Re(py) Im(px)) - x Im(py) + y Im(px),
// linear in the domain point, so its constant part is formed once per thread;
// each thread takes COMBINE_ROWS rows (wave-strided, so every load stays
// coalesced) and inverts their norms |den|^2 with one M31 inversion per
// accumulation (Montgomery's trick) instead of one per row.
#include "common.hpp"
#include "field.hpp"

namespace vx {
namespace {

constexpr uint32_t DEEP_THREADS = 256;
constexpr uint32_t COMBINE_ROWS = 8;  // rows per thread of quotients_combine

// x < 2^64 -> x mod p (2^31 = 1 mod p: x = lo + mid 2^31 + hi 2^62 = lo + mid + hi).
__device__ __forceinline__ uint32_t reduce64(uint64_t x) {
    const uint64_t s = (x & P) + ((x >> 31) & P) + (x >> 62);  // < 2^32 + 2
    return canon((uint32_t)(s & P) + (uint32_t)(s >> 31));
}

__global__ __launch_bounds__(DEEP_THREADS) void accumulate_numerators_kernel(
    const uint32_t* const* __restrict__ col_ptrs, const uint32_t* __restrict__ col_indices,
    const uint32_t* __restrict__ b_coeffs, const uint32_t* __restrict__ c_coeffs, uint32_t n_batch, uint32_t n_rows,
    uint32_t* __restrict__ out0, uint32_t* __restrict__ out1, uint32_t* __restrict__ out2,
    uint32_t* __restrict__ out3) {
    const uint32_t row = blockIdx.x * DEEP_THREADS + threadIdx.x;
    if (row >= n_rows) return;
    // sum_i b_i: the same for every row
    uint32_t bs[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < n_batch; i++)
#pragma unroll
        for (int k = 0; k < 4; k++) bs[k] = madd(bs[k], b_coeffs[4 * i + k]);
    uint64_t acc[4] = {0, 0, 0, 0};
    uint32_t i = 0;
    // three products of < 2^62 on top of a reduced value stay below 2^64
    for (; i + 3 <= n_batch; i += 3) {
        const uint32_t f0 = col_ptrs[col_indices[i]][row];
        const uint32_t f1 = col_ptrs[col_indices[i + 1]][row];
        const uint32_t f2 = col_ptrs[col_indices[i + 2]][row];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint64_t a = acc[k];
            a += (uint64_t)c_coeffs[4 * i + k] * f0;
            a += (uint64_t)c_coeffs[4 * (i + 1) + k] * f1;
            a += (uint64_t)c_coeffs[4 * (i + 2) + k] * f2;
            acc[k] = reduce64(a);
        }
    }
    for (; i < n_batch; i++) {
        const uint32_t f = col_ptrs[col_indices[i]][row];
#pragma unroll
        for (int k = 0; k < 4; k++) acc[k] = reduce64(acc[k] + (uint64_t)c_coeffs[4 * i + k] * f);
    }
    out0[row] = msub((uint32_t)acc[0], bs[0]);
    out1[row] = msub((uint32_t)acc[1], bs[1]);
    out2[row] = msub((uint32_t)acc[2], bs[2]);
    out3[row] = msub((uint32_t)acc[3], bs[3]);
}

// Thread t of workgroup g takes rows base + 64 * r + lane, r < COMBINE_ROWS, where
// base = (g * 4 + wave) * 64 * COMBINE_ROWS: consecutive lanes, consecutive rows.
__global__ __launch_bounds__(DEEP_THREADS) void quotients_combine_kernel(
    const uint32_t* __restrict__ spx, const uint32_t* __restrict__ spy, const uint32_t* __restrict__ fla,
    const uint32_t* const* __restrict__ n0, const uint32_t* const* __restrict__ n1,
    const uint32_t* const* __restrict__ n2, const uint32_t* const* __restrict__ n3,
    const uint32_t* __restrict__ acc_logs, uint32_t n_accs, const uint32_t* __restrict__ dxs,
    const uint32_t* __restrict__ dys, uint32_t lifting_log, uint32_t n_rows, uint32_t* __restrict__ out0,
    uint32_t* __restrict__ out1, uint32_t* __restrict__ out2, uint32_t* __restrict__ out3) {
    constexpr uint32_t R = COMBINE_ROWS;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t base = ((uint64_t)blockIdx.x * (DEEP_THREADS / 64) + (threadIdx.x >> 6)) * (64u * R);
    if (base >= n_rows) return;
    uint32_t dx[R], dy[R];
    bool live[R];
#pragma unroll
    for (uint32_t r = 0; r < R; r++) {
        const uint64_t row = base + 64u * r + lane;
        live[r] = row < n_rows;
        dx[r] = live[r] ? dxs[row] : 1u;
        dy[r] = live[r] ? dys[row] : 0u;
    }
    QM31 q[R];
#pragma unroll
    for (uint32_t r = 0; r < R; r++) q[r] = {{0, 0, 0, 0}};
    for (uint32_t j = 0; j < n_accs; j++) {
        const CM31 prx = {spx[4 * j], spx[4 * j + 1]}, pix = {spx[4 * j + 2], spx[4 * j + 3]};
        const CM31 pry = {spy[4 * j], spy[4 * j + 1]}, piy = {spy[4 * j + 2], spy[4 * j + 3]};
        const CM31 k0 = csub(cmul(prx, piy), cmul(pry, pix));
        const QM31 a = {{fla[4 * j], fla[4 * j + 1], fla[4 * j + 2], fla[4 * j + 3]}};
        const uint32_t sh = lifting_log - acc_logs[j] + 1;
        const uint32_t *p0 = n0[j], *p1 = n1[j], *p2 = n2[j], *p3 = n3[j];
        CM31 den[R];
        uint32_t nrm[R], pre[R];  // norms |den|^2 and their prefix products
        uint32_t run = 1;
#pragma unroll
        for (uint32_t r = 0; r < R; r++) {
            // den = k0 - x Im(py) + y Im(px); a dead row gets a unit denominator
            const uint64_t da = (uint64_t)dy[r] * pix.a + (uint64_t)(P - dx[r]) * piy.a + k0.a;
            const uint64_t db = (uint64_t)dy[r] * pix.b + (uint64_t)(P - dx[r]) * piy.b + k0.b;
            den[r] = live[r] ? CM31{reduce64(da), reduce64(db)} : CM31{1u, 0u};
            nrm[r] = reduce64((uint64_t)den[r].a * den[r].a + (uint64_t)den[r].b * den[r].b);
            pre[r] = run;
            // a zero denominator (a domain point on the sample line) inverts to 0, as
            // cm31_inv(0) does; it is kept out of the running product
            run = mmul(run, nrm[r] ? nrm[r] : 1u);
        }
        uint32_t inv = minv(run);
#pragma unroll
        for (int r = R - 1; r >= 0; r--) {
            const uint32_t ninv = nrm[r] ? mmul(inv, pre[r]) : 0u;
            inv = mmul(inv, nrm[r] ? nrm[r] : 1u);
            const CM31 dinv = {mmul(den[r].a, ninv), mneg(mmul(den[r].b, ninv))};
            if (!live[r]) continue;
            const uint64_t row = base + 64u * r + lane;
            const uint64_t li = ((row >> sh) << 1) + (row & 1);
            const QM31 nv = {{p0[li], p1[li], p2[li], p3[li]}};
            const QM31 full = qsub(nv, qmul_m31(a, dy[r]));
            const CM31 fa = cmul({full.v[0], full.v[1]}, dinv), fb = cmul({full.v[2], full.v[3]}, dinv);
            q[r] = qadd(q[r], {{fa.a, fa.b, fb.a, fb.b}});
        }
    }
#pragma unroll
    for (uint32_t r = 0; r < R; r++) {
        if (!live[r]) continue;
        const uint64_t row = base + 64u * r + lane;
        out0[row] = q[r].v[0];
        out1[row] = q[r].v[1];
        out2[row] = q[r].v[2];
        out3[row] = q[r].v[3];
    }
}

// VEC: both arrays 16-byte aligned, one uint4 per element; else four words.
template <bool VEC>
__global__ __launch_bounds__(DEEP_THREADS) void accumulate_lift_kernel(uint32_t* __restrict__ col,
                                                                       const uint32_t* __restrict__ curr,
                                                                       uint32_t col_n, uint32_t log_ratio) {
    const uint32_t i = blockIdx.x * DEEP_THREADS + threadIdx.x;
    if (i >= col_n) return;
    const uint32_t s = ((i >> (log_ratio + 1)) << 1) | (i & 1u);
    if (VEC) {
        uint4* c4 = (uint4*)col + i;
        const uint4 a = *c4, b = ((const uint4*)curr)[s];
        *c4 = make_uint4(madd(a.x, b.x), madd(a.y, b.y), madd(a.z, b.z), madd(a.w, b.w));
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) col[4ull * i + k] = madd(col[4ull * i + k], curr[4ull * s + k]);
    }
}

}  // namespace
}  // namespace vx

using namespace vx;

extern "C" {

void cuda_accumulate_numerators(const uint32_t* const* col_ptrs, const uint32_t* col_indices,
                                const uint32_t* b_coeffs, const uint32_t* c_coeffs, uint32_t n_batch_cols,
                                uint32_t n_rows, uint32_t* out0, uint32_t* out1, uint32_t* out2, uint32_t* out3) {
    if (!n_rows) return;
    accumulate_numerators_kernel<<<ceil_div(n_rows, DEEP_THREADS), DEEP_THREADS, 0, current_stream()>>>(
        col_ptrs, col_indices, b_coeffs, c_coeffs, n_batch_cols, n_rows, out0, out1, out2, out3);
    check_launch("cuda_accumulate_numerators");
}

void cuda_compute_quotients_combine(const uint32_t* sample_points_x, const uint32_t* sample_points_y,
                                    const uint32_t* first_linear_acc, const uint32_t* const* numer_ptrs0,
                                    const uint32_t* const* numer_ptrs1, const uint32_t* const* numer_ptrs2,
                                    const uint32_t* const* numer_ptrs3, const uint32_t* acc_log_sizes,
                                    uint32_t n_accs, const uint32_t* domain_xs, const uint32_t* domain_ys,
                                    uint32_t lifting_log_size, uint32_t n_rows, uint32_t* out0, uint32_t* out1,
                                    uint32_t* out2, uint32_t* out3) {
    if (!n_rows) return;
    const uint64_t rows_per_block = (uint64_t)DEEP_THREADS * COMBINE_ROWS;
    quotients_combine_kernel<<<ceil_div(n_rows, rows_per_block), DEEP_THREADS, 0, current_stream()>>>(
        sample_points_x, sample_points_y, first_linear_acc, numer_ptrs0, numer_ptrs1, numer_ptrs2, numer_ptrs3,
        acc_log_sizes, n_accs, domain_xs, domain_ys, lifting_log_size, n_rows, out0, out1, out2, out3);
    check_launch("cuda_compute_quotients_combine");
}

void cuda_accumulate_lift(uint32_t* col, const uint32_t* curr, uint32_t col_n, uint32_t log_ratio) {
    if (!col_n) return;
    if ((((uintptr_t)col | (uintptr_t)curr) & 15) == 0)
        accumulate_lift_kernel<true>
            <<<ceil_div(col_n, DEEP_THREADS), DEEP_THREADS, 0, current_stream()>>>(col, curr, col_n, log_ratio);
    else
        accumulate_lift_kernel<false>
            <<<ceil_div(col_n, DEEP_THREADS), DEEP_THREADS, 0, current_stream()>>>(col, curr, col_n, log_ratio);
    check_launch("cuda_accumulate_lift");
}

}  // extern "C"
