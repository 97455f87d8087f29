// stwo DEEP quotient on the device: the three kernels stwo's QuotientOps and
// AccumulationOps call (cuda/quotient.cu:25-253, cuda/accumulate_lift.cu:13-51;
// callers crates/cuda-backend/src/quotient_ops.rs:35-223,
// accumulation_ops.rs:46-88).
//
//  accumulate_numerators   out[row] = sum_i (c_i * col[idx_i][row] - b_i)
//  quotients_combine       out[row] = sum_j (numer_j[lift_j(row)] - a_j * y) / den_j(row)
//  accumulate_lift         col[i] += curr[lift(i)]   (QM31, 4 words per element)
//
// accumulate_numerators is bound by HBM (4 bytes of column a term): each thread
// takes eight rows (two uint4 loads a column, a wave reading 2 KiB of it), the sum of products c_i * f is kept
// in 64 bits (one v_mad_u64_u32 per coordinate and column, a partial reduction
// every three columns, the full one at the end), and sum_i b_i is formed once
// per thread. quotients_combine is bound by the VALU: den_j(row) is the CM31
//   (Re(px) - x) * Im(py) - (Re(py) - y) * Im(px)
//     = (Re(px) Im(py) - Re(py) Im(px)) - x Im(py) + y Im(px),
// linear in the domain point, so its constant part is formed once per thread;
// each thread takes COMBINE_ROWS rows (wave-strided, so every load stays
// coalesced) and inverts their norms |den|^2 with one M31 inversion per
// accumulation (Montgomery's trick) instead of one per row.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "common.hpp"
#include "field.hpp"
#include "kernels.hpp"

namespace vx {
namespace {

constexpr uint32_t DEEP_THREADS = 256;
constexpr uint32_t COMBINE_ROWS = 8;  // rows per thread of quotients_combine

// x < 2^64 -> x mod p (2^31 = 1 mod p: x = lo + mid 2^31 + hi 2^62 = lo + mid + hi).
__device__ __forceinline__ uint32_t reduce64(uint64_t x) {
    const uint64_t s = (x & P) + ((x >> 31) & P) + (x >> 62);  // < 2^32 + 2
    return canon((uint32_t)(s & P) + (uint32_t)(s >> 31));
}

// ROWS rows per thread (row_of): 8 (uint4 loads and stores; n_rows % 512 == 0
// and 16-byte aligned outputs, checked on the host; a column whose pointer is
// not 16-byte aligned is read word by word, a wave-uniform branch) or 1.
// Rows of a thread: ROWS = 1: row0; ROWS = 8: row0 .. row0 + 3 and row0 + 256 ..
// row0 + 259, so that a wave reads 2 KiB of a column contiguously (lane l at
// row0 = 512 * wave + 4 * l). Wider contiguous reads per wave and column stream
// faster from HBM when a thread reads many columns.
template <uint32_t ROWS>
__device__ __forceinline__ uint64_t row_of(uint64_t row0, uint32_t r) {
    return ROWS == 1 ? row0 : row0 + (r & 3) + 256 * (r >> 2);
}
template <uint32_t ROWS>
__device__ __forceinline__ void load_rows(const uint32_t* col, uint64_t row0, uint32_t f[ROWS]) {
    if (ROWS == 8 && ((uintptr_t)col & 15) == 0) {
        const uint4 v = *(const uint4*)(col + row0), w = *(const uint4*)(col + row0 + 256);
        f[0] = v.x;
        f[1] = v.y;
        f[2] = v.z;
        f[3] = v.w;
        f[4] = w.x;
        f[5] = w.y;
        f[6] = w.z;
        f[7] = w.w;
    } else {
#pragma unroll
        for (uint32_t r = 0; r < ROWS; r++) f[r] = col[row_of<ROWS>(row0, r)];
    }
}

template <uint32_t ROWS>
__global__ __launch_bounds__(DEEP_THREADS) void accumulate_numerators_kernel(
    const uint32_t* const* __restrict__ col_ptrs, const uint32_t* __restrict__ col_indices,
    const uint32_t* __restrict__ b_coeffs, const uint32_t* __restrict__ c_coeffs, uint32_t n_batch, uint32_t n_rows,
    uint32_t* __restrict__ out0, uint32_t* __restrict__ out1, uint32_t* __restrict__ out2,
    uint32_t* __restrict__ out3) {
    const uint64_t row0 = ROWS == 1 ? (uint64_t)blockIdx.x * DEEP_THREADS + threadIdx.x
                                    : ((uint64_t)blockIdx.x * (DEEP_THREADS / 64) + (threadIdx.x >> 6)) * 64 * ROWS +
                                          4 * (threadIdx.x & 63);
    if (row0 >= n_rows) return;
    // sum_i b_i: the same for every row
    uint32_t bs[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < n_batch; i++)
#pragma unroll
        for (int k = 0; k < 4; k++) bs[k] = madd(bs[k], b_coeffs[4 * i + k]);
    // The accumulators stay below 2^62 + 4 between groups: a group adds three
    // products of at most (p - 1)^2 < 2^62 - 2^32 (three v_mad_u64_u32, the 64-bit
    // addend chained), which stays below 2^64, and then folds bits 62..63 back
    // onto bit 0 (2^62 = 1 mod p): a partial reduction of a few 32-bit ops.
    constexpr uint64_t M62 = (1ull << 62) - 1;
    uint64_t acc[ROWS][4] = {};
    uint32_t i = 0;
    for (; i + 3 <= n_batch; i += 3) {
        uint32_t f[3][ROWS];
#pragma unroll
        for (int g = 0; g < 3; g++) load_rows<ROWS>(col_ptrs[col_indices[i + g]], row0, f[g]);
#pragma unroll
        for (uint32_t r = 0; r < ROWS; r++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                uint64_t a = acc[r][k];
                a = (uint64_t)c_coeffs[4 * i + k] * f[0][r] + a;
                a = (uint64_t)c_coeffs[4 * (i + 1) + k] * f[1][r] + a;
                a = (uint64_t)c_coeffs[4 * (i + 2) + k] * f[2][r] + a;
                acc[r][k] = (a & M62) + (a >> 62);
            }
    }
    for (; i < n_batch; i++) {
        uint32_t f[ROWS];
        load_rows<ROWS>(col_ptrs[col_indices[i]], row0, f);
#pragma unroll
        for (uint32_t r = 0; r < ROWS; r++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint64_t a = (uint64_t)c_coeffs[4 * i + k] * f[r] + acc[r][k];
                acc[r][k] = (a & M62) + (a >> 62);
            }
    }
    uint32_t o[4][ROWS];
#pragma unroll
    for (uint32_t r = 0; r < ROWS; r++)
#pragma unroll
        for (int k = 0; k < 4; k++) o[k][r] = msub(reduce64(acc[r][k]), bs[k]);
    uint32_t* outs[4] = {out0, out1, out2, out3};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (ROWS == 8) {
            *(uint4*)(outs[k] + row0) = make_uint4(o[k][0], o[k][1], o[k][2], o[k][3]);
            *(uint4*)(outs[k] + row0 + 256) = make_uint4(o[k][4], o[k][5], o[k][6], o[k][7]);
        } else {
            outs[k][row0] = o[k][0];
        }
    }
}

// Thread t of workgroup g takes rows base + 64 * r + lane, r < R, where
// base = (g * 4 + wave) * 64 * R: consecutive lanes, consecutive rows. Each
// accumulation's numerators are loaded before its denominators are formed and
// inverted, so the loads are in flight during that work.
template <uint32_t R>
__global__ __launch_bounds__(DEEP_THREADS) void quotients_combine_kernel(
    const uint32_t* __restrict__ spx, const uint32_t* __restrict__ spy, const uint32_t* __restrict__ fla,
    const uint32_t* const* __restrict__ n0, const uint32_t* const* __restrict__ n1,
    const uint32_t* const* __restrict__ n2, const uint32_t* const* __restrict__ n3,
    const uint32_t* __restrict__ acc_logs, uint32_t n_accs, const uint32_t* __restrict__ dxs,
    const uint32_t* __restrict__ dys, uint32_t lifting_log, uint32_t n_rows, uint32_t* __restrict__ out0,
    uint32_t* __restrict__ out1, uint32_t* __restrict__ out2, uint32_t* __restrict__ out3) {
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
        const uint32_t sh = lifting_log - acc_logs[j] + 1;
        const uint32_t *p0 = n0[j], *p1 = n1[j], *p2 = n2[j], *p3 = n3[j];
        uint32_t nv[R][4];
#pragma unroll
        for (uint32_t r = 0; r < R; r++) {
            const uint64_t row = base + 64u * r + lane;
            const uint64_t li = live[r] ? ((row >> sh) << 1) + (row & 1) : 0;
            nv[r][0] = p0[li];
            nv[r][1] = p1[li];
            nv[r][2] = p2[li];
            nv[r][3] = p3[li];
        }
        const CM31 prx = {spx[4 * j], spx[4 * j + 1]}, pix = {spx[4 * j + 2], spx[4 * j + 3]};
        const CM31 pry = {spy[4 * j], spy[4 * j + 1]}, piy = {spy[4 * j + 2], spy[4 * j + 3]};
        const CM31 k0 = csub(cmul(prx, piy), cmul(pry, pix));
        const QM31 a = {{fla[4 * j], fla[4 * j + 1], fla[4 * j + 2], fla[4 * j + 3]}};  // canonical
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
            // 1 / den = d0 - d1 i
            const uint64_t d0 = mmul(den[r].a, ninv), d1 = mmul(den[r].b, ninv);
            // full = numer - a y, then full * (d0 - d1 i), one reduction per coordinate:
            // (f0 + f1 i)(d0 - d1 i) = (f0 d0 + f1 d1) + (f1 d0 - f0 d1) i, likewise for f2, f3
            const uint64_t y = dy[r];
            const uint64_t f0 = reduce64(nv[r][0] + (P - a.v[0]) * y), f1 = reduce64(nv[r][1] + (P - a.v[1]) * y);
            const uint64_t f2 = reduce64(nv[r][2] + (P - a.v[2]) * y), f3 = reduce64(nv[r][3] + (P - a.v[3]) * y);
            q[r].v[0] = madd(q[r].v[0], reduce64(f0 * d0 + f1 * d1));
            q[r].v[1] = madd(q[r].v[1], reduce64(f1 * d0 + (P - f0) * d1));
            q[r].v[2] = madd(q[r].v[2], reduce64(f2 * d0 + f3 * d1));
            q[r].v[3] = madd(q[r].v[3], reduce64(f3 * d0 + (P - f2) * d1));
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

// ---------------------------------------------------------------- fused form
// One launch for the whole quotient (no reference counterpart): for every row
// of the lifting domain,
//   out[row] = sum_p (sum_c (c_pc * col_c[lift_c(row)] - b_pc) - (sum_c a_pc) * y) / den_p(row),
// which is what accumulate_numerators per (point, size) batch followed by
// quotients_combine compute, summed in another order (exact field arithmetic,
// so the same words). Each column value is read once for all sample points and
// each (row, point) forms and inverts its denominator once, where the
// reference pipeline reads the columns once per point, writes and re-reads the
// numerators, and inverts once per (row, accumulation).
//
// Metadata (device words): per point x[4] y[4] A[4] B[4] (A, B the sums of the
// point's a and b coefficients); per column ptr lo, ptr hi, lift shift, 0; per
// (column, point) c[4].
constexpr uint32_t DEEP_MAX_POINTS = 4;

// Values of columns c .. c + G - 1 at the rows of this thread (lifted indices).
template <uint32_t G, uint32_t R>
__device__ __forceinline__ void fused_load(const uint32_t* __restrict__ cm, uint32_t c, const uint32_t row[R],
                                           uint32_t f[3][R]) {
#pragma unroll
    for (uint32_t g = 0; g < G; g++) {
        const uint32_t* ptr = (const uint32_t*)(uintptr_t)(cm[4 * (c + g)] | ((uint64_t)cm[4 * (c + g) + 1] << 32));
        const uint32_t sh = cm[4 * (c + g) + 2];
#pragma unroll
        for (uint32_t r = 0; r < R; r++) f[g][r] = ptr[((row[r] >> sh) << 1) | (row[r] & 1)];
    }
}

// acc += c_pc * f for columns c .. c + G - 1 (G <= 3): three products on top of
// a partially reduced value (< 2^62 + 4) stay below 2^64, then bits 62..63 fold
// onto bit 0.
template <uint32_t G, uint32_t NP, uint32_t R>
__device__ __forceinline__ void fused_accumulate(const uint32_t* __restrict__ cf, uint32_t c, const uint32_t f[3][R],
                                                 uint64_t acc[NP][R][4]) {
    constexpr uint64_t M62 = (1ull << 62) - 1;
#pragma unroll
    for (uint32_t p = 0; p < NP; p++)
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (uint32_t r = 0; r < R; r++) {
                uint64_t a = acc[p][r][k];
#pragma unroll
                for (uint32_t g = 0; g < G; g++) a = (uint64_t)cf[((c + g) * NP + p) * 4 + k] * f[g][r] + a;
                acc[p][r][k] = (a & M62) + (a >> 62);
            }
}

template <uint32_t NP, uint32_t R>
__global__ __launch_bounds__(DEEP_THREADS) void deep_fused_kernel(const uint32_t* __restrict__ meta, uint32_t n_cols,
                                                                  const uint32_t* __restrict__ dxs,
                                                                  const uint32_t* __restrict__ dys, uint32_t n_rows,
                                                                  uint32_t* __restrict__ out0,
                                                                  uint32_t* __restrict__ out1,
                                                                  uint32_t* __restrict__ out2,
                                                                  uint32_t* __restrict__ out3) {
    const uint32_t lane = threadIdx.x & 63u;
    // n_rows <= 2^30: row indices fit 32 bits
    const uint32_t base = (blockIdx.x * (DEEP_THREADS / 64) + (threadIdx.x >> 6)) * (64u * R);
    if (base >= n_rows) return;
    const uint32_t* pts = meta;
    const uint32_t* cm = meta + 16 * NP;
    const uint32_t* cf = cm + 4 * n_cols;
    uint32_t row[R];  // rows past the end read the last row; their results are dropped
    bool live[R];
#pragma unroll
    for (uint32_t r = 0; r < R; r++) {
        const uint32_t x = base + 64u * r + lane;
        live[r] = x < n_rows;
        row[r] = live[r] ? x : n_rows - 1;
    }
    uint64_t acc[NP][R][4] = {};
    // The one or two columns past a multiple of three first (while the first
    // whole group's loads are in flight), then whole groups of three, the next
    // group's loads in flight while a group is summed.
    const uint32_t c0 = n_cols % 3;
    uint32_t fa[3][R], fb[3][R];
    if (c0 < n_cols) fused_load<3, R>(cm, c0, row, fa);
    if (c0 == 1) {
        fused_load<1, R>(cm, 0, row, fb);
        fused_accumulate<1, NP, R>(cf, 0, fb, acc);
    } else if (c0 == 2) {
        fused_load<2, R>(cm, 0, row, fb);
        fused_accumulate<2, NP, R>(cf, 0, fb, acc);
    }
    if (c0 < n_cols) {
        for (uint32_t c = c0;;) {
            const bool more_b = c + 3 < n_cols;
            if (more_b) fused_load<3, R>(cm, c + 3, row, fb);
            fused_accumulate<3, NP, R>(cf, c, fa, acc);
            c += 3;
            if (!more_b) break;
            const bool more_a = c + 3 < n_cols;
            if (more_a) fused_load<3, R>(cm, c + 3, row, fa);
            fused_accumulate<3, NP, R>(cf, c, fb, acc);
            c += 3;
            if (!more_a) break;
        }
    }
    uint32_t dx[R], dy[R];
#pragma unroll
    for (uint32_t r = 0; r < R; r++) {
        dx[r] = live[r] ? dxs[row[r]] : 1u;
        dy[r] = live[r] ? dys[row[r]] : 0u;
    }
    QM31 q[R];
#pragma unroll
    for (uint32_t r = 0; r < R; r++) q[r] = {{0, 0, 0, 0}};
#pragma unroll
    for (uint32_t p = 0; p < NP; p++) {
        const uint32_t* pp = pts + 16 * p;
        const CM31 prx = {pp[0], pp[1]}, pix = {pp[2], pp[3]};
        const CM31 pry = {pp[4], pp[5]}, piy = {pp[6], pp[7]};
        const CM31 k0 = csub(cmul(prx, piy), cmul(pry, pix));
        CM31 den[R];
        uint32_t nrm[R], pre[R];
        uint32_t run = 1;
#pragma unroll
        for (uint32_t r = 0; r < R; r++) {
            const uint64_t da = (uint64_t)dy[r] * pix.a + (uint64_t)(P - dx[r]) * piy.a + k0.a;
            const uint64_t db = (uint64_t)dy[r] * pix.b + (uint64_t)(P - dx[r]) * piy.b + k0.b;
            den[r] = live[r] ? CM31{reduce64(da), reduce64(db)} : CM31{1u, 0u};
            nrm[r] = reduce64((uint64_t)den[r].a * den[r].a + (uint64_t)den[r].b * den[r].b);
            pre[r] = run;
            run = mmul(run, nrm[r] ? nrm[r] : 1u);  // a zero denominator inverts to 0 (cm31_inv(0))
        }
        uint32_t inv = minv(run);
#pragma unroll
        for (int r = R - 1; r >= 0; r--) {
            const uint32_t ninv = nrm[r] ? mmul(inv, pre[r]) : 0u;
            inv = mmul(inv, nrm[r] ? nrm[r] : 1u);
            const uint64_t d0 = mmul(den[r].a, ninv), d1 = mmul(den[r].b, ninv);
            const uint64_t y = dy[r];
            uint64_t f[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                // numerator - B - A y, A and B canonical
                const uint32_t num = msub(reduce64(acc[p][r][k]), pp[12 + k]);
                f[k] = reduce64(num + (P - pp[8 + k]) * y);
            }
            q[r].v[0] = madd(q[r].v[0], reduce64(f[0] * d0 + f[1] * d1));
            q[r].v[1] = madd(q[r].v[1], reduce64(f[1] * d0 + (P - f[0]) * d1));
            q[r].v[2] = madd(q[r].v[2], reduce64(f[2] * d0 + f[3] * d1));
            q[r].v[3] = madd(q[r].v[3], reduce64(f[3] * d0 + (P - f[2]) * d1));
        }
    }
#pragma unroll
    for (uint32_t r = 0; r < R; r++) {
        if (!live[r]) continue;
        out0[row[r]] = q[r].v[0];
        out1[row[r]] = q[r].v[1];
        out2[row[r]] = q[r].v[2];
        out3[row[r]] = q[r].v[3];
    }
}

// Points of the canonic circle domain of size 2^log_n at bit-reversed rows:
// natural index k < n/2 is half_odds(log_n - 1).at(k), the rest their conjugates.
__global__ __launch_bounds__(DEEP_THREADS) void domain_points_kernel(CPt initial, CPt step, uint32_t log_n,
                                                                     uint32_t* __restrict__ xs,
                                                                     uint32_t* __restrict__ ys) {
    const uint64_t row = (uint64_t)blockIdx.x * DEEP_THREADS + threadIdx.x;
    if (row >= (1ull << log_n)) return;
    const uint32_t k = brev_bits((uint32_t)row, log_n), half = (1u << log_n) >> 1;
    const CPt pt = cp_pow_mul(initial, step, k < half ? k : k - half);
    xs[row] = pt.x;
    ys[row] = k < half ? pt.y : mneg(pt.y);
}

inline uint32_t hadd31(uint32_t a, uint32_t b) {
    const uint32_t r = a + b;
    return r >= P ? r - P : r;
}

}  // namespace
}  // namespace vx

using namespace vx;

// Metadata goes through two pinned staging / device slots used in turn: a call
// waits (host) only for the copy of the call two back, and its stream waits for
// that call's kernel before overwriting the device slot.
// Metadata of one fused DEEP launch, written to h (deep_meta_words words):
// per point 16 words (x, y, then the sums A = sum_c a_pc and B = sum_c b_pc),
// per column 4 words (pointer, lifting shift, 0), then per (column, point) the
// 4 words of c_pc.
static uint64_t deep_meta_words(uint32_t n_cols, uint32_t n_points) {
    return 16ull * n_points + 4ull * n_cols + 4ull * n_cols * n_points;
}
static void deep_meta(uint32_t* h, const uint32_t* const* cols, const uint32_t* col_log_sizes, uint32_t n_cols,
                      uint32_t n_points, const uint32_t* points, const uint32_t* line_coeffs,
                      uint32_t lifting_log_size) {
    const uint32_t NP = n_points;
    for (uint32_t p = 0; p < NP; p++) {
        uint32_t* pp = h + 16 * p;
        for (int k = 0; k < 8; k++) pp[k] = points[8 * p + k] % P;
        for (int k = 0; k < 8; k++) pp[8 + k] = 0;
        for (uint32_t c = 0; c < n_cols; c++) {
            const uint32_t* lc = line_coeffs + 12ull * ((uint64_t)p * n_cols + c);
            for (int k = 0; k < 4; k++) {
                pp[8 + k] = hadd31(pp[8 + k], lc[k] % P);       // A += a
                pp[12 + k] = hadd31(pp[12 + k], lc[4 + k] % P);  // B += b
            }
        }
    }
    uint32_t* cm = h + 16 * NP;
    uint32_t* cf = cm + 4 * n_cols;
    for (uint32_t c = 0; c < n_cols; c++) {
        const uint64_t ptr = (uint64_t)(uintptr_t)cols[c];
        cm[4 * c] = (uint32_t)ptr;
        cm[4 * c + 1] = (uint32_t)(ptr >> 32);
        cm[4 * c + 2] = lifting_log_size - col_log_sizes[c] + 1;
        cm[4 * c + 3] = 0;
        for (uint32_t p = 0; p < NP; p++)
            for (int k = 0; k < 4; k++)
                cf[(c * NP + p) * 4 + k] = line_coeffs[12ull * ((uint64_t)p * n_cols + c) + 8 + k] % P;
    }
}

// Argument checks shared by the fused DEEP entry points.
static bool deep_args_ok(const char* who, const uint32_t* const* cols, const uint32_t* col_log_sizes, uint32_t n_cols,
                         uint32_t max_cols, uint32_t n_points, uint32_t max_points, const uint32_t* points,
                         const uint32_t* line_coeffs, uint32_t lifting_log_size, const uint32_t* domain_xs,
                         const uint32_t* domain_ys, const uint32_t* out0, const uint32_t* out1, const uint32_t* out2,
                         const uint32_t* out3) {
    std::string msg;
    if (!cols || !col_log_sizes || !points || !line_coeffs || !domain_xs || !domain_ys || !out0 || !out1 || !out2 ||
        !out3) {
        msg = std::string(who) + ": null argument";
        set_error(hipErrorInvalidValue, msg.c_str());
        return false;
    }
    if (!n_cols || n_cols > max_cols || !n_points || n_points > max_points || lifting_log_size < 1 ||
        lifting_log_size > 30) {
        msg = std::string(who) + ": column or point count beyond the limit, or lifting log outside 1..30";
        set_error(-2, msg.c_str());
        return false;
    }
    for (uint32_t c = 0; c < n_cols; c++)
        if (col_log_sizes[c] < 1 || col_log_sizes[c] > lifting_log_size || !cols[c]) {
            msg = std::string(who) + ": column log sizes must be in 1..lifting_log_size";
            set_error(-2, msg.c_str());
            return false;
        }
    return true;
}

// One fused DEEP launch over 2^lifting_log_size rows with its metadata at meta (device).
static bool launch_deep_fused(const uint32_t* meta, uint32_t NP, uint32_t n_cols, uint32_t lifting_log_size,
                              const uint32_t* domain_xs, const uint32_t* domain_ys, uint32_t* out0, uint32_t* out1,
                              uint32_t* out2, uint32_t* out3, hipStream_t s) {
    const uint32_t n_rows = 1u << lifting_log_size;
    switch (NP) {
#define VX_DEEP_LAUNCH(np, r)                                                                                  \
    case np:                                                                                                   \
        deep_fused_kernel<np, r><<<ceil_div(n_rows, (uint64_t)DEEP_THREADS * r), DEEP_THREADS, 0, s>>>(        \
            meta, n_cols, domain_xs, domain_ys, n_rows, out0, out1, out2, out3);                               \
        break;
        VX_DEEP_LAUNCH(1, 4)
        VX_DEEP_LAUNCH(2, 4)
        VX_DEEP_LAUNCH(3, 2)
        VX_DEEP_LAUNCH(4, 2)
#undef VX_DEEP_LAUNCH
    }
    return check_launch("deep_fused_kernel");
}

struct vx_deep {
    uint32_t max_cols = 0, max_points = 0;
    size_t bytes = 0;
    uint32_t* h_meta[2] = {nullptr, nullptr};  // pinned staging
    uint32_t* d_meta[2] = {nullptr, nullptr};
    hipEvent_t copied[2] = {nullptr, nullptr}, done[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
    uint32_t next = 0;
};

// ---------------------------------------------------------------- allocation-free form
// vx_deep_quotient_noalloc keeps its launch metadata in this file-scope device
// array instead of memory the library allocates: each GPU holds its own copy
// until the process ends, and every call on a GPU reuses it (calls that overlap
// on one GPU race on it). Host code must never pass the array's name to a
// kernel or a copy (on the host it is only a stand-in for the symbol); it uses
// the device address deep_scratch() looked up with hipGetSymbolAddress.
constexpr uint32_t DEEP_SCRATCH_WORDS = 8192;  // 32 KiB
static __device__ __align__(256) uint32_t scratch[DEEP_SCRATCH_WORDS];

// Metadata words per staging launch: they travel by value in the kernel
// arguments (2 KiB), so staging needs no host buffer the device reads.
constexpr uint32_t META_CHUNK_WORDS = 512;
struct MetaChunk {
    uint32_t w[META_CHUNK_WORDS];
};
__global__ void stage_meta_kernel(MetaChunk c, uint32_t* dst, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = c.w[i];
}

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_addr[MAX_DEVICES] = {};
}  // namespace

// Device address of `scratch` on the current device: looked up on the first
// call made with that device current, then kept.
static uint32_t* deep_scratch() {
    int dev = 0;
    if (!check(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error(-2, "vx_deep_quotient_noalloc: device index beyond 63");
        return nullptr;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        if (!check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)") || !p)
            return nullptr;
        g_scratch_addr[dev] = (uint32_t*)p;
    }
    return g_scratch_addr[dev];
}

extern "C" {

void cuda_accumulate_numerators(const uint32_t* const* col_ptrs, const uint32_t* col_indices,
                                const uint32_t* b_coeffs, const uint32_t* c_coeffs, uint32_t n_batch_cols,
                                uint32_t n_rows, uint32_t* out0, uint32_t* out1, uint32_t* out2, uint32_t* out3) {
    if (!n_rows) return;
    // whole 512-row wave chunks and 16-byte aligned outputs: the wide form
    const bool vec = n_rows % 512 == 0 && (((uintptr_t)out0 | (uintptr_t)out1 | (uintptr_t)out2 | (uintptr_t)out3) & 15) == 0;
    if (vec)
        accumulate_numerators_kernel<8><<<ceil_div(n_rows, 8 * DEEP_THREADS), DEEP_THREADS, 0, current_stream()>>>(
            col_ptrs, col_indices, b_coeffs, c_coeffs, n_batch_cols, n_rows, out0, out1, out2, out3);
    else
        accumulate_numerators_kernel<1><<<ceil_div(n_rows, DEEP_THREADS), DEEP_THREADS, 0, current_stream()>>>(
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
    quotients_combine_kernel<COMBINE_ROWS><<<ceil_div(n_rows, rows_per_block), DEEP_THREADS, 0, current_stream()>>>(
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

vx_deep* vx_deep_create(uint32_t max_cols, uint32_t max_points) {
    if (!max_cols || max_cols > (1u << 20) || !max_points || max_points > DEEP_MAX_POINTS) {
        set_error(-2, "vx_deep_create: need 1..2^20 columns and 1..4 sample points");
        return nullptr;
    }
    vx_deep* d = new vx_deep();
    d->max_cols = max_cols;
    d->max_points = max_points;
    d->bytes = 4ull * (16ull * max_points + 4ull * max_cols + 4ull * max_cols * max_points);
    bool ok = true;
    for (int k = 0; ok && k < 2; k++)
        ok = check(hipHostMalloc((void**)&d->h_meta[k], d->bytes), "hipHostMalloc") &&
             check(hipMalloc((void**)&d->d_meta[k], d->bytes), "hipMalloc") &&
             check(hipEventCreateWithFlags(&d->copied[k], hipEventDisableTiming), "hipEventCreate") &&
             check(hipEventCreateWithFlags(&d->done[k], hipEventDisableTiming), "hipEventCreate");
    if (!ok) {
        vx_deep_destroy(d);
        return nullptr;
    }
    return d;
}

void vx_deep_destroy(vx_deep* d) {
    if (!d) return;
    for (int k = 0; k < 2; k++) {
        if (d->pending[k]) (void)hipEventSynchronize(d->done[k]);
        if (d->h_meta[k]) (void)hipHostFree(d->h_meta[k]);
        if (d->d_meta[k]) (void)hipFree(d->d_meta[k]);
        if (d->copied[k]) (void)hipEventDestroy(d->copied[k]);
        if (d->done[k]) (void)hipEventDestroy(d->done[k]);
    }
    delete d;
}

int32_t vx_deep_quotient(vx_deep* d, const uint32_t* const* cols, const uint32_t* col_log_sizes, uint32_t n_cols,
                         uint32_t n_points, const uint32_t* points, const uint32_t* line_coeffs,
                         uint32_t lifting_log_size, const uint32_t* domain_xs, const uint32_t* domain_ys,
                         uint32_t* out0, uint32_t* out1, uint32_t* out2, uint32_t* out3, void* stream) {
    if (!d) {
        set_error(hipErrorInvalidValue, "vx_deep_quotient: null argument");
        return -1;
    }
    if (!deep_args_ok("vx_deep_quotient", cols, col_log_sizes, n_cols, d->max_cols, n_points, d->max_points, points,
                      line_coeffs, lifting_log_size, domain_xs, domain_ys, out0, out1, out2, out3))
        return -1;
    hipStream_t s = (hipStream_t)stream;
    // this slot may still feed the copy of the call two back, and its device copy that call's kernel
    const uint32_t slot = d->next;
    if (d->pending[slot] && !(check(hipEventSynchronize(d->copied[slot]), "hipEventSynchronize") &&
                              check(hipStreamWaitEvent(s, d->done[slot], 0), "hipStreamWaitEvent")))
        return -1;
    uint32_t* h = d->h_meta[slot];
    const uint32_t NP = n_points;
    deep_meta(h, cols, col_log_sizes, n_cols, n_points, points, line_coeffs, lifting_log_size);
    const size_t bytes = 4ull * deep_meta_words(n_cols, NP);
    if (!check(hipMemcpyAsync(d->d_meta[slot], h, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync") ||
        !check(hipEventRecord(d->copied[slot], s), "hipEventRecord"))
        return -1;
    if (!launch_deep_fused(d->d_meta[slot], NP, n_cols, lifting_log_size, domain_xs, domain_ys, out0, out1, out2,
                           out3, s))
        return -1;
    if (!check(hipEventRecord(d->done[slot], s), "hipEventRecord")) return -1;
    d->pending[slot] = true;
    d->next ^= 1;
    return 0;
}

int32_t vx_deep_quotient_noalloc(const uint32_t* const* cols, const uint32_t* col_log_sizes, uint32_t n_cols,
                                 uint32_t n_points, const uint32_t* points, const uint32_t* line_coeffs,
                                 uint32_t lifting_log_size, const uint32_t* domain_xs, const uint32_t* domain_ys,
                                 uint32_t* out0, uint32_t* out1, uint32_t* out2, uint32_t* out3, void* stream) {
    if (!deep_args_ok("vx_deep_quotient_noalloc", cols, col_log_sizes, n_cols, 1u << 20, n_points, DEEP_MAX_POINTS,
                      points, line_coeffs, lifting_log_size, domain_xs, domain_ys, out0, out1, out2, out3))
        return -1;
    const uint64_t words = deep_meta_words(n_cols, n_points);
    if (words > DEEP_SCRATCH_WORDS) {
        set_error(-2, "vx_deep_quotient_noalloc: metadata of this many columns and points exceeds the 8192-word "
                      "device scratch (16 n_points + 4 n_cols + 4 n_cols n_points words)");
        return -1;
    }
    uint32_t* dst = deep_scratch();
    if (!dst) return -1;
    hipStream_t s = (hipStream_t)stream;
    // metadata travels as kernel arguments into the scratch array, chunk by chunk
    MetaChunk c;
    uint32_t h[DEEP_SCRATCH_WORDS];
    deep_meta(h, cols, col_log_sizes, n_cols, n_points, points, line_coeffs, lifting_log_size);
    for (uint64_t off = 0; off < words; off += META_CHUNK_WORDS) {
        const uint32_t n = (uint32_t)std::min<uint64_t>(META_CHUNK_WORDS, words - off);
        std::memcpy(c.w, h + off, 4ull * n);
        stage_meta_kernel<<<1, 256, 0, s>>>(c, dst + off, n);
        if (!check_launch("stage_meta_kernel")) return -1;
    }
    return launch_deep_fused(dst, n_points, n_cols, lifting_log_size, domain_xs, domain_ys, out0, out1, out2, out3, s)
               ? 0
               : -1;
}

int32_t vx_circle_domain_points(uint32_t log_size, uint32_t* xs, uint32_t* ys, void* stream) {
    if (!xs || !ys || log_size < 1 || log_size > 30) {
        set_error(-2, "vx_circle_domain_points: log size outside 1..30 or null output");
        return -1;
    }
    HostPt hi, hs;
    host_coset(2, log_size - 1, hi, hs);  // half_odds(log_size - 1)
    domain_points_kernel<<<ceil_div(1ull << log_size, DEEP_THREADS), DEEP_THREADS, 0, (hipStream_t)stream>>>(
        CPt{hi.x, hi.y}, CPt{hs.x, hs.y}, log_size, xs, ys);
    return check_launch("domain_points_kernel") ? 0 : -1;
}

}  // extern "C"
