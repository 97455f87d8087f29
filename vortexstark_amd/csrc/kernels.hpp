// Internal (C++) interface between the kernel files and the prover.
#pragma once
#include <functional>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "field.hpp"

namespace vx {

// ---------------------------------------------------------------- NTT
// Twiddle tables of the native circle NTT for one domain of size 2^log_n
// (layout of src/circle.rs:357-451: line layers concatenated, layer l at
// offs[l]; circle twiddles separate).
struct NttTables {
    const uint32_t* line = nullptr;
    const uint32_t* circle = nullptr;
    uint32_t offs[32] = {0};
    uint32_t log_n = 0;
    bool stwo = false;  // line is a flat stwo twiddle buffer; circle twiddles derived from it
};
// Tables of the stwo-basis transform of size 2^log_n over a flat twiddle buffer
// of 2^(log_n-1) words (src/ntt.rs:262-324): layer b at n/2 - n/2^b.
NttTables stwo_tables(const uint32_t* flat, uint32_t log_n);

// Forward transform (coefficients -> values). Reads src (src_n valid words,
// the rest taken as zero: the implicit zero pad of cuda/constraint.cu:81-90)
// and writes dst (2^log_n words). src may equal dst when src_n == 2^log_n.
void ntt_forward(uint32_t* dst, const uint32_t* src, uint64_t src_n, const NttTables& t, hipStream_t s);
// Inverse transform including the 2^-log_n scale. src may equal dst.
void ntt_inverse(uint32_t* dst, const uint32_t* src, const NttTables& t, hipStream_t s);
// Batched forms over a device array of column pointers (one launch per pass for
// all columns). Forward works in place and reads src_n words of each column (the
// rest taken as zero; 0 = all). Inverse reads d_srcs when given, else in place.
// aligned: every column pointer is 16-byte aligned (tiles then move as uint4).
void ntt_forward_batch(uint32_t* const* d_cols, uint32_t n_cols, const NttTables& t, hipStream_t s,
                       uint64_t src_n = 0, bool aligned = false);
void ntt_inverse_batch(uint32_t* const* d_cols, uint32_t n_cols, const NttTables& t, hipStream_t s,
                       const uint32_t* const* d_srcs = nullptr, bool aligned = false);
// Blowup-4 LDE of n_cols columns: interpolation of 2^ti.log_n points (from
// srcs, or in place) into the first quarter of each column, then evaluation of
// 2^te.log_n = 4 * 2^ti.log_n points in place (top passes fused when the plans
// allow).
// mark("interpolate") / mark("evaluate") follow the two parts when given.
void lde_batch(uint32_t* const* cols, uint32_t n_cols, const NttTables& ti, const NttTables& te, hipStream_t s,
               const uint32_t* const* srcs, bool src_aligned,
               const std::function<void(const char*)>& mark = nullptr);

// Builds the twiddles of coset (initial, step) of size 2^log_n directly on the
// device: line (2^log_n - 1 words), circle (2^(log_n-1) words) and, when the
// i* pointers are non-null, their inverses. offs receives the layer offsets.
void build_ntt_tables(CPt initial, CPt step, uint32_t log_n, uint32_t* line, uint32_t* circle, uint32_t* iline,
                      uint32_t* icircle, uint32_t offs[32], hipStream_t s);
// FRI fold twiddles of a domain: out[i] = 1 / coord(initial * step^bitrev(2i)), i < 2^(log_n-1).
void build_fold_twiddles(CPt initial, CPt step, uint32_t log_n, bool use_y, uint32_t* out, uint32_t* scratch,
                         hipStream_t s);
void batch_inverse(const uint32_t* in, uint32_t* out, uint64_t n, hipStream_t s);

// Host-side circle arithmetic used to derive coset parameters.
struct HostPt {
    uint32_t x, y;
};
HostPt host_pmul(HostPt a, HostPt b);
HostPt host_repeated_double(HostPt p, uint32_t n);
HostPt host_generator();
// kind: 0 subgroup, 1 half_coset, 2 half_odds (src/circle.rs:98-143)
void host_coset(int kind, uint32_t log_size, HostPt& initial, HostPt& step);

}  // namespace vx
