// M31 / CM31 / QM31 arithmetic for gfx950 device code.
//
// Semantics follow the reference field tower exactly (canonical u32 in [0, p),
// p = 2^31 - 1): src/field/m31.rs:24-121, cm31.rs:78-87, qm31.rs:114-132 and
// cuda/include/{m31,cm31,qm31}.cuh. The instruction choices are CDNA-specific:
// the canonical reduction of r in [0, 2p) is min(r, r - p) (v_sub + v_min_u32,
// no compare/select), and the 62-bit product is folded with one v_alignbit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vx {

constexpr uint32_t P = 0x7FFFFFFFu;

// r in [0, 2p) -> r mod p.
__device__ __forceinline__ uint32_t canon(uint32_t r) { return min(r, r - P); }

__device__ __forceinline__ uint32_t madd(uint32_t a, uint32_t b) { return canon(a + b); }
__device__ __forceinline__ uint32_t msub(uint32_t a, uint32_t b) { return canon(a + (P - b)); }
__device__ __forceinline__ uint32_t mneg(uint32_t a) { return a == 0 ? 0 : P - a; }
__device__ __forceinline__ uint32_t mdbl(uint32_t a) { return canon(a << 1); }

__device__ __forceinline__ uint32_t mmul(uint32_t a, uint32_t b) {
    uint64_t p = (uint64_t)a * b;
    uint32_t lo = (uint32_t)p, hi = (uint32_t)(p >> 32);
    // p mod (2^31 - 1) = (p & p31) + (p >> 31); the sum is < 2p for canonical inputs.
    uint32_t top = __builtin_amdgcn_alignbit(hi, lo, 31);
    return canon((lo & P) + top);
}

__device__ __forceinline__ uint32_t mpow(uint32_t b, uint32_t e) {
    uint32_t r = 1;
    while (e) {
        if (e & 1) r = mmul(r, b);
        b = mmul(b, b);
        e >>= 1;
    }
    return r;
}
__device__ __forceinline__ uint32_t msqr_n(uint32_t a, int n) {
    for (int i = 0; i < n; i++) a = mmul(a, a);
    return a;
}
// a^(p-2) = a^(2^31 - 3) by an addition chain: x_k = a^(2^k - 1) from
// x_1, x_2, x_4, x_8, x_16, x_24, x_28, x_29, then x_29^4 * a
// (30 squarings and 8 products, against 61 for square-and-multiply). 0 -> 0.
__device__ __forceinline__ uint32_t minv(uint32_t a) {
    const uint32_t x2 = mmul(mmul(a, a), a);
    const uint32_t x4 = mmul(msqr_n(x2, 2), x2);
    const uint32_t x8 = mmul(msqr_n(x4, 4), x4);
    const uint32_t x16 = mmul(msqr_n(x8, 8), x8);
    const uint32_t x24 = mmul(msqr_n(x16, 8), x8);
    const uint32_t x28 = mmul(msqr_n(x24, 4), x4);
    const uint32_t x29 = mmul(mmul(x28, x28), a);
    return mmul(msqr_n(x29, 2), a);
}

struct CM31 {
    uint32_t a, b;
};
__device__ __forceinline__ CM31 cadd(CM31 x, CM31 y) { return {madd(x.a, y.a), madd(x.b, y.b)}; }
__device__ __forceinline__ CM31 csub(CM31 x, CM31 y) { return {msub(x.a, y.a), msub(x.b, y.b)}; }
// (a + bi)(c + di) with three products: the imaginary part is
// (a + b)(c + d) - ac - bd (Karatsuba; equal to ad + bc mod p).
__device__ __forceinline__ CM31 cmul(CM31 x, CM31 y) {
    const uint32_t ac = mmul(x.a, y.a), bd = mmul(x.b, y.b);
    const uint32_t s = mmul(madd(x.a, x.b), madd(y.a, y.b));
    return {msub(ac, bd), msub(msub(s, ac), bd)};
}

// QM31 = a + b*u with u^2 = 2 + i; stored as {a.a, a.b, b.a, b.b}.
struct QM31 {
    uint32_t v[4];
};
__device__ __forceinline__ QM31 qadd(QM31 x, QM31 y) {
    return {{madd(x.v[0], y.v[0]), madd(x.v[1], y.v[1]), madd(x.v[2], y.v[2]), madd(x.v[3], y.v[3])}};
}
__device__ __forceinline__ QM31 qsub(QM31 x, QM31 y) {
    return {{msub(x.v[0], y.v[0]), msub(x.v[1], y.v[1]), msub(x.v[2], y.v[2]), msub(x.v[3], y.v[3])}};
}
__device__ __forceinline__ QM31 qmul_m31(QM31 x, uint32_t s) {
    return {{mmul(x.v[0], s), mmul(x.v[1], s), mmul(x.v[2], s), mmul(x.v[3], s)}};
}
// (A + Bu)(C + Du) = (AC + (2 + i) BD) + ((A + B)(C + D) - AC - BD) u: three
// CM31 products (nine M31 products) instead of four (sixteen).
__device__ __forceinline__ QM31 qmul(QM31 x, QM31 y) {
    const CM31 xa = {x.v[0], x.v[1]}, xb = {x.v[2], x.v[3]};
    const CM31 ya = {y.v[0], y.v[1]}, yb = {y.v[2], y.v[3]};
    const CM31 ac = cmul(xa, ya), bd = cmul(xb, yb);
    const CM31 s = cmul(cadd(xa, xb), cadd(ya, yb));
    // bd * (2 + i) = (2p - q) + (p + 2q) i
    const CM31 bdu = {msub(mdbl(bd.a), bd.b), madd(bd.a, mdbl(bd.b))};
    const CM31 ra = cadd(ac, bdu), rb = csub(csub(s, ac), bd);
    return {{ra.a, ra.b, rb.a, rb.b}};
}

// FRI fold of one pair: (f0 + f1) + alpha * ((f0 - f1) * tw)   (cuda/fri.cu:26-33)
__device__ __forceinline__ QM31 fold_pair(QM31 f0, QM31 f1, QM31 alpha, uint32_t tw) {
    return qadd(qadd(f0, f1), qmul(alpha, qmul_m31(qsub(f0, f1), tw)));
}

// Circle group (src/circle.rs:33-39)
struct CPt {
    uint32_t x, y;
};
__device__ __forceinline__ CPt cpmul(CPt p, CPt q) {
    return {msub(mmul(p.x, q.x), mmul(p.y, q.y)), madd(mmul(p.x, q.y), mmul(p.y, q.x))};
}
// initial * step^e by square-and-multiply.
__device__ __forceinline__ CPt cp_pow_mul(CPt initial, CPt step, uint32_t e) {
    CPt r = {1, 0};
    while (e) {
        if (e & 1) r = cpmul(r, step);
        step = cpmul(step, step);
        e >>= 1;
    }
    return cpmul(initial, r);
}

__device__ __forceinline__ uint32_t brev_bits(uint32_t x, uint32_t bits) {
    return bits == 0 ? 0u : (__builtin_bitreverse32(x) >> (32 - bits));
}

}  // namespace vx
