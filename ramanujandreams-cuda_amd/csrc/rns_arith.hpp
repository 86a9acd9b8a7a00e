// rns_arith.hpp -- modular arithmetic for the RNS walk on gfx950.
//
// Residues mod m < 2^30 (odd) are kept in Montgomery form aR mod m with
// R = 2^32.  The product of two residues is one 32x32->64 multiply-add
// (v_mad_u64_u32) followed by a Montgomery reduction of the 64-bit sum;
// because every operand is < m < 2^30, up to three products and an
// addend a*2^32 can be accumulated in 64 bits before one reduction.  All
// results are the exact residues the reference's `(a*b) % m` arithmetic
// produces (cuda/cmf_walk_rns.cu:45-66), in another representation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cmfw {

struct MontMod {
    uint32_t m;      // modulus (odd, < 2^30)
    uint32_t mp;     // -m^-1 mod 2^32
    uint32_t r1;     // R mod m      (Montgomery 1)
    uint32_t r2;     // R^2 mod m    (to Montgomery form)
    double inv;      // 1/m rounded
};

// Montgomery reduction of T < 2^63: T * R^-1 mod m, in [0, m)
__device__ __forceinline__ uint32_t redc(uint64_t T, const MontMod &M) {
    const uint32_t q = (uint32_t)T * M.mp;
    const uint64_t U = T + (uint64_t)q * M.m;  // < 2^63 + 2^62, divisible by 2^32
    uint32_t r = (uint32_t)(U >> 32);          // < 2^31 + 2^30
    r = (r >= M.m) ? r - M.m : r;
    r = (r >= M.m) ? r - M.m : r;
    return r;
}
// Product of two reduced residues a, b < m: T < m^2, so the reduction's
// (T + q m) / 2^32 < m^2 / 2^32 + m < 1.25 m (m < 2^30) and one conditional
// subtraction leaves it in [0, m).
__device__ __forceinline__ uint32_t mmul(uint32_t a, uint32_t b, const MontMod &M) {
    const uint64_t T = (uint64_t)a * b;
    const uint32_t q = (uint32_t)T * M.mp;
    const uint32_t r = (uint32_t)((T + (uint64_t)q * M.m) >> 32);
    return (r >= M.m) ? r - M.m : r;
}
// a*b + c (all Montgomery): c enters as c*R so one reduction suffices
__device__ __forceinline__ uint32_t mfma(uint32_t a, uint32_t b, uint32_t c, const MontMod &M) {
    return redc((uint64_t)a * b + ((uint64_t)c << 32), M);
}
__device__ __forceinline__ uint32_t madd(uint32_t a, uint32_t b, uint32_t m) {
    const uint32_t r = a + b;
    return r >= m ? r - m : r;
}
__device__ __forceinline__ uint32_t msub(uint32_t a, uint32_t b, uint32_t m) {
    return a >= b ? a - b : a + m - b;
}
__device__ __forceinline__ uint32_t to_mont(uint32_t a, const MontMod &M) { return mmul(a, M.r2, M); }
__device__ __forceinline__ uint32_t from_mont(uint32_t a, const MontMod &M) { return redc(a, M); }

// x^(m-2) mod m for x in Montgomery form (the reference's mod_inv,
// cuda/cmf_walk_rns.cu:68-78: an inverse for prime m, a fixed power
// otherwise; powers are multiplicative either way).
__device__ __forceinline__ uint32_t mpow_m2(uint32_t x, const MontMod &M) {
    uint32_t e = M.m - 2u;
    uint32_t r = M.r1, b = x;
    while (e) {
        const uint32_t rb = mmul(r, b, M);
        r = (e & 1u) ? rb : r;
        b = mmul(b, b, M);
        e >>= 1;
    }
    return r;
}

// llround with the restatement's out-of-range policy (NaN or |x| >= 2^63
// -> INT64_MIN), then the mathematical residue in [0, m): to_mod,
// cuda/cmf_walk_rns.cu:61-66 (its two branches compute v mod m for every
// int64 v, INT64_MIN included).  Doubles below 2^52 reduce in floating
// point: q may be off by one, r = x - q m is exact.
__device__ __forceinline__ uint32_t to_mod_fast(double x, const MontMod &M) {
    if (!(x >= -9223372036854775808.0 && x < 9223372036854775808.0)) {
        const uint64_t neg = 0x8000000000000000ull;  // -(INT64_MIN) as unsigned
        return (uint32_t)((M.m - (uint32_t)(neg % M.m)) % M.m);
    }
    const double v = round(x);
    if (fabs(v) < 4503599627370496.0) {
        const double m = (double)M.m;
        const double q = floor(v * M.inv);
        double r = fma(-q, m, v);
        r = (r < 0.0) ? r + m : r;
        r = (r >= m) ? r - m : r;
        return (uint32_t)r;
    }
    const int64_t iv = (int64_t)v;
    if (iv >= 0) return (uint32_t)((uint64_t)iv % M.m);
    const uint64_t neg = (uint64_t)0 - (uint64_t)iv;
    return (uint32_t)((M.m - (uint32_t)(neg % M.m)) % M.m);
}

}  // namespace cmfw
