// Device-side modular arithmetic for moduli p < 2^62 (exact, integer only).
//
// Conventions follow math/src/zq/mod.rs of fhe.rs: Shoup companions
// w' = floor(w * 2^64 / p); lazy results live in [0, 2p) or [0, 4p) and are
// fully reduced at operation boundaries, so every value this library returns is
// the unique residue in [0, p).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fhehip {

typedef uint64_t u64;
typedef uint32_t u32;
typedef unsigned __int128 u128;

// x mod p for x < 2p: the borrow of x - p selects the result.
__device__ __forceinline__ u64 reduce1(u64 x, u64 p) {
    unsigned long long borrow;
    const u64 d = __builtin_subcll(x, p, 0, &borrow);
    return borrow ? x : d;
}

// Full 64 x 64 -> 128-bit product from four 32 x 32 -> 64 partial products.
__device__ __forceinline__ void mul_wide_fwd(u64 a, u64 b, u64& lo, u64& hi) {
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    const u64 p00 = (u64)a0 * b0;
    const u64 p01 = (u64)a0 * b1;
    const u64 p10 = (u64)a1 * b0;
    const u64 p11 = (u64)a1 * b1;
    const u64 mid = (p00 >> 32) + (u32)p01 + (u32)p10;
    lo = (mid << 32) | (u32)p00;
    hi = p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

// lazy_mul_shoup (math/src/zq/mod.rs:224-234): a*w mod p in [0, 2p) for any a < 2^64.
__device__ __forceinline__ u64 lazy_mul_shoup(u64 a, u64 w, u64 ws, u64 p) {
    u64 q = __umul64hi(a, ws);
    return a * w - q * p;
}

__device__ __forceinline__ u64 mul_shoup(u64 a, u64 w, u64 ws, u64 p) {
    return reduce1(lazy_mul_shoup(a, w, ws, p), p);
}

// Barrett reduction of a 128-bit value with bh:bl = floor(2^128 / p)
// (lazy_reduce_u128, math/src/zq/mod.rs:693-707); result in [0, 2p). As in the
// reference, the input must keep a_hi * bl + a_lo * bh below 2^128 (true for
// a < 2^127 and p >= 2).
__device__ __forceinline__ u64 lazy_reduce_u128(u64 a_lo, u64 a_hi, u64 bl, u64 bh, u64 p) {
    const u64 lolo_hi = __umul64hi(a_lo, bl);
    u64 hl_lo, hl_hi, lh_lo, lh_hi;
    mul_wide_fwd(a_hi, bl, hl_lo, hl_hi);
    mul_wide_fwd(a_lo, bh, lh_lo, lh_hi);
    u64 m = hl_lo + lh_lo;
    u64 carry = m < hl_lo ? 1 : 0;
    const u64 m2 = m + lolo_hi;
    carry += m2 < m ? 1 : 0;
    const u64 q = hl_hi + lh_hi + carry + a_hi * bh;
    return a_lo - q * p;
}

__device__ __forceinline__ u64 reduce_u128(u64 a_lo, u64 a_hi, u64 bl, u64 bh, u64 p) {
    return reduce1(lazy_reduce_u128(a_lo, a_hi, bl, bh, p), p);
}

// Exact a*b mod p for a, b < p.
__device__ __forceinline__ u64 mul_mod(u64 a, u64 b, u64 bl, u64 bh, u64 p) {
    return reduce_u128(a * b, __umul64hi(a, b), bl, bh, p);
}

// Lazy reduction of a u64 to [0, 2p) (math/src/zq/mod.rs:712-722).
__device__ __forceinline__ u64 lazy_reduce_u64(u64 a, u64 bl, u64 bh, u64 p) {
    const u64 lolo_hi = __umul64hi(a, bl);
    u64 lh_lo, lh_hi;
    mul_wide_fwd(a, bh, lh_lo, lh_hi);
    const u64 m = lh_lo + lolo_hi;
    const u64 q = lh_hi + (m < lh_lo ? 1 : 0);
    return a - q * p;
}

__device__ __forceinline__ u64 add_mod(u64 a, u64 b, u64 p) { return reduce1(a + b, p); }

__device__ __forceinline__ void mul_wide(u64 a, u64 b, u64& lo, u64& hi) { mul_wide_fwd(a, b, lo, hi); }

// (lo, hi) += (alo, ahi)
__device__ __forceinline__ void add128(u64& lo, u64& hi, u64 alo, u64 ahi) {
    unsigned long long c;
    lo = __builtin_addcll(lo, alo, 0, &c);
    hi = hi + ahi + c;
}

// (lo, hi, top) += (alo, ahi): 192-bit accumulation of 128-bit terms.
__device__ __forceinline__ void add192(u64& lo, u64& hi, u64& top, u64 alo, u64 ahi) {
    unsigned long long c0, c1;
    lo = __builtin_addcll(lo, alo, 0, &c0);
    hi = __builtin_addcll(hi, ahi, c0, &c1);
    top += c1;
}

// Montgomery reduction with R = 2^64 of t = hi * 2^64 + lo < p * 2^63:
// returns t * R^-1 mod p in [0, 2p). pinv_neg = -p^-1 mod 2^64.
__device__ __forceinline__ u64 mont_redc(u64 lo, u64 hi, u64 p, u64 pinv_neg) {
    const u64 m = lo * pinv_neg;
    return hi + __umul64hi(m, p) + (lo != 0 ? 1 : 0);
}
__device__ __forceinline__ u64 sub_mod(u64 a, u64 b, u64 p) { return reduce1(a + p - b, p); }

// Per-modulus constants kept resident on the device.
struct ModInfo {
    u64 p;
    u64 barrett_lo;  // floor(2^128 / p), low word
    u64 barrett_hi;  // floor(2^128 / p), high word
    u64 n_inv;       // N^-1 mod p
    u64 n_inv_shoup;
    u64 pinv_neg;    // -p^-1 mod 2^64 (Montgomery)
    u64 n_inv_r;     // N^-1 * 2^64 mod p: undoes a Montgomery factor inside an inverse NTT
    u64 n_inv_r_shoup;
    // zeta of the inverse transform's last stage times N^-1 (and times
    // N^-1 2^64): the final scaling folded into that stage.
    u64 last_zn, last_zn_shoup;
    u64 last_zr, last_zr_shoup;
};

}  // namespace fhehip
