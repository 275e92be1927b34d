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
//
// q = floor(a * ws / 2^64) and r = a*w - q*p (mod 2^64) on 32-bit halves
// (n = 2^64 - p), with each partial product formed once:
//   D = a0*s1 + hi(a0*s0)            (< 2^64)
//   E = a1*s0 + D, carry k           (65 bits)
//   q = a1*s1 + (E >> 32) + k*2^32
//   r = a0*w0 + q0*n0 + 2^32 * (a0*w1 + a1*w0 + q0*n1 + q1*n0)
// The carry and the high words are consumed as halves of 64-bit operands,
// so the temporaries are fixed registers (v0..v5, result in v[2:3]). A VALU
// write of vcc or of an SGPR pair is read by a VALU instruction only two or
// more instructions (or nop wait states) later. Every instruction reads at
// most one SGPR operand, so values that are uniform across the wave (the
// modulus, and the twiddle where it is shared) can stay in SGPRs.
#define FHEHIP_SHOUP_ASM                                 \
    "v_mov_b32 v1, 0\n\t"                                \
    "v_mul_hi_u32 v0, %[a0], %[s0]\n\t"                  \
    "v_mad_u64_u32 v[2:3], %[k], %[a0], %[s1], v[0:1]\n\t" \
    "v_mad_u64_u32 v[2:3], %[k], %[a1], %[s0], v[2:3]\n\t" \
    "v_mov_b32 v0, v3\n\t"                               \
    "s_nop 1\n\t"                                        \
    "v_addc_co_u32 v1, %[k], 0, 0, %[k]\n\t"             \
    "v_mad_u64_u32 v[0:1], %[k], %[a1], %[s1], v[0:1]\n\t" \
    "v_mad_u64_u32 v[2:3], %[k], %[a0], %[w0], 0\n\t"    \
    "v_mad_u64_u32 v[2:3], %[k], v0, %[n0], v[2:3]\n\t"  \
    "v_mul_lo_u32 v4, %[a0], %[w1]\n\t"                  \
    "v_mul_lo_u32 v5, %[a1], %[w0]\n\t"                  \
    "v_add3_u32 v3, v3, v4, v5\n\t"                      \
    "v_mul_lo_u32 v4, v0, %[n1]\n\t"                     \
    "v_mul_lo_u32 v5, v1, %[n0]\n\t"                     \
    "v_add3_u32 v3, v3, v4, v5"

// Twiddle-class operands (w, ws, n) in VGPRs ("v") or SGPRs ("s").
#define FHEHIP_SHOUP_IN(C)                                                                           \
    [a0] "v"((u32)a), [a1] "v"((u32)(a >> 32)), [s0] C((u32)ws), [s1] C((u32)(ws >> 32)), [w0] C((u32)w), \
        [w1] C((u32)(w >> 32)), [n0] C((u32)n), [n1] C((u32)(n >> 32))

__device__ __forceinline__ u64 lazy_mul_shoup(u64 a, u64 w, u64 ws, u64 p) {
    const u64 n = 0 - p;
    u64 r, k;
    asm(FHEHIP_SHOUP_ASM : "=&{v[2:3]}"(r), [k] "=&s"(k) : FHEHIP_SHOUP_IN("v") : "v0", "v1", "v4", "v5");
    return r;
}

// lazy_mul_shoup for w, ws and p uniform across the wave (kept in SGPRs).
__device__ __forceinline__ u64 lazy_mul_shoup_u(u64 a, u64 w, u64 ws, u64 p) {
    const u64 n = 0 - p;
    u64 r, k;
    asm(FHEHIP_SHOUP_ASM : "=&{v[2:3]}"(r), [k] "=&s"(k) : FHEHIP_SHOUP_IN("s") : "v0", "v1", "v4", "v5");
    return r;
}

// Harvey forward butterfly (ntt/native.rs:256-269) on x, y < 4p:
// xr = x - 2p if x >= 2p, t = lazy_mul_shoup(y), x' = xr + t, y' = xr + 2p - t.
// Fixed temporaries v0..v7 as in lazy_mul_shoup (t ends in v[2:3]); both
// outputs are written by 64-bit adds straight into x and y after the last
// input read.
// p is uniform across the wave; UT: the twiddle (w, ws) is too.
#define FHEHIP_FWD_BFLY_ASM                                \
    "v_mov_b32 v1, 0\n\t"                                  \
    "v_subrev_co_u32 v4, vcc, %[P0], %[x0]\n\t"            \
    "v_mul_hi_u32 v0, %[a0], %[s0]\n\t"                    \
    "v_mad_u64_u32 v[2:3], %[k], %[a0], %[s1], v[0:1]\n\t" \
    "v_subb_co_u32 v5, vcc, %[x1], %[P1], vcc\n\t"         \
    "v_mad_u64_u32 v[2:3], %[k], %[a1], %[s0], v[2:3]\n\t" \
    "v_mov_b32 v0, v3\n\t"                                 \
    "v_cndmask_b32 v4, v4, %[x0], vcc\n\t"                 \
    "v_addc_co_u32 v1, %[k], 0, 0, %[k]\n\t"               \
    "v_cndmask_b32 v5, v5, %[x1], vcc\n\t"                 \
    "v_mad_u64_u32 v[0:1], %[k], %[a1], %[s1], v[0:1]\n\t" \
    "v_mad_u64_u32 v[2:3], %[k], %[a0], %[w0], 0\n\t"      \
    "v_mul_lo_u32 v6, %[a0], %[w1]\n\t"                    \
    "v_mul_lo_u32 v7, %[a1], %[w0]\n\t"                    \
    "v_mad_u64_u32 v[2:3], %[k], v0, %[n0], v[2:3]\n\t"    \
    "v_mul_lo_u32 v0, v0, %[n1]\n\t"                       \
    "v_mul_lo_u32 v1, v1, %[n0]\n\t"                       \
    "v_add3_u32 v7, v7, v6, v0\n\t"                        \
    "v_sub_co_u32 v6, vcc, %[P0], v2\n\t"                  \
    "v_add3_u32 v3, v3, v7, v1\n\t"                        \
    "s_nop 0\n\t"                                          \
    "v_subb_co_u32 v7, vcc, %[P1], v3, vcc\n\t"            \
    "v_lshl_add_u64 %[xo], v[4:5], 0, v[2:3]\n\t"          \
    "v_lshl_add_u64 %[yo], v[4:5], 0, v[6:7]"

#define FHEHIP_FWD_BFLY_IN(C)                                                                                  \
    [x0] "v"((u32)x), [x1] "v"((u32)(x >> 32)), [a0] "v"((u32)y), [a1] "v"((u32)(y >> 32)), [s0] C((u32)ws), \
        [s1] C((u32)(ws >> 32)), [w0] C((u32)w), [w1] C((u32)(w >> 32)), [n0] "s"((u32)n),                  \
        [n1] "s"((u32)(n >> 32)), [P0] "s"((u32)p2), [P1] "v"((u32)(p2 >> 32))

template <bool UT = false>
__device__ __forceinline__ void fwd_bfly_asm(u64& x, u64& y, u64 w, u64 ws, u64 p, u64 p2) {
    const u64 n = 0 - p;
    u64 k;
    if constexpr (UT)
        asm(FHEHIP_FWD_BFLY_ASM : [xo] "=v"(x), [yo] "=v"(y), [k] "=&s"(k) : FHEHIP_FWD_BFLY_IN("s")
            : "v0", "v1", "v2", "v3", "v4", "v5", "v6", "v7", "vcc");
    else
        asm(FHEHIP_FWD_BFLY_ASM : [xo] "=v"(x), [yo] "=v"(y), [k] "=&s"(k) : FHEHIP_FWD_BFLY_IN("v")
            : "v0", "v1", "v2", "v3", "v4", "v5", "v6", "v7", "vcc");
}

// Gentleman-Sande inverse butterfly (ntt/native.rs:288-300) on x, y < 2p:
// x' = (x + y) mod 2p, y' = lazy_mul_shoup(x + 2p - y). The reduction adds
// (u < 2p ? 0 : -2p) to u = x + y, with -2p masked by a lane mask so that it
// can stay in SGPRs.
#define FHEHIP_INV_BFLY_ASM                                \
    "v_sub_co_u32 v6, %[k], %[x0], %[y0]\n\t"              \
    "v_mov_b32 v1, 0\n\t"                                  \
    "v_lshl_add_u64 v[4:5], %[x], 0, %[y]\n\t"             \
    "v_subb_co_u32 v7, %[k], %[x1], %[y1], %[k]\n\t"       \
    "v_cmp_gt_u64 vcc, %[P], v[4:5]\n\t"                   \
    "v_lshl_add_u64 v[6:7], v[6:7], 0, %[P]\n\t"           \
    "v_mul_hi_u32 v0, v6, %[s0]\n\t"                       \
    "v_mad_u64_u32 v[2:3], %[k], v6, %[s1], v[0:1]\n\t"    \
    "v_mad_u64_u32 v[2:3], %[k], v7, %[s0], v[2:3]\n\t"    \
    "v_mov_b32 v0, v3\n\t"                                 \
    "s_nop 0\n\t"                                          \
    "v_addc_co_u32 v1, %[k], 0, 0, %[k]\n\t"               \
    "v_mad_u64_u32 v[0:1], %[k], v7, %[s1], v[0:1]\n\t"    \
    "v_mad_u64_u32 v[2:3], %[k], v6, %[w0], 0\n\t"         \
    "v_mul_lo_u32 v6, v6, %[w1]\n\t"                       \
    "v_mul_lo_u32 v7, v7, %[w0]\n\t"                       \
    "v_mad_u64_u32 v[2:3], %[k], v0, %[n0], v[2:3]\n\t"    \
    "v_mul_lo_u32 v0, v0, %[n1]\n\t"                       \
    "v_mul_lo_u32 v1, v1, %[n0]\n\t"                       \
    "v_add3_u32 v7, v7, v6, v0\n\t"                        \
    "v_add3_u32 v3, v3, v7, v1\n\t"                        \
    "v_cndmask_b32 v1, -1, 0, vcc\n\t"                     \
    "v_and_b32 v0, %[m0], v1\n\t"                          \
    "v_and_b32 v1, %[m1], v1\n\t"                          \
    "v_mov_b64 %[yo], v[2:3]\n\t"                          \
    "v_lshl_add_u64 %[xo], v[4:5], 0, v[0:1]"

#define FHEHIP_INV_BFLY_IN(C)                                                                                  \
    [x0] "v"((u32)x), [x1] "v"((u32)(x >> 32)), [y0] "v"((u32)y), [y1] "v"((u32)(y >> 32)), [x] "v"(x),       \
        [y] "v"(y), [s0] C((u32)zs), [s1] C((u32)(zs >> 32)), [w0] C((u32)z), [w1] C((u32)(z >> 32)),        \
        [n0] "s"((u32)n), [n1] "s"((u32)(n >> 32)), [P] "s"(p2), [m0] "s"((u32)m2), [m1] "s"((u32)(m2 >> 32))

template <bool UT = false>
__device__ __forceinline__ void inv_bfly_asm(u64& x, u64& y, u64 z, u64 zs, u64 p, u64 p2) {
    const u64 n = 0 - p;
    const u64 m2 = 0 - p2;
    u64 k;
    if constexpr (UT)
        asm(FHEHIP_INV_BFLY_ASM : [xo] "=v"(x), [yo] "=v"(y), [k] "=&s"(k) : FHEHIP_INV_BFLY_IN("s")
            : "v0", "v1", "v2", "v3", "v4", "v5", "v6", "v7", "vcc");
    else
        asm(FHEHIP_INV_BFLY_ASM : [xo] "=v"(x), [yo] "=v"(y), [k] "=&s"(k) : FHEHIP_INV_BFLY_IN("v")
            : "v0", "v1", "v2", "v3", "v4", "v5", "v6", "v7", "vcc");
}

__device__ __forceinline__ u64 mul_shoup(u64 a, u64 w, u64 ws, u64 p) {
    return reduce1(lazy_mul_shoup(a, w, ws, p), p);
}

// mul_shoup with w, ws and p uniform across the wave.
__device__ __forceinline__ u64 mul_shoup_u(u64 a, u64 w, u64 ws, u64 p) {
    return reduce1(lazy_mul_shoup_u(a, w, ws, p), p);
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
    // exact reduction of 128-bit sums (reduce128_shoup): [2^64]_p, its Shoup
    // companion, floor(2^64 / p)
    u64 r64, r64_shoup, w1s;
};

}  // namespace fhehip
