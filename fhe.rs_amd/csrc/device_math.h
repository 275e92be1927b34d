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

__device__ __forceinline__ u64 reduce1(u64 x, u64 p) { return x >= p ? x - p : x; }

// lazy_mul_shoup (math/src/zq/mod.rs:224-234): a*w mod p in [0, 2p) for any a < 2^64.
__device__ __forceinline__ u64 lazy_mul_shoup(u64 a, u64 w, u64 ws, u64 p) {
    u64 q = __umul64hi(a, ws);
    return a * w - q * p;
}

__device__ __forceinline__ u64 mul_shoup(u64 a, u64 w, u64 ws, u64 p) {
    return reduce1(lazy_mul_shoup(a, w, ws, p), p);
}

// Barrett reduction of a 128-bit value with bh:bl = floor(2^128 / p)
// (lazy_reduce_u128, math/src/zq/mod.rs:693-707); result in [0, 2p).
__device__ __forceinline__ u64 lazy_reduce_u128(u64 a_lo, u64 a_hi, u64 bl, u64 bh, u64 p) {
    u64 lolo_hi = __umul64hi(a_lo, bl);
    u128 hilo = (u128)a_hi * bl;
    u128 lohi = (u128)a_lo * bh;
    u128 mid = hilo + lohi + lolo_hi;
    u64 q = (u64)(mid >> 64) + a_hi * bh;
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
    u64 lolo_hi = __umul64hi(a, bl);
    u128 lohi = (u128)a * bh;
    u64 q = (u64)((lohi + lolo_hi) >> 64);
    return a - q * p;
}

__device__ __forceinline__ u64 add_mod(u64 a, u64 b, u64 p) { return reduce1(a + b, p); }
__device__ __forceinline__ u64 sub_mod(u64 a, u64 b, u64 p) { return reduce1(a + p - b, p); }

// Per-modulus constants kept resident on the device.
struct ModInfo {
    u64 p;
    u64 barrett_lo;  // floor(2^128 / p), low word
    u64 barrett_hi;  // floor(2^128 / p), high word
    u64 n_inv;       // N^-1 mod p
    u64 n_inv_shoup;
    u64 pad[3];
};

}  // namespace fhehip
