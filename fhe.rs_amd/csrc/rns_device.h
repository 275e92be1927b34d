// Device restatement of RnsScaler::scale (math/src/rns/scaler.rs:249-352).
//
// One thread handles one coefficient column: it reads the |from| residues of
// the column, forms v = round(sum_j r_j * theta_garner_j / 2^shift) with the
// same 256-bit wrapping fixed-point arithmetic and rounding steps as the
// reference, the signed correction w when the factor is not one, and then
// each output residue y_i = reduce(2 q_i - [v]_{q_i} gamma_i (+/- [w]_{q_i})
// + sum_j r_j omega_ij). The result is bit-identical to the reference,
// including its behaviour at rounding ties.
#pragma once

#include "device_math.h"

namespace fhehip {

// Flat constant block (u64 words) written by the host (bfv.cpp):
//   header[16]: nf, nt, start, count, shift, is_one, tg_lo, tg_hi, tg_sign,
//               flags, omega_mask
//   theta_garner[nf][2], theta_omega[nf][3] (lo, hi, sign)
//   per output i < count: p, barrett_lo, barrett_hi, q - gamma,
//                         [2^64 (q - gamma)]_q, [2^64]_q, [2^128]_q,
//                         shoup([2^64]_q), shoup(1) = floor(2^64 / q), omega[nf]
// flags: SC_F_SUM128 when the host has shown that every output's exact sum
// (below) stays under 2^128 for inputs below 2 q_j; SC_F_TG_ZERO when
// theta_gamma is zero. omega_mask: bit j set when theta_omega_j is not zero
// (t/Q scaling has zero theta_omega for every modulus of the extension part).
enum : int {
    SC_NF = 0, SC_NT, SC_START, SC_COUNT, SC_SHIFT, SC_IS_ONE, SC_TG_LO, SC_TG_HI, SC_TG_SIGN, SC_FLAGS,
    SC_OMEGA_MASK,
    SC_HEADER = 16
};
enum : unsigned { SC_F_SUM128 = 1, SC_F_TG_ZERO = 2 };
__host__ __device__ inline int sc_out_stride(int nf) { return 9 + nf; }
__host__ __device__ inline int sc_out_base(int nf) { return SC_HEADER + 2 * nf + 3 * nf; }

// Scaler arithmetic with explicit 32-bit add-with-carry chains. In these
// long exact sums the chains replace the compare-and-select carries and the
// register-pair moves of the 64-bit formulation (about 28% fewer VALU
// instructions per column).
__device__ __forceinline__ void sc_mul_wide(u64 a, u64 b, u64& lo, u64& hi) {
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    const u64 p00 = (u64)a0 * b0, p01 = (u64)a0 * b1, p10 = (u64)a1 * b0, p11 = (u64)a1 * b1;
    unsigned c1, c2, c3, c4;
    u32 r1 = __builtin_addc((u32)(p00 >> 32), (u32)p01, 0u, &c1);
    r1 = __builtin_addc(r1, (u32)p10, 0u, &c2);
    u32 r2 = __builtin_addc((u32)(p01 >> 32), (u32)(p10 >> 32), c1, &c3);
    r2 = __builtin_addc(r2, (u32)p11, c2, &c4);
    const u32 r3 = (u32)(p11 >> 32) + c3 + c4;
    lo = ((u64)r1 << 32) | (u32)p00;
    hi = ((u64)r3 << 32) | r2;
}

// (lo, hi) += (alo, ahi) modulo 2^128.
__device__ __forceinline__ void sc_add128(u64& lo, u64& hi, u64 alo, u64 ahi) {
    unsigned c;
    const u32 r0 = __builtin_addc((u32)lo, (u32)alo, 0u, &c);
    const u32 r1 = __builtin_addc((u32)(lo >> 32), (u32)(alo >> 32), c, &c);
    const u32 r2 = __builtin_addc((u32)hi, (u32)ahi, c, &c);
    const u32 r3 = (u32)(hi >> 32) + (u32)(ahi >> 32) + c;
    lo = ((u64)r1 << 32) | r0;
    hi = ((u64)r3 << 32) | r2;
}

// (lo, hi, top) += (alo, ahi).
__device__ __forceinline__ void sc_add192(u64& lo, u64& hi, u64& top, u64 alo, u64 ahi) {
    unsigned c;
    const u32 r0 = __builtin_addc((u32)lo, (u32)alo, 0u, &c);
    const u32 r1 = __builtin_addc((u32)(lo >> 32), (u32)(alo >> 32), c, &c);
    const u32 r2 = __builtin_addc((u32)hi, (u32)ahi, c, &c);
    const u32 r3 = __builtin_addc((u32)(hi >> 32), (u32)(ahi >> 32), c, &c);
    const u32 r4 = __builtin_addc((u32)top, 0u, c, &c);
    const u32 r5 = (u32)(top >> 32) + c;
    lo = ((u64)r1 << 32) | r0;
    hi = ((u64)r3 << 32) | r2;
    top = ((u64)r5 << 32) | r4;
}

// High 64 bits of a 64 x 64 product.
__device__ __forceinline__ u64 sc_umulhi(u64 a, u64 b) {
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    const u32 h00 = __umulhi(a0, b0);
    const u64 p01 = (u64)a0 * b1, p10 = (u64)a1 * b0, p11 = (u64)a1 * b1;
    unsigned c1, c2, c3, c4;
    const u32 t = __builtin_addc(h00, (u32)p01, 0u, &c1);
    (void)__builtin_addc(t, (u32)p10, 0u, &c2);
    u32 r2 = __builtin_addc((u32)(p01 >> 32), (u32)(p10 >> 32), c1, &c3);
    r2 = __builtin_addc(r2, (u32)p11, c2, &c4);
    const u32 r3 = (u32)(p11 >> 32) + c3 + c4;
    return ((u64)r3 << 32) | r2;
}

// Exact a mod p for a < 2^127 (Barrett with bh:bl = floor(2^128 / p), as
// lazy_reduce_u128 + reduce1).
__device__ __forceinline__ u64 sc_reduce_u128(u64 a_lo, u64 a_hi, u64 bl, u64 bh, u64 p) {
    const u64 lolo_hi = sc_umulhi(a_lo, bl);
    u64 hl_lo, hl_hi, lh_lo, lh_hi;
    sc_mul_wide(a_hi, bl, hl_lo, hl_hi);
    sc_mul_wide(a_lo, bh, lh_lo, lh_hi);
    u64 m_lo = hl_lo, m_hi = hl_hi;
    sc_add128(m_lo, m_hi, lh_lo, lh_hi);
    sc_add128(m_lo, m_hi, lolo_hi, 0);
    const u64 q = m_hi + a_hi * bh;
    return reduce1(a_lo - q * p, p);
}

// Exact sums of 64 x 64-bit products in column form. With r = r1 2^32 + r0
// and w = w1 2^32 + w0, r w = r0 w0 + 2^32 (r0 w1 + r1 w0) + 2^64 r1 w1; the
// three column sums are kept in 64-bit words a, b, c whose carries out are
// counted in ka, kb, kc, so that
//   S = a + 2^32 b + 2^64 c + 2^64 ka + 2^96 kb + 2^128 kc
// exactly. One product costs four v_mad_u64_u32 (the carry out of each add
// goes to an SGPR pair) and four v_addc_co_u32 counting those carries; every
// carry is read at least two instructions after the VALU wrote it.
struct Acc3 {
    u64 a, b, c;
    u32 ka, kb, kc;
};

__device__ __forceinline__ void acc3_init(Acc3& s, u64 a0) {
    s.a = a0;
    s.b = s.c = 0;
    s.ka = s.kb = s.kc = 0;
}

#define FHEHIP_ACC3_ASM                                    \
    "v_mad_u64_u32 %[a], %[k0], %[r0], %[w0], %[a]\n\t"    \
    "v_mad_u64_u32 %[b], %[k1], %[r0], %[w1], %[b]\n\t"    \
    "v_mad_u64_u32 %[c], %[k2], %[r1], %[w1], %[c]\n\t"    \
    "v_addc_co_u32 %[ka], %[d], %[ka], 0, %[k0]\n\t"       \
    "v_mad_u64_u32 %[b], %[k0], %[r1], %[w0], %[b]\n\t"    \
    "v_addc_co_u32 %[kb], %[d], %[kb], 0, %[k1]\n\t"       \
    "v_addc_co_u32 %[kc], %[d], %[kc], 0, %[k2]\n\t"       \
    "v_addc_co_u32 %[kb], %[d], %[kb], 0, %[k0]"

// s += r * w. SW: w is uniform across the wave (passed in SGPRs).
template <bool SW>
__device__ __forceinline__ void acc3_mul(Acc3& s, u64 r, u64 w) {
    u64 k0, k1, k2, d;
    if constexpr (SW)
        asm(FHEHIP_ACC3_ASM
            : [a] "+v"(s.a), [b] "+v"(s.b), [c] "+v"(s.c), [ka] "+v"(s.ka), [kb] "+v"(s.kb), [kc] "+v"(s.kc),
              [k0] "=&s"(k0), [k1] "=&s"(k1), [k2] "=&s"(k2), [d] "=&s"(d)
            : [r0] "v"((u32)r), [r1] "v"((u32)(r >> 32)), [w0] "s"((u32)w), [w1] "s"((u32)(w >> 32)));
    else
        asm(FHEHIP_ACC3_ASM
            : [a] "+v"(s.a), [b] "+v"(s.b), [c] "+v"(s.c), [ka] "+v"(s.ka), [kb] "+v"(s.kb), [kc] "+v"(s.kc),
              [k0] "=&s"(k0), [k1] "=&s"(k1), [k2] "=&s"(k2), [d] "=&s"(d)
            : [r0] "v"((u32)r), [r1] "v"((u32)(r >> 32)), [w0] "v"((u32)w), [w1] "v"((u32)(w >> 32)));
}

// s = r * w (starts an accumulator without zeroing it first): the only
// carry is the one of the middle column.
#define FHEHIP_ACC3_FIRST_ASM                              \
    "v_mad_u64_u32 %[b], %[k0], %[r0], %[w1], 0\n\t"      \
    "v_mad_u64_u32 %[a], %[k1], %[r0], %[w0], 0\n\t"      \
    "v_mad_u64_u32 %[c], %[k1], %[r1], %[w1], 0\n\t"      \
    "v_mad_u64_u32 %[b], %[k0], %[r1], %[w0], %[b]\n\t"   \
    "v_mov_b32 %[ka], 0\n\t"                              \
    "v_mov_b32 %[kc], 0\n\t"                              \
    "v_addc_co_u32 %[kb], %[k1], 0, 0, %[k0]"

template <bool SW>
__device__ __forceinline__ void acc3_first(Acc3& s, u64 r, u64 w) {
    u64 k0, k1;
    if constexpr (SW)
        asm(FHEHIP_ACC3_FIRST_ASM
            : [a] "=&v"(s.a), [b] "=&v"(s.b), [c] "=&v"(s.c), [ka] "=&v"(s.ka), [kb] "=&v"(s.kb), [kc] "=&v"(s.kc),
              [k0] "=&s"(k0), [k1] "=&s"(k1)
            : [r0] "v"((u32)r), [r1] "v"((u32)(r >> 32)), [w0] "s"((u32)w), [w1] "s"((u32)(w >> 32)));
    else
        asm(FHEHIP_ACC3_FIRST_ASM
            : [a] "=&v"(s.a), [b] "=&v"(s.b), [c] "=&v"(s.c), [ka] "=&v"(s.ka), [kb] "=&v"(s.kb), [kc] "=&v"(s.kc),
              [k0] "=&s"(k0), [k1] "=&s"(k1)
            : [r0] "v"((u32)r), [r1] "v"((u32)(r >> 32)), [w0] "v"((u32)w), [w1] "v"((u32)(w >> 32)));
}

// The 192-bit value of an accumulator (it is below 2^192).
__device__ __forceinline__ void acc3_value(const Acc3& s, u64& lo, u64& hi, u64& top) {
    unsigned long long c1, c2, c3;
    lo = __builtin_addcll(s.a, s.b << 32, 0, &c1);
    hi = __builtin_addcll(s.c, (s.b >> 32) | ((u64)s.kb << 32), c1, &c2);
    hi = __builtin_addcll(hi, (u64)s.ka, 0, &c3);
    top = (u64)s.kc + c2 + c3;
}

// Column accumulator of a sum known to stay below 2^128: the same columns
// as Acc3, but the top column c cannot carry out (its carry count would
// weigh 2^128), so one product costs four v_mad_u64_u32 and three
// v_addc_co_u32. Every carry is read at least three instructions after the
// VALU wrote it.
struct Acc2 {
    u64 a, b, c;
    u32 ka, kb;
};

#define FHEHIP_ACC2_ASM                                    \
    "v_mad_u64_u32 %[a], %[k0], %[r0], %[w0], %[a]\n\t"    \
    "v_mad_u64_u32 %[b], %[k1], %[r0], %[w1], %[b]\n\t"    \
    "v_mad_u64_u32 %[b], %[k2], %[r1], %[w0], %[b]\n\t"    \
    "v_addc_co_u32 %[ka], %[d], %[ka], 0, %[k0]\n\t"       \
    "v_mad_u64_u32 %[c], %[d], %[r1], %[w1], %[c]\n\t"     \
    "v_addc_co_u32 %[kb], %[d], %[kb], 0, %[k1]\n\t"       \
    "v_addc_co_u32 %[kb], %[d], %[kb], 0, %[k2]"

// s += r * w. SW: w is uniform across the wave (passed in SGPRs).
template <bool SW>
__device__ __forceinline__ void acc2_mul(Acc2& s, u64 r, u64 w) {
    u64 k0, k1, k2, d;
    if constexpr (SW)
        asm(FHEHIP_ACC2_ASM
            : [a] "+v"(s.a), [b] "+v"(s.b), [c] "+v"(s.c), [ka] "+v"(s.ka), [kb] "+v"(s.kb), [k0] "=&s"(k0),
              [k1] "=&s"(k1), [k2] "=&s"(k2), [d] "=&s"(d)
            : [r0] "v"((u32)r), [r1] "v"((u32)(r >> 32)), [w0] "s"((u32)w), [w1] "s"((u32)(w >> 32)));
    else
        asm(FHEHIP_ACC2_ASM
            : [a] "+v"(s.a), [b] "+v"(s.b), [c] "+v"(s.c), [ka] "+v"(s.ka), [kb] "+v"(s.kb), [k0] "=&s"(k0),
              [k1] "=&s"(k1), [k2] "=&s"(k2), [d] "=&s"(d)
            : [r0] "v"((u32)r), [r1] "v"((u32)(r >> 32)), [w0] "v"((u32)w), [w1] "v"((u32)(w >> 32)));
}

// s = r * w: the first two products add to zero and cannot carry; the
// carry of the second middle product is read three instructions later.
#define FHEHIP_ACC2_FIRST_ASM                              \
    "v_mad_u64_u32 %[b], %[k1], %[r0], %[w1], 0\n\t"      \
    "v_mad_u64_u32 %[a], %[k1], %[r0], %[w0], 0\n\t"      \
    "v_mad_u64_u32 %[b], %[k0], %[r1], %[w0], %[b]\n\t"   \
    "v_mov_b32 %[ka], 0\n\t"                              \
    "v_mad_u64_u32 %[c], %[k1], %[r1], %[w1], 0\n\t"      \
    "v_addc_co_u32 %[kb], %[k1], 0, 0, %[k0]"

template <bool SW>
__device__ __forceinline__ void acc2_first(Acc2& s, u64 r, u64 w) {
    u64 k0, k1;
    if constexpr (SW)
        asm(FHEHIP_ACC2_FIRST_ASM
            : [a] "=&v"(s.a), [b] "=&v"(s.b), [c] "=&v"(s.c), [ka] "=&v"(s.ka), [kb] "=&v"(s.kb), [k0] "=&s"(k0),
              [k1] "=&s"(k1)
            : [r0] "v"((u32)r), [r1] "v"((u32)(r >> 32)), [w0] "s"((u32)w), [w1] "s"((u32)(w >> 32)));
    else
        asm(FHEHIP_ACC2_FIRST_ASM
            : [a] "=&v"(s.a), [b] "=&v"(s.b), [c] "=&v"(s.c), [ka] "=&v"(s.ka), [kb] "=&v"(s.kb), [k0] "=&s"(k0),
              [k1] "=&s"(k1)
            : [r0] "v"((u32)r), [r1] "v"((u32)(r >> 32)), [w0] "v"((u32)w), [w1] "v"((u32)(w >> 32)));
}

// The 128-bit value of an Acc2 accumulator.
__device__ __forceinline__ void acc2_value(const Acc2& s, u64& lo, u64& hi) {
    unsigned long long c1, c2;
    lo = __builtin_addcll(s.a, s.b << 32, 0, &c1);
    hi = __builtin_addcll(s.c, (s.b >> 32) | ((u64)s.kb << 32), c1, &c2);
    hi += (u64)s.ka;
}

// a - floor(a * ws / 2^64) * p (mod 2^64) with ws = floor(2^64 / p): a mod p
// in [0, 2p) for any a < 2^64 (lazy_mul_shoup by one, math/src/zq/mod.rs
// :224-234, without the product by the multiplier). Temporaries and hazard
// spacing as FHEHIP_SHOUP_ASM.
#define FHEHIP_SHOUP1_ASM                                   \
    "v_mov_b32 v1, 0\n\t"                                   \
    "v_mul_hi_u32 v0, %[a0], %[s0]\n\t"                     \
    "v_mad_u64_u32 v[2:3], %[k], %[a0], %[s1], v[0:1]\n\t"  \
    "v_mad_u64_u32 v[2:3], %[k], %[a1], %[s0], v[2:3]\n\t"  \
    "v_mov_b32 v0, v3\n\t"                                  \
    "s_nop 1\n\t"                                           \
    "v_addc_co_u32 v1, %[k], 0, 0, %[k]\n\t"                \
    "v_mad_u64_u32 v[0:1], %[k], %[a1], %[s1], v[0:1]\n\t"  \
    "v_mad_u64_u32 v[2:3], %[k], v0, %[n0], %[a]\n\t"      \
    "v_mul_lo_u32 v4, v0, %[n1]\n\t"                        \
    "v_mul_lo_u32 v5, v1, %[n0]\n\t"                        \
    "v_add3_u32 v3, v3, v4, v5"

template <bool SW>
__device__ __forceinline__ u64 lazy_reduce_shoup1(u64 a, u64 ws, u64 p) {
    const u64 n = 0 - p;
    u64 r, k;
    if constexpr (SW)
        asm(FHEHIP_SHOUP1_ASM
            : "=&{v[2:3]}"(r), [k] "=&s"(k)
            : [a] "v"(a), [a0] "v"((u32)a), [a1] "v"((u32)(a >> 32)), [s0] "s"((u32)ws), [s1] "s"((u32)(ws >> 32)),
              [n0] "s"((u32)n), [n1] "s"((u32)(n >> 32))
            : "v0", "v1", "v4", "v5");
    else
        asm(FHEHIP_SHOUP1_ASM
            : "=&{v[2:3]}"(r), [k] "=&s"(k)
            : [a] "v"(a), [a0] "v"((u32)a), [a1] "v"((u32)(a >> 32)), [s0] "v"((u32)ws), [s1] "v"((u32)(ws >> 32)),
              [n0] "v"((u32)n), [n1] "v"((u32)(n >> 32))
            : "v0", "v1", "v4", "v5");
    return r;
}

// Exact (lo + 2^64 hi) mod p for p < 2^62: hi [2^64]_p by a lazy Shoup
// product (r64, r64s = shoup([2^64]_p)) and lo by a lazy Shoup reduction
// (w1s = floor(2^64 / p)), each in [0, 2p), then two conditional
// subtractions. SW: the constants are uniform across the wave.
template <bool SW>
__device__ __forceinline__ u64 reduce128_shoup(u64 lo, u64 hi, u64 r64, u64 r64s, u64 w1s, u64 p) {
    const u64 x = (SW ? lazy_mul_shoup_u(hi, r64, r64s, p) : lazy_mul_shoup(hi, r64, r64s, p)) +
                  lazy_reduce_shoup1<SW>(lo, w1s, p);
    return reduce1(reduce1(x, 2 * p), p);
}

struct U256 {
    u64 w[4];
};

__device__ __forceinline__ void u256_zero(U256& a) { a.w[0] = a.w[1] = a.w[2] = a.w[3] = 0; }

// a = (a0 + 2^64 a1 + 2^128 a2) + 2^64 (b0 + 2^64 b1 + 2^128 b2), modulo 2^256.
__device__ __forceinline__ void u256_combine(U256& a, u64 a0, u64 a1, u64 a2, u64 b0, u64 b1, u64 b2) {
    unsigned long long c1, c2, c3;
    a.w[0] = a0;
    a.w[1] = __builtin_addcll(a1, b0, 0, &c1);
    a.w[2] = __builtin_addcll(a2, b1, c1, &c2);
    a.w[3] = __builtin_addcll(b2, 0, c2, &c3);
}

__device__ __forceinline__ void u256_add(U256& a, const U256& b) {
    unsigned long long c0, c1, c2, c3;
    a.w[0] = __builtin_addcll(a.w[0], b.w[0], 0, &c0);
    a.w[1] = __builtin_addcll(a.w[1], b.w[1], c0, &c1);
    a.w[2] = __builtin_addcll(a.w[2], b.w[2], c1, &c2);
    a.w[3] = __builtin_addcll(a.w[3], b.w[3], c2, &c3);
}

__device__ __forceinline__ void u256_sub(U256& a, const U256& b) {
    unsigned long long c0, c1, c2, c3;
    a.w[0] = __builtin_subcll(a.w[0], b.w[0], 0, &c0);
    a.w[1] = __builtin_subcll(a.w[1], b.w[1], c0, &c1);
    a.w[2] = __builtin_subcll(a.w[2], b.w[2], c1, &c2);
    a.w[3] = __builtin_subcll(a.w[3], b.w[3], c2, &c3);
}

// Low 128 bits of (a >> s), 0 < s < 256. The words are picked by selects
// (no indexing of a private array by the shift, which would go to scratch).
__device__ __forceinline__ u128 u256_shr_lo128(const U256& a, unsigned s) {
    const unsigned ws = s >> 6, bs = s & 63;
    auto word = [&](unsigned k) -> u64 {
        return k == 0 ? a.w[0] : k == 1 ? a.w[1] : k == 2 ? a.w[2] : k == 3 ? a.w[3] : 0;
    };
    const u64 x0 = word(ws), x1 = word(ws + 1), x2 = word(ws + 2);
    const u64 r0 = bs ? ((x0 >> bs) | (x1 << (64 - bs))) : x0;
    const u64 r1 = bs ? ((x1 >> bs) | (x2 << (64 - bs))) : x1;
    return ((u128)r1 << 64) | r0;
}

// v = ceil((S >> (shift - 1)) / 2) from the low 128 bits of the shifted
// sum S = sum_j r_j theta_garner_j.
__device__ __forceinline__ u128 sc_round_v(const U256& sum, unsigned shift) {
    const u128 vv = u256_shr_lo128(sum, shift - 1);
    return (vv >> 1) + (vv & 1);
}

// The magnitude of w = round(so / 2^127) with the reference's rounding of
// each sign (so is a signed 256-bit value), and its sign.
__device__ __forceinline__ u128 sc_round_w(const U256& so, bool& w_sign) {
    w_sign = (so.w[3] >> 63) != 0;
    if (w_sign) {
        U256 nso;
#pragma unroll
        for (int i = 0; i < 4; i++) nso.w[i] = ~so.w[i];
        return (u256_shr_lo128(nso, 126) + 1) >> 1;
    }
    const u128 ww = u256_shr_lo128(so, 126);
    return (ww >> 1) + (ww & 1);
}

// A small header field of the constant block as a wave-uniform value: read
// through the scalar cache (SW), or from a copy in LDS and moved to an SGPR.
template <bool SW>
__device__ __forceinline__ u32 sc_uniform(u64 v) {
    if constexpr (SW) return (u32)v;
    else return (u32)__builtin_amdgcn_readfirstlane((int)(u32)v);
}

// The per-column values of RnsScaler::scale that every output uses: v and
// the signed correction w (zero when the factor is one), as 64-bit halves.
struct ScColumn {
    u64 v0, v1, w0, w1;
    bool w_sign;
};

// Header fields of the constant block as wave-uniform values.
template <bool SW>
struct ScHeader {
    int count;
    unsigned shift;
    bool is_one;
    unsigned flags;
    unsigned omega_mask;
    __device__ __forceinline__ explicit ScHeader(const u64* __restrict__ sc)
        : count((int)sc_uniform<SW>(sc[SC_COUNT])),
          shift(sc_uniform<SW>(sc[SC_SHIFT])),
          is_one(sc_uniform<SW>(sc[SC_IS_ONE]) != 0),
          flags(sc_uniform<SW>(sc[SC_FLAGS])),
          omega_mask(sc_uniform<SW>(sc[SC_OMEGA_MASK])) {}
};

// v and w of one column with NF input residues; `rests(j)` yields residue j.
// SW: `sc` is global memory and the constants enter the products from SGPRs;
// otherwise `sc` is a copy in LDS and they enter from VGPRs.
template <int NF, bool SW, class Rests>
__device__ __forceinline__ ScColumn sc_column(const u64* __restrict__ sc, Rests rests, const ScHeader<SW>& h) {
    constexpr int nf = NF;
    const u64* tg = sc + SC_HEADER;
    const u64* tom = tg + 2 * nf;

    // v = ceil((sum r_j theta_garner_j >> (shift - 1)) / 2). The 256-bit sum
    // is formed as A + 2^64 B with A = sum r_j lo_j and B = sum r_j hi_j, each
    // an exact 192-bit sum (no term exceeds 2^128), combined once.
    U256 sum;
    {
        Acc3 sa, sb;
        acc3_first<SW>(sa, rests(0), tg[0]);
        acc3_first<SW>(sb, rests(0), tg[1]);
#pragma unroll
        for (int j = 1; j < nf; j++) {
            acc3_mul<SW>(sa, rests(j), tg[2 * j]);
            acc3_mul<SW>(sb, rests(j), tg[2 * j + 1]);
        }
        u64 a0, a1, a2, b0, b1, b2;
        acc3_value(sa, a0, a1, a2);
        acc3_value(sb, b0, b1, b2);
        u256_combine(sum, a0, a1, a2, b0, b1, b2);
    }
    const u128 v = sc_round_v(sum, h.shift);

    bool w_sign = false;
    u128 w = 0;
    if (!h.is_one) {
        // sum_j +-r_j theta_omega_j with the signs of the constants, without
        // a branch on the sign: a negative theta is taken as its 128-bit
        // complement 2^128 - theta, so that
        //   sum = sum_j r_j theta'_j - 2^128 sum_{negative j} r_j
        // exactly (the signs and constants are uniform, so theta' and the
        // mask are scalar values). Zero thetas (omega_mask) add nothing and
        // are skipped by uniform branches.
        Acc3 wa, wb;
        acc3_init(wa, 0);
        acc3_init(wb, 0);
        u64 rn0 = 0, rn1 = 0;
#pragma unroll
        for (int j = 0; j < nf; j++) {
            if (!((h.omega_mask >> j) & 1)) continue;
            const u64 lo = tom[3 * j], hi = tom[3 * j + 1];
            const bool neg = tom[3 * j + 2] != 0;
            const u64 clo = neg ? 0 - lo : lo;
            const u64 chi = neg ? ~hi + (lo == 0 ? 1 : 0) : hi;
            acc3_mul<SW>(wa, rests(j), clo);
            acc3_mul<SW>(wb, rests(j), chi);
            if (sc_uniform<SW>(tom[3 * j + 2])) add128(rn0, rn1, rests(j), 0);
        }
        u64 p0, p1, p2, q0, q1, q2;
        acc3_value(wa, p0, p1, p2);
        acc3_value(wb, q0, q1, q2);
        U256 so, corr;
        u256_combine(so, p0, p1, p2, q0, q1, q2);
        corr.w[0] = corr.w[1] = 0;
        corr.w[2] = rn0;
        corr.w[3] = rn1;
        u256_sub(so, corr);
        if (!(h.flags & SC_F_TG_ZERO)) {
            // v * theta_gamma (128 x 128 -> 256)
            const u64 tgl = sc[SC_TG_LO], tgh = sc[SC_TG_HI];
            U256 vt;
            {
                u64 a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0, l, h2;
                sc_mul_wide((u64)v, tgl, l, h2);
                sc_add192(a0, a1, a2, l, h2);
                sc_mul_wide((u64)(v >> 64), tgl, l, h2);
                sc_add192(b0, b1, b2, l, h2);
                sc_mul_wide((u64)v, tgh, l, h2);
                sc_add192(b0, b1, b2, l, h2);
                u256_combine(vt, a0, a1, a2, b0, b1, b2);
                sc_mul_wide((u64)(v >> 64), tgh, l, h2);  // times 2^128
                vt.w[2] += l;
                vt.w[3] += h2 + (vt.w[2] < l ? 1 : 0);
            }
            if (sc_uniform<SW>(sc[SC_TG_SIGN])) u256_add(so, vt);
            else u256_sub(so, vt);
        }
        w = sc_round_w(so, w_sign);
    }
    return ScColumn{(u64)v, (u64)(v >> 64), (u64)w, (u64)(w >> 64), w_sign};
}

// Each output residue is the exact value of
//   2 q - [v]_q gamma (+/- [w]_q) + sum_j r_j omega_ij   mod q,
// which the reference reduces term by term. Here every term is an exact
// product added into one column accumulator (v and w enter through their
// 64-bit halves: v = v1 2^64 + v0 contributes v0 (q - gamma) +
// v1 [2^64 (q - gamma)]_q), and the sum is reduced once:
// - SC_F_SUM128 (the sum stays below 2^128): Acc2 and reduce128_shoup;
// - otherwise an Acc3 sum lo + 2^64 hi + 2^128 top, folded to
//   lo + hi [2^64]_q + top [2^128]_q (< 2^127) and reduced by Barrett.
// +w, or -w = w (q - 1) mod q, enters without a branch on the column's sign
// as w0 c0 + w1 c1 with (c0, c1) = (1, [2^64]_q) or (q - 1, q - [2^64]_q).
template <int NF, bool SW, class Rests>
__device__ __forceinline__ u64 sc_output(const u64* __restrict__ o, Rests rests, const ScColumn& c,
                                         const ScHeader<SW>& h) {
    constexpr int nf = NF;
    const u64 q = o[0];
    // o[3] = q - gamma, o[4] = [2^64 (q - gamma)]_q, o[5] = [2^64]_q, o[6] = [2^128]_q,
    // o[7] = shoup([2^64]_q), o[8] = floor(2^64 / q)
    const u64* om = o + 9;
    if (h.flags & SC_F_SUM128) {
        Acc2 acc;
        acc2_first<SW>(acc, c.v0, o[3]);
        acc2_mul<SW>(acc, c.v1, o[4]);
        if (!h.is_one) {
            const u64 c0 = c.w_sign ? q - 1 : 1;
            const u64 c1 = c.w_sign ? q - o[5] : o[5];
            acc2_mul<false>(acc, c.w0, c0);
            acc2_mul<false>(acc, c.w1, c1);
        }
#pragma unroll
        for (int j = 0; j < nf; j++) acc2_mul<SW>(acc, rests(j), om[j]);
        u64 lo, hi;
        acc2_value(acc, lo, hi);
        return reduce128_shoup<SW>(lo, hi, o[5], o[7], o[8], q);
    }
    Acc3 acc;
    acc3_first<SW>(acc, c.v0, o[3]);
    acc3_mul<SW>(acc, c.v1, o[4]);
    if (!h.is_one) {
        const u64 c0 = c.w_sign ? q - 1 : 1;
        const u64 c1 = c.w_sign ? q - o[5] : o[5];
        acc3_mul<false>(acc, c.w0, c0);
        acc3_mul<false>(acc, c.w1, c1);
    }
#pragma unroll
    for (int j = 0; j < nf; j++) acc3_mul<SW>(acc, rests(j), om[j]);
    u64 lo, hi, top;
    acc3_value(acc, lo, hi, top);
    Acc3 fold;
    acc3_init(fold, lo);
    acc3_mul<SW>(fold, hi, o[5]);
    acc3_mul<SW>(fold, top, o[6]);
    u64 l2, h2, t2;
    acc3_value(fold, l2, h2, t2);
    return sc_reduce_u128(l2, h2, o[1], o[2], q);
}

// One column of RnsScaler::scale with NF input residues. `rests(j)` yields
// residue j of the column; `emit(i, value)` receives output i (0-based, output
// modulus start + i).
template <int NF, bool SW = true, class Rests, class Emit>
__device__ __forceinline__ void rns_scale_column(const u64* __restrict__ sc, Rests rests, Emit emit) {
    const ScHeader<SW> h(sc);
    const ScColumn c = sc_column<NF, SW>(sc, rests, h);
    const u64* outc = sc + sc_out_base(NF);
    for (int i = 0; i < h.count; i++) emit(i, sc_output<NF, SW>(outc + (size_t)i * sc_out_stride(NF), rests, c, h));
}

}  // namespace fhehip
