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
//   header[16]: nf, nt, start, count, shift, is_one, tg_lo, tg_hi, tg_sign
//   theta_garner[nf][2], theta_omega[nf][3] (lo, hi, sign)
//   per output i < count: p, barrett_lo, barrett_hi, q - gamma,
//                         [2^64 (q - gamma)]_q, [2^64]_q, [2^128]_q, omega[nf]
enum : int {
    SC_NF = 0, SC_NT, SC_START, SC_COUNT, SC_SHIFT, SC_IS_ONE, SC_TG_LO, SC_TG_HI, SC_TG_SIGN,
    SC_HEADER = 16
};
__host__ __device__ inline int sc_out_stride(int nf) { return 7 + nf; }
__host__ __device__ inline int sc_out_base(int nf) { return SC_HEADER + 2 * nf + 3 * nf; }

struct U256 {
    u64 w[4];
};

__device__ __forceinline__ void u256_zero(U256& a) { a.w[0] = a.w[1] = a.w[2] = a.w[3] = 0; }

// a += r * (lo + 2^64 hi)  (product < 2^192), wrapping at 2^256
__device__ __forceinline__ void u256_mac(U256& a, u64 r, u64 lo, u64 hi, bool subtract) {
    u64 p0l, p0h, p1l, p1h;
    mul_wide(r, lo, p0l, p0h);
    mul_wide(r, hi, p1l, p1h);
    const u64 m0 = p0l;
    const u64 m1 = p0h + p1l;
    const u64 m2 = p1h + (m1 < p0h ? 1 : 0);
    if (!subtract) {
        u128 s = (u128)a.w[0] + m0;
        a.w[0] = (u64)s;
        s = (s >> 64) + a.w[1] + m1;
        a.w[1] = (u64)s;
        s = (s >> 64) + a.w[2] + m2;
        a.w[2] = (u64)s;
        a.w[3] += (u64)(s >> 64);
    } else {
        u64 b0 = a.w[0] < m0;
        a.w[0] -= m0;
        u128 t = (u128)m1 + b0;
        u64 b1 = (u128)a.w[1] < t;
        a.w[1] -= (u64)t;
        t = (u128)m2 + b1;
        u64 b2 = (u128)a.w[2] < t;
        a.w[2] -= (u64)t;
        a.w[3] -= b2;
    }
}

__device__ __forceinline__ void u256_add(U256& a, const U256& b) {
    u128 s = (u128)a.w[0] + b.w[0];
    a.w[0] = (u64)s;
    s = (s >> 64) + a.w[1] + b.w[1];
    a.w[1] = (u64)s;
    s = (s >> 64) + a.w[2] + b.w[2];
    a.w[2] = (u64)s;
    a.w[3] = a.w[3] + b.w[3] + (u64)(s >> 64);
}

__device__ __forceinline__ void u256_sub(U256& a, const U256& b) {
    u64 borrow = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        u64 x = a.w[i], y = b.w[i];
        u64 d = x - y - borrow;
        borrow = (x < y) || (x - y < borrow);
        a.w[i] = d;
    }
}

// Low 128 bits of (a >> s), 0 < s < 256.
__device__ __forceinline__ u128 u256_shr_lo128(const U256& a, unsigned s) {
    unsigned ws = s >> 6, bs = s & 63;
    u64 r[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        unsigned k = i + ws;
        u64 lo = k < 4 ? a.w[k] : 0;
        u64 hi = (k + 1) < 4 ? a.w[k + 1] : 0;
        r[i] = bs ? ((lo >> bs) | (hi << (64 - bs))) : lo;
    }
    return ((u128)r[1] << 64) | r[0];
}

// One column of RnsScaler::scale with NF input residues. `rests(j)` yields
// residue j of the column; `emit(i, value)` receives output i (0-based, output
// modulus start + i).
template <int NF, class Rests, class Emit>
__device__ __forceinline__ void rns_scale_column(const u64* __restrict__ sc, Rests rests, Emit emit) {
    constexpr int nf = NF;
    const int count = (int)sc[SC_COUNT];
    const unsigned shift = (unsigned)sc[SC_SHIFT];
    const bool is_one = sc[SC_IS_ONE] != 0;
    const u64* tg = sc + SC_HEADER;
    const u64* tom = tg + 2 * nf;

    // v = ceil((sum r_j theta_garner_j >> (shift - 1)) / 2)
    U256 sum;
    u256_zero(sum);
#pragma unroll
    for (int j = 0; j < nf; j++) u256_mac(sum, rests(j), tg[2 * j], tg[2 * j + 1], false);
    u128 vv = u256_shr_lo128(sum, shift - 1);
    const u128 v = (vv >> 1) + (vv & 1);

    bool w_sign = false;
    u128 w = 0;
    if (!is_one) {
        U256 so;
        u256_zero(so);
#pragma unroll
        for (int j = 0; j < nf; j++) u256_mac(so, rests(j), tom[3 * j], tom[3 * j + 1], tom[3 * j + 2] != 0);
        // v * theta_gamma (128 x 128 -> 256)
        const u64 tgl = sc[SC_TG_LO], tgh = sc[SC_TG_HI];
        U256 vt;
        u256_zero(vt);
        u256_mac(vt, (u64)v, tgl, tgh, false);
        U256 vh;
        u256_zero(vh);
        u256_mac(vh, (u64)(v >> 64), tgl, tgh, false);
        U256 vhs;
        vhs.w[0] = 0;
        vhs.w[1] = vh.w[0];
        vhs.w[2] = vh.w[1];
        vhs.w[3] = vh.w[2];
        u256_add(vt, vhs);
        if (sc[SC_TG_SIGN]) u256_add(so, vt);
        else u256_sub(so, vt);
        w_sign = (so.w[3] >> 63) != 0;
        if (w_sign) {
            U256 nso;
#pragma unroll
            for (int i = 0; i < 4; i++) nso.w[i] = ~so.w[i];
            w = u256_shr_lo128(nso, 126) + 1;
            w >>= 1;
        } else {
            u128 ww = u256_shr_lo128(so, 126);
            w = (ww >> 1) + (ww & 1);
        }
    }

    // Each output residue is the exact value of
    //   2 q - [v]_q gamma (+/- [w]_q) + sum_j r_j omega_ij   mod q,
    // which the reference reduces term by term. Here every term is an exact
    // product added into a 192-bit accumulator (v and w enter through their
    // 64-bit halves: v = v1 2^64 + v0 contributes v0 (q - gamma) +
    // v1 [2^64 (q - gamma)]_q), and the sum is reduced once.
    const u64* outc = sc + sc_out_base(nf);
    const int ostride = sc_out_stride(nf);
    const u64 v0 = (u64)v, v1 = (u64)(v >> 64);
    const u64 w0 = (u64)w, w1 = (u64)(w >> 64);
    for (int i = 0; i < count; i++) {
        const u64* o = outc + (size_t)i * ostride;
        const u64 q = o[0], bl = o[1], bh = o[2];
        // o[3] = q - gamma, o[4] = [2^64 (q - gamma)]_q, o[5] = [2^64]_q, o[6] = [2^128]_q
        u64 lo, hi, top = 0;
        mul_wide(v0, o[3], lo, hi);
        {
            u64 l, h;
            mul_wide(v1, o[4], l, h);
            add192(lo, hi, top, l, h);
        }
        if (!is_one) {
            // +w or -w = w (q - 1) mod q
            u64 l, h;
            if (w_sign) {
                mul_wide(w0, q - 1, l, h);
                add192(lo, hi, top, l, h);
                mul_wide(w1, q - o[5], l, h);  // w1 2^64 (q - 1) = -w1 [2^64]_q
                add192(lo, hi, top, l, h);
            } else {
                add192(lo, hi, top, w0, 0);
                mul_wide(w1, o[5], l, h);
                add192(lo, hi, top, l, h);
            }
        }
        const u64* om = o + 7;
#pragma unroll
        for (int j = 0; j < nf; j++) {
            u64 l, h;
            mul_wide(rests(j), om[j], l, h);
            add192(lo, hi, top, l, h);
        }
        // lo + 2^64 hi + 2^128 top  ->  lo + hi [2^64]_q + top [2^128]_q  (< 2^127) -> Barrett
        u64 l2, h2;
        mul_wide(hi, o[5], l2, h2);
        add128(l2, h2, lo, 0);
        u64 l3, h3;
        mul_wide(top, o[6], l3, h3);
        add128(l2, h2, l3, h3);
        emit(i, reduce_u128(l2, h2, bl, bh, q));
    }
}

}  // namespace fhehip
