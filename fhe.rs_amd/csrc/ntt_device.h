// Negacyclic NTT building blocks for one polynomial row held by one workgroup.
//
// Semantics follow NttOperator (math/src/ntt/native.rs:77-234): the forward
// transform is the Cooley-Tukey network with merged psi twiddles
// omegas[k] = psi^brv(k) and bit-reversed output; the inverse is the
// Gentleman-Sande network with zetas_inv[k] = psi^-(brv(k)+1) followed by a
// multiplication by N^-1. The butterflies are the Harvey lazy ones of
// native.rs:256-300 (values < 4p forward, < 2p inverse).
//
// Layout of the work: a row of N = 2^LOGN u64 is split over T = N / 16
// threads, each holding E = 16 values in registers. A "pass" runs up to four
// consecutive stages on those registers; between passes the row is exchanged
// through LDS (N + N/16 words with one pad word every 16 to spread banks).
#pragma once

#include "device_math.h"

namespace fhehip {

// RE_WANT = log2 of the values each thread holds (4: 16 values, 1024 threads
// for N = 16384; 5: 32 values, 512 threads).
template <int LOGN, int RE_WANT = 4>
struct NttShape {
    static constexpr int N = 1 << LOGN;
    static constexpr int RE = LOGN >= RE_WANT ? RE_WANT : LOGN;  // log2 elements per thread
    static constexpr int E = 1 << RE;
    static constexpr int T = N / E;  // threads per row
    static constexpr int LDS_WORDS = N + (N >> 4) + 1;
};

// Workgroups are dispatched round-robin over the chip's 8 XCDs, each with its
// own L2: workgroup b of a 1-D grid runs on XCD b % 8. This returns the
// logical index of workgroup b in an order that gives each XCD a contiguous
// range, so consecutive logical indices (workgroups reading the same rows)
// run on one XCD at about the same time and share its L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned total) {
    const unsigned xcd = b & 7, slot = b >> 3, base = total >> 3, rem = total & 7;
    return xcd * base + (xcd < rem ? xcd : rem) + slot;
}

// LDS word of element j: one pad word every 16 spreads the strided exchange
// patterns over the banks.
__device__ __forceinline__ int lds_slot(int j) { return j + (j >> 4); }

// Forward butterfly (native.rs:256-269): x, y < 4p -> x', y' < 4p. UT: the
// twiddle is the same for the whole wave.
template <bool UT = false>
__device__ __forceinline__ void fwd_bfly(u64& x, u64& y, u64 w, u64 ws, u64 p, u64 p2) {
    fwd_bfly_asm<UT>(x, y, w, ws, p, p2);
}

// Inverse butterfly (native.rs:288-300): x, y < 2p -> x', y' < 2p.
template <bool UT = false>
__device__ __forceinline__ void inv_bfly(u64& x, u64& y, u64 z, u64 zs, u64 p, u64 p2) {
    inv_bfly_asm<UT>(x, y, z, zs, p, p2);
}

// Half h of the N = 32768 stage-0 butterfly (native.rs:256-269) for one
// coefficient pair (a, c) = (x_j, x_{j + 16384}), a, c < 4p: with
// ar = a mod 2p and wc = lazy_mul_shoup(c) (< 2p), half 0 takes ar + wc and
// half 1 ar + 2p - wc (both < 4p). A workgroup that transforms one half of
// a row forms its own half of the first stage this way on load.
__device__ __forceinline__ u64 split_half(u64 a, u64 c, u32 h, u64 w, u64 ws, u64 p, u64 p2) {
    const u64 wc = lazy_mul_shoup_u(c, w, ws, p);
    const u64 ar = reduce1(a, p2);
    return h ? ar + p2 - wc : ar + wc;
}

// Final scaling of an inverse transform folded into its last stage:
// x' = (x + y) s and y' = (x - y) zeta s, with zs = zeta s; outputs reduced.
struct InvScale {
    u64 s, ss, zs, zss;
};

// Element index owned by thread t in slot k for a forward pass whose
// (virtual) first stage is S0V: stage s pairs bit LOGN-1-s.
template <int LOGN, int S0V, int R = 4>
__device__ __forceinline__ int fwd_index(int t, int k) {
    constexpr int RE = NttShape<LOGN, R>::RE;
    constexpr int LOB = LOGN - S0V - RE;
    return ((t >> LOB) << (LOGN - S0V)) | (k << LOB) | (t & ((1 << LOB) - 1));
}

// Element index for an inverse pass whose (virtual) first stage pairs bit B0V.
template <int LOGN, int B0V, int R = 4>
__device__ __forceinline__ int inv_index(int t, int k) {
    constexpr int RE = NttShape<LOGN, R>::RE;
    return ((t >> B0V) << (B0V + RE)) | (k << B0V) | (t & ((1 << B0V) - 1));
}

// Forward stages S0V .. S0V+RE-1, skipping stages below SMIN.
// tw: interleaved (omega, omega_shoup) pairs of the row's modulus.
template <int LOGN, int S0V, int SMIN, int R = 4>
__device__ __forceinline__ void fwd_pass(u64 (&x)[(NttShape<LOGN, R>::E)], int t, const ulonglong2* __restrict__ tw,
                                         u64 p, u64 p2) {
    constexpr int RE = NttShape<LOGN, R>::RE;
    constexpr int LOB = LOGN - S0V - RE;
    // In the first pass every thread uses the same twiddles (hi == 0 because
    // t < T = 2^LOB); spelling it out lets them live in scalar registers.
    const int hi = (S0V == 0) ? 0 : (t >> LOB);
#pragma unroll
    for (int u = 0; u < RE; u++) {
        const int s = S0V + u;
        if (s < SMIN) continue;
        const int half = 1 << (RE - 1 - u);
#pragma unroll
        for (int g = 0; g < (1 << u); g++) {
            const ulonglong2 w = tw[(1 << s) + (hi << u) + g];
#pragma unroll
            for (int q = 0; q < half; q++) {
                const int k0 = (g << (RE - u)) + q;
                fwd_bfly<S0V == 0>(x[k0], x[k0 + half], w.x, w.y, p, p2);
            }
        }
        // Keep the next stage's twiddle loads from being hoisted above this
        // one (bounds the live registers).
        if constexpr (RE >= 4) asm volatile("" ::: "memory");
    }
}

// Inverse stages pairing bits B0V .. B0V+RE-1, skipping bits below VMIN.
// tw: interleaved (zeta_inv, zeta_inv_shoup) pairs of the row's modulus.
// With SCALE, the last stage (bit LOGN-1, one group) applies `sc`.
template <int LOGN, int B0V, int VMIN, int R = 4, bool SCALE = false>
__device__ __forceinline__ void inv_pass(u64 (&x)[(NttShape<LOGN, R>::E)], int t, const ulonglong2* __restrict__ tw,
                                         u64 p, u64 p2, const InvScale& sc) {
    constexpr int RE = NttShape<LOGN, R>::RE;
    constexpr int N = 1 << LOGN;
    const int hi = t >> B0V;
#pragma unroll
    for (int u = 0; u < RE; u++) {
        const int v = B0V + u;
        if (v < VMIN) continue;
        const int half = 1 << u;
        if (SCALE && v == LOGN - 1) {
#pragma unroll
            for (int q = 0; q < half; q++) {
                const u64 a = x[q], b = x[q + half];
                x[q] = mul_shoup_u(a + b, sc.s, sc.ss, p);
                x[q + half] = mul_shoup_u(a + p2 - b, sc.zs, sc.zss, p);
            }
            continue;
        }
        const int offset = N - (N >> v);
#pragma unroll
        for (int g = 0; g < (1 << (RE - 1 - u)); g++) {
            const ulonglong2 z = tw[offset + (hi << (RE - 1 - u)) + g];
#pragma unroll
            for (int q = 0; q < half; q++) {
                const int k0 = (g << (u + 1)) + q;
                // In the last pass (B0V = LOGN - RE) hi is 0: shared twiddles.
                inv_bfly<B0V == LOGN - RE>(x[k0], x[k0 + half], z.x, z.y, p, p2);
            }
        }
        if constexpr (RE >= 4) asm volatile("" ::: "memory");
    }
}

template <int LOGN, int S0V, int R = 4>
__device__ __forceinline__ void lds_put_fwd(u64* lds, const u64 (&x)[(NttShape<LOGN, R>::E)], int t) {
#pragma unroll
    for (int k = 0; k < NttShape<LOGN, R>::E; k++) lds[lds_slot(fwd_index<LOGN, S0V, R>(t, k))] = x[k];
}
template <int LOGN, int S0V, int R = 4>
__device__ __forceinline__ void lds_get_fwd(const u64* lds, u64 (&x)[(NttShape<LOGN, R>::E)], int t) {
#pragma unroll
    for (int k = 0; k < NttShape<LOGN, R>::E; k++) x[k] = lds[lds_slot(fwd_index<LOGN, S0V, R>(t, k))];
}
template <int LOGN, int B0V, int R = 4>
__device__ __forceinline__ void lds_put_inv(u64* lds, const u64 (&x)[(NttShape<LOGN, R>::E)], int t) {
#pragma unroll
    for (int k = 0; k < NttShape<LOGN, R>::E; k++) lds[lds_slot(inv_index<LOGN, B0V, R>(t, k))] = x[k];
}
template <int LOGN, int B0V, int R = 4>
__device__ __forceinline__ void lds_get_inv(const u64* lds, u64 (&x)[(NttShape<LOGN, R>::E)], int t) {
#pragma unroll
    for (int k = 0; k < NttShape<LOGN, R>::E; k++) x[k] = lds[lds_slot(inv_index<LOGN, B0V, R>(t, k))];
}

// Exchange layouts in which each wave only touches the block of 64 * E
// consecutive elements it owns (elements j with j >> (6 + RE) equal to its
// wave index t >> 6). Moving data between two such layouts needs no
// workgroup barrier, only ordering within the wave: LDS instructions of one
// wave execute in program order, so a compiler fence at wavefront scope
// suffices.
template <int LOGN, int S0V, int R = 4>
constexpr bool fwd_local() {
    return LOGN - S0V - NttShape<LOGN, R>::RE <= 6;
}
template <int B0V>
constexpr bool inv_local() {
    return B0V <= 6;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Ordering between an exchange's LDS writes and reads (or reads and the next
// writes): a workgroup barrier unless every wave stays inside its own block.
template <bool LOCAL>
__device__ __forceinline__ void exchange_sync() {
    if constexpr (LOCAL)
        wave_sync();
    else
        __syncthreads();
}

constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// The last pass of either direction always starts at LOGN - RE, so that its
// element order is fwd_index<LOGN, LOGN-RE> (16 consecutive values per thread)
// or inv_index<LOGN, LOGN-RE> (value k of thread t at k * T + t).
template <int LOGN, int R = 4>
constexpr int last_start() {
    return LOGN - NttShape<LOGN, R>::RE;
}

template <int LOGN, int S0V, int R = 4>
__device__ __forceinline__ void fwd_rest(u64* lds, u64 (&x)[(NttShape<LOGN, R>::E)], int t, const ulonglong2* tw,
                                         u64 p, u64 p2) {
    constexpr int RE = NttShape<LOGN, R>::RE;
    // Called after the pass starting at S0V has run; continues with the next one.
    if constexpr (S0V + RE < LOGN) {
        constexpr int NEXT = (S0V + 2 * RE <= LOGN) ? S0V + RE : LOGN - RE;
        constexpr int SMIN = S0V + RE;
        // The written layout is S0V's, the read one NEXT's; the next writes
        // (the following exchange, or store_fwd_order) use NEXT's again.
        constexpr bool A = fwd_local<LOGN, S0V, R>(), B = fwd_local<LOGN, NEXT, R>();
        lds_put_fwd<LOGN, S0V, R>(lds, x, t);
        exchange_sync<A && B>();
        lds_get_fwd<LOGN, NEXT, R>(lds, x, t);
        exchange_sync<B>();
        fwd_pass<LOGN, NEXT, SMIN, R>(x, t, tw, p, p2);
        fwd_rest<LOGN, NEXT, R>(lds, x, t, tw, p, p2);
    }
}

template <int LOGN, int B0V, int R = 4, bool SCALE = false>
__device__ __forceinline__ void inv_rest(u64* lds, u64 (&x)[(NttShape<LOGN, R>::E)], int t, const ulonglong2* tw,
                                         u64 p, u64 p2, const InvScale& sc) {
    constexpr int RE = NttShape<LOGN, R>::RE;
    if constexpr (B0V + RE < LOGN) {
        // Bits below LW are paired inside one wave's block. When the next
        // full pass would straddle LW, a shorter wave-local pass up to LW
        // (layout LW - RE, stages from VMIN) keeps that exchange barrier-free,
        // provided it costs no extra pass.
        constexpr int LW = 6 + RE;
        constexpr int PLAIN = (B0V + 2 * RE <= LOGN) ? B0V + RE : LOGN - RE;
        constexpr bool SPLIT = B0V + RE < LW && LW < B0V + 2 * RE && LW - RE > B0V && LOGN > LW &&
                               1 + cdiv(LOGN - LW, RE) <= cdiv(LOGN - B0V - RE, RE);
        constexpr int NEXT = SPLIT ? LW - RE : PLAIN;
        constexpr int VMIN = B0V + RE;
        constexpr bool A = inv_local<B0V>(), B = inv_local<NEXT>();
        lds_put_inv<LOGN, B0V, R>(lds, x, t);
        exchange_sync<A && B>();
        lds_get_inv<LOGN, NEXT, R>(lds, x, t);
        // After the last exchange the LDS is not written again (the callers
        // store the result from registers), so its reads need no fence.
        if constexpr (NEXT + RE < LOGN) exchange_sync<B>();
        inv_pass<LOGN, NEXT, VMIN, R, SCALE>(x, t, tw, p, p2, sc);
        inv_rest<LOGN, NEXT, R, SCALE>(lds, x, t, tw, p, p2, sc);
    }
}

// Full forward transform on registers: input order fwd_index<LOGN,0>, output
// order fwd_index<LOGN, last_start<LOGN>()> (E consecutive values per
// thread), values < 4p.
template <int LOGN, int R = 4>
__device__ __forceinline__ void ntt_fwd_regs(u64* lds, u64 (&x)[(NttShape<LOGN, R>::E)], int t, const ulonglong2* tw,
                                             u64 p) {
    const u64 p2 = 2 * p;
    fwd_pass<LOGN, 0, 0, R>(x, t, tw, p, p2);
    fwd_rest<LOGN, 0, R>(lds, x, t, tw, p, p2);
}

// Full inverse transform on registers: input order inv_index<LOGN,0> (E
// consecutive values per thread), output order inv_index<LOGN,
// last_start<LOGN>()>. Its last LDS reads are not fenced: the caller must
// not write the LDS afterwards. The final multiplication (by N^-1, or by N^-1 2^64
// after Montgomery products) is folded into the last stage through `sc`;
// the outputs are reduced.
template <int LOGN, int R = 4>
__device__ __forceinline__ void ntt_inv_regs(u64* lds, u64 (&x)[(NttShape<LOGN, R>::E)], int t, const ulonglong2* tw,
                                             u64 p, const InvScale& sc) {
    const u64 p2 = 2 * p;
    inv_pass<LOGN, 0, 0, R, true>(x, t, tw, p, p2, sc);
    inv_rest<LOGN, 0, R, true>(lds, x, t, tw, p, p2, sc);
}

__device__ __forceinline__ InvScale inv_scale_n(const ModInfo& mi) {
    return InvScale{mi.n_inv, mi.n_inv_shoup, mi.last_zn, mi.last_zn_shoup};
}
__device__ __forceinline__ InvScale inv_scale_nr(const ModInfo& mi) {
    return InvScale{mi.n_inv_r, mi.n_inv_r_shoup, mi.last_zr, mi.last_zr_shoup};
}

// Lanes per wave that hold part of the row (a row of fewer than 64 * E
// values lives in one partial wave), and the first element of the block of
// LANES * E consecutive elements that the wave of thread t owns.
template <int LOGN, int R = 4>
struct WaveBlock {
    static constexpr int T = NttShape<LOGN, R>::T;
    static constexpr int LB = T < 64 ? LOGN - NttShape<LOGN, R>::RE : 6;  // log2 lanes
    static constexpr int RE = NttShape<LOGN, R>::RE;
    __device__ static int base(int t) { return (t >> LB) << (LB + RE); }
    __device__ static int lane(int t) { return t & ((1 << LB) - 1); }
};

// Coalesced global row -> registers in the inverse-transform input order
// (value e of thread t is element t*E + e). Each wave reads its own block of
// 64 * E elements with consecutive lanes on consecutive elements and
// transposes it through its own part of the LDS, so no wave waits for
// another: waves that have their data start transforming while others are
// still loading. Must be followed by the inverse transform (whose first
// exchange writes the same block).
template <int LOGN, int UNROLL = 4, int R = 4, class Load>
__device__ __forceinline__ void load_inv_order(u64* lds, u64 (&x)[(NttShape<LOGN, R>::E)], int t, Load load) {
    using S = NttShape<LOGN, R>;
    using W = WaveBlock<LOGN, R>;
    const int base = W::base(t), lane = W::lane(t);
    // Partial unrolling bounds the number of loads in flight (and registers).
#pragma unroll UNROLL
    for (int e = 0; e < S::E; e++) {
        const int j = base + (e << W::LB) + lane;
        lds[lds_slot(j)] = load(j);
    }
    wave_sync();
#pragma unroll
    for (int e = 0; e < S::E; e++) x[e] = lds[lds_slot((t << S::RE) + e)];
    wave_sync();
}

// Registers in the inverse-transform output order -> row stores: value e of
// thread t goes to element inv_index<LOGN, last_start>(t, e). The thread
// index passes through an empty asm so that the addresses are formed here
// and not hoisted to the kernel start, where they would stay live in
// registers across the whole transform.
template <int LOGN, int R = 4>
__device__ __forceinline__ void store_inv_order(u64* __restrict__ dst, const u64 (&x)[(NttShape<LOGN, R>::E)], int t) {
    int tl = t;
    asm volatile("" : "+v"(tl));
#pragma unroll
    for (int e = 0; e < NttShape<LOGN, R>::E; e++) dst[inv_index<LOGN, last_start<LOGN, R>(), R>(tl, e)] = x[e];
}

// Registers in the forward-transform output order (value e of thread t is
// element t*E + e) -> coalesced `store(j, value)`, transposed per wave
// through the wave's own block of the LDS (as in load_inv_order). Must follow
// a forward transform, whose last exchange reads the same block.
template <int LOGN, int R = 4, int UNROLL = 4, class Store>
__device__ __forceinline__ void store_fwd_order(u64* lds, const u64 (&x)[(NttShape<LOGN, R>::E)], int t, Store store) {
    using S = NttShape<LOGN, R>;
    using W = WaveBlock<LOGN, R>;
#pragma unroll
    for (int e = 0; e < S::E; e++) lds[lds_slot((t << S::RE) + e)] = x[e];
    wave_sync();
    const int base = W::base(t), lane = W::lane(t);
#pragma unroll UNROLL
    for (int e = 0; e < S::E; e++) {
        const int j = base + (e << W::LB) + lane;
        store(j, lds[lds_slot(j)]);
    }
}

}  // namespace fhehip
