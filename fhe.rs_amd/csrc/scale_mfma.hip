// RnsScaler::scale (math/src/rns/scaler.rs:249-352) on the matrix cores: the
// column sums are 8-bit integer MFMA products of host-built digit tables
// with the balanced base-256 digits of the column's residues (arithmetic and
// table layout in scale_mfma.h). The results are those of rns_scale_kernel,
// bit for bit.
//
// A wave takes 64 columns as two blocks of 32 (the N dimension of
// v_mfma_i32_32x32x16_i8 or v_mfma_i32_32x32x32_i8, KW = 16 or 32 k per
// MFMA). Per K step, lane l supplies the digits of KW / 16 residues of
// column l % 32 of each block as the B operand; the A operand (32 table
// rows) is the same for both blocks. After a tile's products, lanes l and
// l + 32 swap halves (v_permlane32_swap) so that every lane holds the 32
// rows of its own column, and each lane finishes one column: the rounding
// of v and w as rns_scale_column, then the outputs. The blocks are columns
// c .. c + 31 and c + 32 .. c + 63, or, for rows left as two open halves,
// c .. c + 31 and c + N/2 .. c + N/2 + 31 (PAIR: the stages pairing the
// halves run inside). A workgroup of 4 waves stages the tables in LDS once
// and covers `chunks` runs of 256 columns (128 pairs);
// grid = (N / (256 chunks), parts, instances).
#include <type_traits>

#include "bfv_launch.h"
#include "profiler.h"
#include "rns_device.h"
#include "scale_mfma.h"

namespace fhehip {

typedef int i32x16 __attribute__((ext_vector_type(16)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

constexpr u64 kDigitBias = 0x8080808080808080ull;

// The bytes of the result, read as signed 8-bit digits, sum to x times
// their powers of 256 (x below 2^64 - kDigitBias).
__device__ __forceinline__ u64 digits64(u64 x) { return (x + kDigitBias) ^ kDigitBias; }

// Sixteen digits of the 128-bit two's complement value (lo, hi), whose
// magnitude is far below 2^127: d0 holds digits 0..7, d1 digits 8..15.
__device__ __forceinline__ void digits128(u64 lo, u64 hi, u64& d0, u64& d1) {
    unsigned long long c;
    const u64 a = __builtin_addcll(lo, kDigitBias, 0, &c);
    d0 = a ^ kDigitBias;
    d1 = (hi + kDigitBias + c) ^ kDigitBias;
}

// The A or B operand of one lane: 8 digit bytes (kw = 16) or 16 (kw = 32).
template <int KW>
using Operand = typename std::conditional<KW == 16, long, i32x4>::type;

__device__ __forceinline__ void mfma(i32x16& c, long a, long b) {
    c = __builtin_amdgcn_mfma_i32_32x32x16_i8(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ void mfma(i32x16& c, const i32x4& a, const i32x4& b) {
    c = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}

// Sixteen digit bytes from two 8-digit words.
__device__ __forceinline__ i32x4 pack16(u64 lo, u64 hi) {
    i32x4 r;
    r[0] = (int)(u32)lo;
    r[1] = (int)(u32)(lo >> 32);
    r[2] = (int)(u32)hi;
    r[3] = (int)(u32)(hi >> 32);
    return r;
}

// The A operand of one lane from its table words at `w` (an LDS copy).
template <int KW>
__device__ __forceinline__ Operand<KW> table_operand(const u64* w) {
    if constexpr (KW == 16) {
        return (long)*w;
    } else {
        const ulonglong2 v = *(const ulonglong2*)w;
        return pack16(v.x, v.y);
    }
}

__device__ __forceinline__ i32x16 zero16() {
    i32x16 z;
#pragma unroll
    for (int r = 0; r < 16; r++) z[r] = 0;
    return z;
}

// Before: c0 (block 0) and c1 (block 1) in the MFMA layout, register r of
// lane l holding row 8 (r / 4) + 4 (l / 32) + r % 4 of column l % 32. After:
// lane l holds column l of the wave, row 8 (r / 4) + r % 4 in c0[r] and row
// 8 (r / 4) + 4 + r % 4 in c1[r].
__device__ __forceinline__ void own_column(i32x16& c0, i32x16& c1) {
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const auto p = __builtin_amdgcn_permlane32_swap((unsigned)c0[r], (unsigned)c1[r], false, false);
        c0[r] = (int)p[0];
        c1[r] = (int)p[1];
    }
}

// Row b of the lane's own column (after own_column).
template <int B>
__device__ __forceinline__ long long row_of(const i32x16& c0, const i32x16& c1) {
    return (B >> 2) & 1 ? c1[4 * (B >> 3) + (B & 3)] : c0[4 * (B >> 3) + (B & 3)];
}

// A 16-digit value known per column (lane l holds column l's, digits 0..7
// in d0, 8..15 in d1) as the B operands of one 16-deep K step for both
// column blocks: lanes below 32 supply d0, the others d1.
__device__ __forceinline__ void value_operands(u64 d0, u64 d1, long& b0, long& b1) {
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)d0, (unsigned)d1, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(d0 >> 32), (unsigned)(d1 >> 32), false, false);
    b0 = (long)((u64)lo[0] | ((u64)hi[0] << 32));
    b1 = (long)((u64)lo[1] | ((u64)hi[1] << 32));
}

// Two 16-digit values known per column (x at k < 16, y at 16 <= k < 32) as
// the B operands of one 32-deep K step for both column blocks: lanes below
// 32 supply x, the others y.
__device__ __forceinline__ void value_operands(const i32x4& x, const i32x4& y, i32x4& b0, i32x4& b1) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const auto p = __builtin_amdgcn_permlane32_swap((unsigned)x[i], (unsigned)y[i], false, false);
        b0[i] = (int)p[0];
        b1[i] = (int)p[1];
    }
}

// acc + x m, x a signed 32-bit value (one v_mad_i64_i32; the sign extension
// and the 64-bit add come with it). m is a power of 256 below 2^31.
__device__ __forceinline__ long long mad_i64(int x, int m, long long acc) {
    long long r;
    u64 ovf;
    asm("v_mad_i64_i32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(ovf) : "v"(x), "s"(m), "v"(acc));
    return r;
}

// acc + sum_{v < 4} P_{4u+v} 256^v over rows 4u .. 4u + 3 of the lane's column.
template <int U>
__device__ __forceinline__ long long word_rows(const i32x16& c0, const i32x16& c1, long long acc) {
    acc = mad_i64((int)row_of<4 * U>(c0, c1), 1, acc);
    acc = mad_i64((int)row_of<4 * U + 1>(c0, c1), 256, acc);
    acc = mad_i64((int)row_of<4 * U + 2>(c0, c1), 65536, acc);
    return mad_i64((int)row_of<4 * U + 3>(c0, c1), 16777216, acc);
}

// sum_b P_b 256^b over the 32 rows of the lane's column, modulo 2^256.
__device__ __forceinline__ U256 combine32(const i32x16& c0, const i32x16& c1) {
    u32 d[8];
    long long acc = word_rows<0>(c0, c1, 0);
    d[0] = (u32)acc;
    acc = word_rows<1>(c0, c1, acc >> 32);
    d[1] = (u32)acc;
    acc = word_rows<2>(c0, c1, acc >> 32);
    d[2] = (u32)acc;
    acc = word_rows<3>(c0, c1, acc >> 32);
    d[3] = (u32)acc;
    acc = word_rows<4>(c0, c1, acc >> 32);
    d[4] = (u32)acc;
    acc = word_rows<5>(c0, c1, acc >> 32);
    d[5] = (u32)acc;
    acc = word_rows<6>(c0, c1, acc >> 32);
    d[6] = (u32)acc;
    acc = word_rows<7>(c0, c1, acc >> 32);
    d[7] = (u32)acc;
    U256 r;
#pragma unroll
    for (int k = 0; k < 4; k++) r.w[k] = d[2 * k] | ((u64)d[2 * k + 1] << 32);
    return r;
}

// sc_round_v for a shift of 65 .. 128 (the table builder declines others):
// the shifted bits come from words 1 .. 3 of the sum.
__device__ __forceinline__ u128 round_v(const U256& sum, unsigned shift) {
    const unsigned bs = shift - 65;  // 0 .. 63
    const u64 r0 = bs ? (sum.w[1] >> bs) | (sum.w[2] << (64 - bs)) : sum.w[1];
    const u64 r1 = bs ? (sum.w[2] >> bs) | (sum.w[3] << (64 - bs)) : sum.w[2];
    return (((u128)r1 << 64 | r0) >> 1) + (r0 & 1);
}

// a w mod p in [0, 2p) for a below 2^32 (Shoup's product with ws =
// floor(w 2^64 / p), its quotient estimate below 2^32).
__device__ __forceinline__ u64 mul_shoup32(u32 a, u64 w, u64 ws, u64 p) {
    const u64 t = (u64)a * (u32)ws;
    const u32 q = (u32)(((u64)a * (u32)(ws >> 32) + (t >> 32)) >> 32);
    return (u64)a * w - (u64)q * p;
}

// The stages pairing two half-row transforms (PAIR): column block 1 is
// block 0 shifted by N/2, so lanes l and l + 32 hold the columns c and
// c + N/2 of a pair. isp (optional): per source limb j, words 6 j .. 6 j + 5
// = (w, w', z, z', n^-1, n^-1') of the last inverse stage, applied on load;
// osp (optional): per output i, words 6 i, 6 i + 1 = (w, w') of the first
// forward stage, applied on store.
struct PairStages {
    const u64* isp;
    const ModInfo* in_mods;
    const u64* osp;
};

// X = (x + y) n^-1 at c, Y = (x - y) zeta n^-1 at c + N/2 for x at c and y
// at c + N/2 below 2p (native.rs:288-300, 130-131); both reduced. w: the
// LDS copy of the limb's (p, zeta n^-1, its Shoup, n^-1, its Shoup).
__device__ __forceinline__ void pair_inverse(u64& x, u64& y, const u64* w) {
    const u64 p = w[0];
    const u64 a = x + y, b = x + 2 * p - y;
    x = mul_shoup(a, w[3], w[4], p);
    y = mul_shoup(b, w[1], w[2], p);
}

// One wave's 64 columns (the tables already in LDS): block 0 from col0,
// block 1 from col0 + off1 (32, or N/2 for a pair). Lane l supplies, per K
// step kappa, the residues j = (KW / 8) kappa + (KW / 16) (l / 32) .. +
// KW / 16 - 1 of column l % 32 of each block (zero from nf on).
template <int KW, int KS1, int NTO, bool HAS_W, bool PAIR>
__device__ __forceinline__ void rns_scale_mfma_cols(const u64* __restrict__ in, u64* __restrict__ out,
                                                    const u64* __restrict__ sc, const u64* tab, u32 n, int nf,
                                                    u32 col0, u32 off1, const PairStages& ps, u32 pc) {
    constexpr int WL = KW / 16, KS3 = KW == 32 ? 1 : (HAS_W ? 2 : 1);
    constexpr int T1 = 1 + (HAS_W ? 1 : 0) + NTO;
    constexpr int A2 = T1 * KS1 * 64 * WL, A3 = A2 + (HAS_W ? 64 * WL : 0), AM = A3 + NTO * KS3 * 64 * WL;
    const int lane = threadIdx.x & 63, h = lane >> 5;
    const u64* src = in + col0 + (lane & 31);
    Operand<KW> b0[KS1], b1[KS1];
#pragma unroll
    for (int k = 0; k < KS1; k++) {
        const int j = (KW / 8) * k + WL * h;
        u64 r[4] = {0, 0, 0, 0};
        if (j < nf) {
            r[0] = src[(size_t)j * n];
            r[1] = src[(size_t)j * n + off1];
            if (PAIR && ps.isp) pair_inverse(r[0], r[1], tab + pc + 5 * j);
        }
        if constexpr (KW == 16) {
            b0[k] = (long)digits64(r[0]);
            b1[k] = (long)digits64(r[1]);
        } else {
            if (j + 1 < nf) {
                r[2] = src[(size_t)(j + 1) * n];
                r[3] = src[(size_t)(j + 1) * n + off1];
                if (PAIR && ps.isp) pair_inverse(r[2], r[3], tab + pc + 5 * (j + 1));
            }
            b0[k] = pack16(digits64(r[0]), digits64(r[2]));
            b1[k] = pack16(digits64(r[1]), digits64(r[3]));
        }
    }
    const ScHeader<true> hd(sc);
    i32x16 c0, c1;
    auto pass1 = [&](int tile) {
        c0 = zero16();
        c1 = zero16();
#pragma unroll
        for (int k = 0; k < KS1; k++) {
            const Operand<KW> a = table_operand<KW>(tab + WL * ((tile * KS1 + k) * 64 + lane));
            mfma(c0, a, b0[k]);
            mfma(c1, a, b1[k]);
        }
    };
    // S -> v
    pass1(0);
    own_column(c0, c1);
    const u128 v = round_v(combine32(c0, c1), hd.shift);
    u64 v0, v1, w0 = 0, w1 = 0;
    digits128((u64)v, (u64)(v >> 64), v0, v1);
    // W (residue part, then v times theta_gamma) -> signed w
    if constexpr (HAS_W) {
        pass1(1);
        Operand<KW> vb0, vb1;
        if constexpr (KW == 16) value_operands(v0, v1, vb0, vb1);
        else value_operands(pack16(v0, v1), pack16(0, 0), vb0, vb1);
        const Operand<KW> a = table_operand<KW>(tab + A2 + WL * lane);
        mfma(c0, a, vb0);
        mfma(c1, a, vb1);
        own_column(c0, c1);
        bool w_sign;
        const u128 w = sc_round_w(combine32(c0, c1), w_sign);
        const u128 ws = w_sign ? (u128)0 - w : w;
        digits128((u64)ws, (u64)(ws >> 64), w0, w1);
    }
    // the B operands of v and w for the output tiles
    Operand<KW> ob0[KS3], ob1[KS3];
    if constexpr (KW == 16) {
        value_operands(v0, v1, ob0[0], ob1[0]);
        if constexpr (HAS_W) value_operands(w0, w1, ob0[KS3 - 1], ob1[KS3 - 1]);
    } else {
        value_operands(pack16(v0, v1), pack16(w0, w1), ob0[0], ob1[0]);
    }
    const int side = lane >> 5;
    u64* dst = out + col0 + (side ? off1 + lane - 32 : lane);
    const u64* outc = sc + sc_out_base(nf);
    const int count = hd.count;
#pragma unroll
    for (int t = 0; t < NTO; t++) {
        pass1(T1 - NTO + t);
#pragma unroll
        for (int k = 0; k < KS3; k++) {
            const Operand<KW> a = table_operand<KW>(tab + A3 + WL * ((t * KS3 + k) * 64 + lane));
            mfma(c0, a, ob0[k]);
            mfma(c1, a, ob1[k]);
        }
        own_column(c0, c1);
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int i = 4 * t + s;
            if (i >= count) break;
            // X = sum_t P_t 256^t (rows 8 s .. 8 s + 7), then X + M_i >= 0
            long long acc = mad_i64(c0[4 * s], 1, 0);
            acc = mad_i64(c0[4 * s + 1], 256, acc);
            acc = mad_i64(c0[4 * s + 2], 65536, acc);
            acc = mad_i64(c0[4 * s + 3], 16777216, acc);
            const u32 lo32 = (u32)acc;
            acc = mad_i64(c1[4 * s], 1, acc >> 32);
            acc = mad_i64(c1[4 * s + 1], 256, acc);
            acc = mad_i64(c1[4 * s + 2], 65536, acc);
            acc = mad_i64(c1[4 * s + 3], 16777216, acc);
            unsigned long long cy;
            const u64 lo = __builtin_addcll((u64)lo32 | ((u64)(u32)acc << 32), tab[AM + 2 * i], 0, &cy);
            const u32 hi = (u32)((u64)(acc >> 32) + tab[AM + 2 * i + 1] + cy);  // X + M < 2^80
            const u64* o = outc + (size_t)i * sc_out_stride(nf);
            const u64 q = o[0];
            const u64 x = lazy_reduce_shoup1<true>(lo, o[8], q) + mul_shoup32(hi, o[5], o[7], q);
            u64 y = reduce1(reduce1(x, 2 * q), q);
            if (PAIR && ps.osp) {
                // Harvey forward butterfly of reduced inputs (native.rs:256-269),
                // outputs < 4q: the lane at c + N/2 forms t = w y and sends it,
                // the lane at c sends its y
                const u64 t = lazy_mul_shoup_u(y, ps.osp[6 * i], ps.osp[6 * i + 1], q);
                const u64 send = side ? t : y;
                const auto lo2 = __builtin_amdgcn_permlane32_swap((unsigned)send, (unsigned)send, false, false);
                const auto hi2 =
                    __builtin_amdgcn_permlane32_swap((unsigned)(send >> 32), (unsigned)(send >> 32), false, false);
                const u64 got = side ? ((u64)hi2[0] << 32 | lo2[0]) : ((u64)hi2[1] << 32 | lo2[1]);
                y = side ? got + 2 * q - t : y + got;
            }
            dst[(size_t)i * n] = y;
        }
    }
}

template <int KW, int KS1, int NTO, bool HAS_W, bool PAIR>
__global__ __launch_bounds__(256) void rns_scale_mfma_kernel(const u64* __restrict__ in, size_t in_inst,
                                                             size_t in_part, u64* __restrict__ out, size_t out_inst,
                                                             size_t out_part, const u64* __restrict__ sc,
                                                             const u64* __restrict__ mt, u32 mt_words, u32 n,
                                                             u32 chunks, const u64* __restrict__ isp,
                                                             const ModInfo* __restrict__ in_mods,
                                                             const u64* __restrict__ osp) {
    extern __shared__ u64 tab[];
    const int nf = (int)sc[SC_NF];
    for (u32 i = threadIdx.x; i < mt_words; i += 256) tab[i] = mt[i];
    // the pair's inverse-stage constants per source limb after the tables
    if (PAIR && isp)
        for (u32 i = threadIdx.x; i < 5 * (u32)nf; i += 256) {
            const u32 j = i / 5, k = i % 5;
            tab[mt_words + i] = k == 0 ? in_mods[j].p : isp[6 * j + 1 + k];
        }
    __syncthreads();
    const u64* src = in + blockIdx.z * in_inst + blockIdx.y * in_part;
    u64* dst = out + blockIdx.z * out_inst + blockIdx.y * out_part;
    const PairStages ps{isp, in_mods, osp};
    // a run: 256 columns, or 128 pairs (c, c + N/2)
    constexpr u32 RUN = PAIR ? 128 : 256, WAVE = PAIR ? 32 : 64;
    const u32 off1 = PAIR ? n / 2 : 32;
    for (u32 ch = 0; ch < chunks; ch++)
        rns_scale_mfma_cols<KW, KS1, NTO, HAS_W, PAIR>(src, dst, sc, tab, n, nf,
                                                       (blockIdx.x * chunks + ch) * RUN + (threadIdx.x >> 6) * WAVE,
                                                       off1, ps, mt_words);
}

// One scaling launch: rows of part k of instance b at in + b in_inst +
// k in_part (likewise out); pair: the stages pairing half rows (PairStages).
struct MfmaLaunch {
    const u64 *in, *sc, *mt;
    size_t in_inst, in_part;
    u64* out;
    size_t out_inst, out_part;
    u32 words, n, chunks;
    bool pair;
    PairStages ps;
};

template <int KW, int KS1, int NTO, bool HAS_W>
static void mfma_go(dim3 grid, size_t lds, hipStream_t s, const MfmaLaunch& m) {
    if (m.pair)
        hipLaunchKernelGGL((rns_scale_mfma_kernel<KW, KS1, NTO, HAS_W, true>), grid, dim3(256), lds, s, m.in, m.in_inst,
                           m.in_part, m.out, m.out_inst, m.out_part, m.sc, m.mt, m.words, m.n, m.chunks, m.ps.isp,
                           m.ps.in_mods, m.ps.osp);
    else
        hipLaunchKernelGGL((rns_scale_mfma_kernel<KW, KS1, NTO, HAS_W, false>), grid, dim3(256), lds, s, m.in,
                           m.in_inst, m.in_part, m.out, m.out_inst, m.out_part, m.sc, m.mt, m.words, m.n, m.chunks,
                           nullptr, nullptr, nullptr);
}

template <int KW, int KS1>
static void mfma_ks(const ScaleMfmaLayout& lay, dim3 grid, size_t lds, hipStream_t s, const MfmaLaunch& m) {
    if (lay.nto == 1) {
        if (lay.has_w) mfma_go<KW, KS1, 1, true>(grid, lds, s, m);
        else mfma_go<KW, KS1, 1, false>(grid, lds, s, m);
    } else {
        if (lay.has_w) mfma_go<KW, KS1, 2, true>(grid, lds, s, m);
        else mfma_go<KW, KS1, 2, false>(grid, lds, s, m);
    }
}

static bool mfma_launch(MfmaLaunch m, const ScaleMfmaLayout& lay, u32 parts, u32 inst, hipStream_t s,
                        const char* name) {
    const u32 n = m.n;
    if (!m.mt || (lay.kw != 16 && lay.kw != 32) || lay.ks1 < 1 || lay.ks1 > (lay.kw == 16 ? 8 : 4) || lay.nto < 1 ||
        lay.nto > 2 || n < 256 || n % 256)
        return false;
    if (!parts || !inst) return true;
    ProfScope ps(name, 31 - __builtin_clz(n), s);
    // Each workgroup stages the tables once for up to 8 runs of 256 columns,
    // as long as the launch keeps at least 2048 workgroups.
    u32 chunks = 1;
    while (chunks < 8 && (n / 256) % (2 * chunks) == 0 && (size_t)(n / 256 / (2 * chunks)) * parts * inst >= 2048)
        chunks *= 2;
    m.chunks = chunks;
    m.words = (u32)lay.words();
    const dim3 grid(n / 256 / chunks, parts, inst);
    const size_t lds = (lay.words() + (m.pair && m.ps.isp ? 5 * (size_t)lay.nf : 0)) * sizeof(u64);
#define FHEHIP_MFMA_KS(W, K) \
    case K: mfma_ks<W, K>(lay, grid, lds, s, m); break;
    if (lay.kw == 16) {
        switch (lay.ks1) {
            FHEHIP_MFMA_KS(16, 1) FHEHIP_MFMA_KS(16, 2) FHEHIP_MFMA_KS(16, 3) FHEHIP_MFMA_KS(16, 4)
            FHEHIP_MFMA_KS(16, 5) FHEHIP_MFMA_KS(16, 6) FHEHIP_MFMA_KS(16, 7) FHEHIP_MFMA_KS(16, 8)
        }
    } else {
        switch (lay.ks1) { FHEHIP_MFMA_KS(32, 1) FHEHIP_MFMA_KS(32, 2) FHEHIP_MFMA_KS(32, 3) FHEHIP_MFMA_KS(32, 4) }
    }
#undef FHEHIP_MFMA_KS
    HIPCHK(hipGetLastError());
    return true;
}

bool launch_rns_scale_mfma(const u64* in, size_t in_inst, size_t in_part, u64* out, size_t out_inst,
                           size_t out_part, const u64* d_sc, const u64* d_mt, const ScaleMfmaLayout& lay, u32 n,
                           u32 parts, u32 inst, hipStream_t s, const char* name) {
    MfmaLaunch m{in, d_sc, d_mt, in_inst, in_part, out, out_inst, out_part, 0, n, 1, false, PairStages{}};
    return mfma_launch(m, lay, parts, inst, s, name);
}

bool launch_rns_scale_mfma_pair(const ScalePairArgs& a, const u64* d_mt, const ScaleMfmaLayout& lay, u32 parts,
                                u32 inst, hipStream_t s, const char* name) {
    if (a.in_split && !a.in_mods) throw Error(-1, "paired scaling: the inverse stage needs the source moduli");
    MfmaLaunch m{a.in, a.sc, d_mt, a.in_inst, a.in_part, a.out, a.out_inst, a.out_part, 0, a.n, 1, true,
                 PairStages{(const u64*)a.in_split, a.in_mods, (const u64*)a.out_split}};
    return mfma_launch(m, lay, parts, inst, s, name);
}

}  // namespace fhehip
