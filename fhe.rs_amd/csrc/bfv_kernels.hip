// Kernels of the BFV evaluation path: RNS scaling columns, INTT with optional
// Galois gather, fused tensor + INTT, key switching (digit lift + lazy NTT +
// Shoup multiply-accumulate), and fused switch-down + add + NTT.
#include <type_traits>

#include "bfv_launch.h"
#include "ntt_device.h"
#include "profiler.h"
#include "rns_device.h"
#include "tuning.h"

namespace fhehip {

// ---------------------------------------------------------------------------
// RnsScaler::scale over the columns of a batch (rq::Scaler::scale conversion
// loop, math/src/rq/scaler.rs:84-93). Power-basis rows of part k of instance b
// start at in + b * in_inst + k * in_part; outputs likewise.
// grid = (N / 256, parts, instances)
//
// LDS: the workgroup first copies the constant block (sc_words words) into
// LDS. A column reads about 2 + count constants per source modulus; through
// the scalar cache these arrive in dependent rounds of a few loads each,
// which a launch of a few waves per SIMD cannot hide (single calls). Large
// launches keep the scalar-cache form, whose constants cost no VGPRs.
// ---------------------------------------------------------------------------
template <int NF, bool LDS>
__global__ __launch_bounds__(256) void rns_scale_kernel(const u64* __restrict__ in, size_t in_inst, size_t in_part,
                                                        u64* __restrict__ out, size_t out_inst, size_t out_part,
                                                        const u64* __restrict__ sc, u32 n, u32 sc_words) {
    extern __shared__ u64 sc_lds[];
    const u32 col = blockIdx.x * 256 + threadIdx.x;
    const bool live = col < n;
    const u64* src = in + blockIdx.z * in_inst + blockIdx.y * in_part + col;
    u64* dst = out + blockIdx.z * out_inst + blockIdx.y * out_part + col;
    // the column's residues are in flight while the constants are staged
    u64 r[NF];
    if (live) {
#pragma unroll
        for (int j = 0; j < NF; j++) r[j] = src[(size_t)j * n];
    }
    if constexpr (LDS) {
        for (u32 i = threadIdx.x; i < sc_words; i += 256) sc_lds[i] = sc[i];
        __syncthreads();
    }
    if (!live) return;
    rns_scale_column<NF, !LDS>(
        LDS ? sc_lds : sc, [&](int j) { return r[j]; }, [&](int i, u64 v) { dst[(size_t)i * n] = v; });
}

template <int NF>
static void rns_scale_launch(dim3 grid, bool lds, hipStream_t s, const u64* in, size_t in_inst, size_t in_part,
                             u64* out, size_t out_inst, size_t out_part, const u64* d_sc, u32 n, u32 sc_words) {
    if (lds)
        hipLaunchKernelGGL((rns_scale_kernel<NF, true>), grid, dim3(256), (size_t)sc_words * 8, s, in, in_inst, in_part,
                           out, out_inst, out_part, d_sc, n, sc_words);
    else
        hipLaunchKernelGGL((rns_scale_kernel<NF, false>), grid, dim3(256), 0, s, in, in_inst, in_part, out, out_inst,
                           out_part, d_sc, n, sc_words);
}

void launch_rns_scale(const u64* in, size_t in_inst, size_t in_part, u64* out, size_t out_inst, size_t out_part,
                      const u64* d_sc, u32 sc_words, u32 nf, u32 n, u32 parts, u32 inst, hipStream_t s,
                      const char* name) {
    if (!parts || !inst) return;
    ProfScope ps(name, 31 - __builtin_clz(n), s);
    dim3 grid((n + 255) / 256, parts, inst);
    const bool lds = (size_t)grid.x * parts * inst <= kScaleLdsMaxBlocks && sc_words * 8 <= 64 * 1024;
#define FHEHIP_SC_CASE(K) \
    case K: rns_scale_launch<K>(grid, lds, s, in, in_inst, in_part, out, out_inst, out_part, d_sc, n, sc_words); break;
    switch (nf) {
        FHEHIP_SC_CASE(1) FHEHIP_SC_CASE(2) FHEHIP_SC_CASE(3) FHEHIP_SC_CASE(4) FHEHIP_SC_CASE(5)
        FHEHIP_SC_CASE(6) FHEHIP_SC_CASE(7) FHEHIP_SC_CASE(8) FHEHIP_SC_CASE(9) FHEHIP_SC_CASE(10)
        FHEHIP_SC_CASE(11) FHEHIP_SC_CASE(12) FHEHIP_SC_CASE(13) FHEHIP_SC_CASE(14) FHEHIP_SC_CASE(15)
        FHEHIP_SC_CASE(16) FHEHIP_SC_CASE(17) FHEHIP_SC_CASE(18) FHEHIP_SC_CASE(19) FHEHIP_SC_CASE(20)
        FHEHIP_SC_CASE(21) FHEHIP_SC_CASE(22) FHEHIP_SC_CASE(23) FHEHIP_SC_CASE(24) FHEHIP_SC_CASE(25)
        FHEHIP_SC_CASE(26) FHEHIP_SC_CASE(27) FHEHIP_SC_CASE(28) FHEHIP_SC_CASE(29) FHEHIP_SC_CASE(30)
        FHEHIP_SC_CASE(31) FHEHIP_SC_CASE(32)
        default: throw Error(-2, "more than 32 source moduli in one scaling");
    }
#undef FHEHIP_SC_CASE
    HIPCHK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Out-of-place inverse NTT of rows, with an optional NTT-domain gather
// (Poly::substitute in the Ntt representation, math/src/rq/mod.rs:324-346,
// as q[m] = p[perm[m]]). grid = (limbs, parts, instances).
// ---------------------------------------------------------------------------
// The gather and the plain copy are separate kernels (PERM), so that each
// gets its own register allocation.
template <int LOGN, bool PERM>
__device__ __forceinline__ void intt_rows_body(u64* lds, const InttRowsArgs& a) {
    using S = NttShape<LOGN>;
    const int t = threadIdx.x;
    const u32 l = blockIdx.x, k = blockIdx.y, b = blockIdx.z;
    const u64* src = a.src[k] + (size_t)b * a.src_inst_stride + ((size_t)l << LOGN);
    u64* dst = a.dst + (size_t)b * a.dst_inst_stride + (size_t)k * a.dst_part_stride + ((size_t)l << LOGN);
    const u32 limb = l + a.table_offset;
    const ulonglong2* tw = a.inv_tw + ((size_t)limb << LOGN);
    const ModInfo mi = a.mods[limb];
    u64 x[S::E];
    if constexpr (PERM) {
        const u32* perm = a.perm;
        load_inv_order<LOGN, 16>(lds, x, t, [&](int j) { return src[perm[j]]; });
    } else {
        load_inv_order<LOGN, 16>(lds, x, t, [&](int j) { return src[j]; });
    }
    ntt_inv_regs<LOGN>(lds, x, t, tw, mi.p, inv_scale_n(mi));
    store_inv_order<LOGN>(dst, x, t);
}

template <int LOGN>
__global__ __launch_bounds__(NttShape<LOGN>::T) void intt_rows_kernel(InttRowsArgs a) {
    __shared__ u64 lds[NttShape<LOGN>::LDS_WORDS];
    intt_rows_body<LOGN, false>(lds, a);
}

template <int LOGN>
__global__ __launch_bounds__(NttShape<LOGN>::T) void intt_rows_gather_kernel(InttRowsArgs a) {
    __shared__ u64 lds[NttShape<LOGN>::LDS_WORDS];
    intt_rows_body<LOGN, true>(lds, a);
}

// ---------------------------------------------------------------------------
// Tensor product fused with the inverse NTT (Multiplicator::multiply,
// fhe/src/bfv/ops/mul.rs:181-185 then the INTT of the down-scaler, rq/scaler.rs
// :68-78): c0 = a0 b0, c1 = a0 b1 + a1 b0, c2 = a1 b1 over the M limbs of the
// multiplication context. 1-D grid over (instance, limb, product), product fastest.
// ---------------------------------------------------------------------------
template <int LOGN, int R>
__global__ __launch_bounds__((NttShape<LOGN, R>::T)) void tensor_intt_kernel(TensorArgs a) {
    using S = NttShape<LOGN, R>;
    __shared__ u64 lds[S::LDS_WORDS];
    const int t = threadIdx.x;
    // The three products of one (limb, instance) read the same rows: they are
    // consecutive logical indices, so they run on one XCD.
    const u32 l = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const u32 c = l % 3, m = (l / 3) % a.M, b = l / (3 * a.M);
    const ModInfo mi = a.mods[m];
    // Row pointers of the four extended inputs a0, a1, b0, b1 at limb m.
    const u64* rows[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const u32 common = k < 2 ? a.common_lhs : a.common_rhs;
        if (m < common) {
            const u64* base = k < 2 ? a.lhs : a.rhs;
            rows[k] = base + (size_t)b * a.in_inst_stride + (size_t)(k & 1) * a.in_part_stride + ((size_t)m << LOGN);
        } else {
            rows[k] = a.ext + (size_t)b * a.ext_inst_stride + (size_t)k * a.ext_part_stride +
                      ((size_t)(m - common) << LOGN);
        }
    }
    // The products are Montgomery products (each carries a factor 2^-64); the
    // inverse transform is linear, so its final scaling by N^-1 * 2^64 returns
    // exactly INTT(c) (c0 = a0 b0, c1 = a0 b1 + a1 b0, c2 = a1 b1 mod q_m).
    // The product is elementwise, so it is formed in the coalesced order and
    // transposed through LDS into the transform's input order.
    const u64 p = mi.p, pn = mi.pinv_neg;
    u64 x[S::E];
    if (c == 1) {
        const u64 *ra0 = rows[0], *ra1 = rows[1], *rb0 = rows[2], *rb1 = rows[3];
        load_inv_order<LOGN, 2, R>(lds, x, t, [&](int j) {
            u64 lo, hi, l2, h2;
            mul_wide(ra0[j], rb1[j], lo, hi);
            mul_wide(ra1[j], rb0[j], l2, h2);
            add128(lo, hi, l2, h2);
            return mont_redc(lo, hi, p, pn);  // < 2p: a valid inverse-transform input
        });
    } else {
        const u64* ra = rows[c == 0 ? 0 : 1];
        const u64* rb = rows[c == 0 ? 2 : 3];
        load_inv_order<LOGN, 4, R>(lds, x, t, [&](int j) {
            u64 lo, hi;
            mul_wide(ra[j], rb[j], lo, hi);
            return mont_redc(lo, hi, p, pn);
        });
    }
    const ulonglong2* tw = a.inv_tw + ((size_t)m << LOGN);
    ntt_inv_regs<LOGN, R>(lds, x, t, tw, mi.p, inv_scale_nr(mi));
    u64* dst = a.out + (size_t)b * a.out_inst_stride + (size_t)c * a.out_part_stride + ((size_t)m << LOGN);
    store_inv_order<LOGN, R>(dst, x, t);
}

// ---------------------------------------------------------------------------
// Rows transformed as two halves of 2^LOGH (degree 32768 always; smaller
// degrees for small batches, so that a row spreads over two workgroups): the
// inverse transform of each half with the Galois gather or the tensor
// products formed on load, using the half tables (which leave out N^-1; the
// Montgomery factor 2^-64 is undone by their last stage as in
// tensor_intt_kernel). The last inverse stage, pairing the halves, follows
// as ntt_split_inv on the outputs.
// grid = (2 limbs, parts, instances) for the rows: blockIdx.x = 2 limb + half.
template <int LOGH, bool PERM>
__global__ __launch_bounds__(NttShape<LOGH>::T) void intt_rows_half_kernel(InttRowsArgs a) {
    using S = NttShape<LOGH>;
    __shared__ u64 lds[S::LDS_WORDS];
    const int t = threadIdx.x;
    const u32 l = blockIdx.x >> 1, h = blockIdx.x & 1, k = blockIdx.y, b = blockIdx.z;
    const u64* src = a.src[k] + (size_t)b * a.src_inst_stride + ((size_t)l << (LOGH + 1));
    u64* dst = a.dst + (size_t)b * a.dst_inst_stride + (size_t)k * a.dst_part_stride + ((size_t)l << (LOGH + 1)) +
               ((size_t)h << LOGH);
    const u32 row = 2 * (l + a.table_offset) + h;  // half-table row
    const ModInfo mi = a.mods[row];
    u64 x[S::E];
    const u32 off = h << LOGH;
    if constexpr (PERM) {
        const u32* perm = a.perm;
        load_inv_order<LOGH, 16>(lds, x, t, [&](int j) { return src[perm[off + j]]; });
    } else {
        load_inv_order<LOGH, 16>(lds, x, t, [&](int j) { return src[off + j]; });
    }
    ntt_inv_regs<LOGH>(lds, x, t, a.inv_tw + ((size_t)row << LOGH), mi.p, inv_scale_n(mi));
    store_inv_order<LOGH>(dst, x, t);
}

// 1-D grid over (instance, limb, half, product), product fastest.
template <int LOGH>
__global__ __launch_bounds__(NttShape<LOGH>::T) void tensor_intt_half_kernel(TensorArgs a) {
    using S = NttShape<LOGH>;
    __shared__ u64 lds[S::LDS_WORDS];
    const int t = threadIdx.x;
    const u32 l = xcd_block(blockIdx.x, gridDim.x);
    const u32 c = l % 3, h = (l / 3) & 1, m = (l / 6) % a.M, b = l / (6 * a.M);
    const ModInfo mi = a.mods[2 * m + h];  // half-table constants
    const u32 off = h << LOGH;
    const u64* rows[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const u32 common = k < 2 ? a.common_lhs : a.common_rhs;
        if (m < common) {
            const u64* base = k < 2 ? a.lhs : a.rhs;
            rows[k] = base + (size_t)b * a.in_inst_stride + (size_t)(k & 1) * a.in_part_stride + ((size_t)m << (LOGH + 1)) + off;
        } else {
            rows[k] = a.ext + (size_t)b * a.ext_inst_stride + (size_t)k * a.ext_part_stride +
                      ((size_t)(m - common) << (LOGH + 1)) + off;
        }
    }
    const u64 p = mi.p, pn = mi.pinv_neg;
    u64 x[S::E];
    if (c == 1) {
        const u64 *ra0 = rows[0], *ra1 = rows[1], *rb0 = rows[2], *rb1 = rows[3];
        load_inv_order<LOGH, 2>(lds, x, t, [&](int j) {
            u64 lo, hi, l2, h2;
            mul_wide(ra0[j], rb1[j], lo, hi);
            mul_wide(ra1[j], rb0[j], l2, h2);
            add128(lo, hi, l2, h2);
            return mont_redc(lo, hi, p, pn);
        });
    } else {
        const u64* ra = rows[c == 0 ? 0 : 1];
        const u64* rb = rows[c == 0 ? 2 : 3];
        load_inv_order<LOGH>(lds, x, t, [&](int j) {
            u64 lo, hi;
            mul_wide(ra[j], rb[j], lo, hi);
            return mont_redc(lo, hi, p, pn);
        });
    }
    ntt_inv_regs<LOGH>(lds, x, t, a.inv_tw + ((size_t)(2 * m + h) << LOGH), p, inv_scale_nr(mi));
    u64* dst = a.out + (size_t)b * a.out_inst_stride + (size_t)c * a.out_part_stride + ((size_t)m << (LOGH + 1)) + off;
    store_inv_order<LOGH>(dst, x, t);
}

// ---------------------------------------------------------------------------
// Key switching (KeySwitchingKey::key_switch, key_switching_key.rs:214-241) for
// one key limb j of one instance: for every digit i, the power-basis row i of
// the input is reduced modulo q'_j (lazy_reduce_vec), forward transformed
// lazily (forward_vt_lazy; together the digit lift of rq/mod.rs:518-540) and
// multiplied by the key rows (Shoup, exact) into two accumulators, so `acc`
// receives the NTT-domain (c0, c1) of the key switch: rows acc + inst *
// acc_inst_stride + {0, 1} * acc_part_stride + j * N. grid = (L', instances).
// ---------------------------------------------------------------------------
// Digit lift and lazy forward NTT of key switching (rq/mod.rs:518-540 inside
// key_switching_key.rs:227-239): Y[b][i][j] = forward_vt_lazy(
// lazy_reduce_{q'_j}(p_i)) for ciphertext limb (digit) i and key limb j,
// stored < 2 q'_j. Each transformed digit is written once; the key inner
// product reads it back in the kernels below. 1-D grid over
// (instance, digit, key limb), key limb fastest.
// ---------------------------------------------------------------------------
template <int LOGN, int R>
__global__ __launch_bounds__((NttShape<LOGN, R>::T)) void lift_ntt_kernel(KeySwitchArgs a, u64* __restrict__ y) {
    using S = NttShape<LOGN, R>;
    __shared__ u64 lds[S::LDS_WORDS];
    const int t = threadIdx.x;
    // The key limbs of one digit read the same source row: consecutive
    // logical indices, run on one XCD.
    const u32 l = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const u32 j = l % a.key_limbs, i = (l / a.key_limbs) % a.digits, b = l / (a.key_limbs * a.digits);
    const ModInfo mi = a.key_mods[j];
    const u64 p = mi.p;
    const u64* src = a.src + (size_t)b * a.src_inst_stride + ((size_t)(a.log_base ? 0 : i) << LOGN);
    u64 x[S::E];
    // Digit i is below q_i (the ciphertext moduli are the first key moduli).
    // Below 4 q'_j it is already a valid input of the lazy forward transform:
    // the stored outputs are then other representatives (< 2 q'_j) of the
    // same residues, which the exact key inner product reduces. A
    // decomposition digit (below 2^31 < q'_j) always is.
    if (a.log_base) {
        const u32 sh = i * a.log_base;
        const u64 mask = ((u64)1 << a.log_base) - 1;
#pragma unroll
        for (int e = 0; e < S::E; e++) x[e] = (src[fwd_index<LOGN, 0, R>(t, e)] >> sh) & mask;
    } else if (a.key_mods[i].p <= 4 * p) {
#pragma unroll
        for (int e = 0; e < S::E; e++) x[e] = src[fwd_index<LOGN, 0, R>(t, e)];
    } else {
#pragma unroll
        for (int e = 0; e < S::E; e++)
            x[e] = lazy_reduce_u64(src[fwd_index<LOGN, 0, R>(t, e)], mi.barrett_lo, mi.barrett_hi, p);
    }
    ntt_fwd_regs<LOGN, R>(lds, x, t, a.key_fwd_tw + ((size_t)j << LOGN), p);
    u64* dst = y + ((((size_t)b * a.digits + i) * a.key_limbs + j) << LOGN);
    const u64 p2 = 2 * p;
    store_fwd_order<LOGN, R>(lds, x, t, [&](int k, u64 v) { dst[k] = reduce1(v, p2); });
}

// ---------------------------------------------------------------------------
// Switch-down of the key-switch output to the ciphertext context (Poly::
// switch_down repeated, math/src/rq/mod.rs:390-462), plus an optional
// power-basis addend, forward NTT, plus an optional NTT-domain gathered
// addend (the substituted c0 of a Galois rotation). Output NTT rows are
// reduced. grid = (L, parts, instances).
// ---------------------------------------------------------------------------
template <int LOGN>
__global__ __launch_bounds__(NttShape<LOGN>::T) void sd_add_ntt_kernel(SdNttArgs a) {
    using S = NttShape<LOGN>;
    __shared__ u64 lds[S::LDS_WORDS];
    const int t = threadIdx.x;
    const u32 i = blockIdx.x, part = blockIdx.y, b = blockIdx.z;
    const ModInfo mi = a.mods[i];
    const u64 p = mi.p;
    const u64* ks = a.ks + (size_t)b * a.ks_inst_stride + (size_t)part * a.ks_part_stride;
    u64 x[S::E];
#pragma unroll
    for (int e = 0; e < S::E; e++) x[e] = ks[((size_t)i << LOGN) + fwd_index<LOGN, 0>(t, e)];
    // Drop the extra limbs L'-1, ..., L, one at a time (last first).
    for (int s = (int)a.extra - 1; s >= 0; s--) {
        const u32 last = a.L + s;  // limb being dropped
        const u64 ql = a.extra_mods[s];
        const u64 half = ql >> 1;
        const u64* lrow = ks + ((size_t)last << LOGN);
        // Limbs L..last-1 are also modified by earlier drops; recompute their
        // chain for the column before using them.
        u64 y[S::E];
#pragma unroll
        for (int e = 0; e < S::E; e++) y[e] = lrow[fwd_index<LOGN, 0>(t, e)];
        for (int s2 = (int)a.extra - 1; s2 > s; s2--) {
            // apply the drop of limb L+s2 to limb `last`
            const u64 ql2 = a.extra_mods[s2];
            const u64 half2 = ql2 >> 1;
            const u64* l2 = ks + ((size_t)(a.L + s2) << LOGN);
            const SdConst cst = a.sd[s2 * a.sd_stride + last];
#pragma unroll
            for (int e = 0; e < S::E; e++) {
                u64 z = l2[fwd_index<LOGN, 0>(t, e)];
                // (chains deeper than two extra limbs are rejected on the host)
                z = z + half2;
                z = z >= ql2 ? z - ql2 : z;
                u64 tmp = lazy_reduce_u64(z, cst.bl, cst.bh, cst.q) + cst.half_mod;
                y[e] = mul_shoup_u(y[e] + 3 * cst.q - tmp, cst.inv, cst.inv_shoup, cst.q);
            }
        }
        const SdConst cst = a.sd[s * a.sd_stride + i];
#pragma unroll
        for (int e = 0; e < S::E; e++) {
            u64 z = y[e] + half;
            z = z >= ql ? z - ql : z;
            u64 tmp = lazy_reduce_u64(z, mi.barrett_lo, mi.barrett_hi, p) + cst.half_mod;
            x[e] = mul_shoup_u(x[e] + 3 * p - tmp, cst.inv, cst.inv_shoup, p);
        }
    }
    if (a.add_pb) {
        const u64* ad = a.add_pb + (size_t)b * a.add_inst_stride + (size_t)part * a.add_part_stride +
                        ((size_t)i << LOGN);
#pragma unroll
        for (int e = 0; e < S::E; e++) x[e] = add_mod(x[e], ad[fwd_index<LOGN, 0>(t, e)], p);
    }
    ntt_fwd_regs<LOGN>(lds, x, t, a.fwd_tw + ((size_t)i << LOGN), p);
    const u64 p2 = 2 * p;
    u64* dst = a.out + (size_t)b * a.out_inst_stride + (size_t)part * a.out_part_stride + ((size_t)i << LOGN);
    if (a.add_ntt && part < a.add_ntt_parts) {
        const u64* g = a.add_ntt + (size_t)b * a.add_ntt_inst_stride + (size_t)part * a.add_ntt_part_stride +
                       ((size_t)i << LOGN);
        const u32* perm = a.perm;
        store_fwd_order<LOGN, 4, 16>(lds, x, t, [&](int j, u64 v) {
            dst[j] = add_mod(reduce1(reduce1(v, p2), p), g[perm ? perm[j] : (u32)j], p);
        });
    } else {
        store_fwd_order<LOGN>(lds, x, t, [&](int j, u64 v) { dst[j] = reduce1(reduce1(v, p2), p); });
    }
}

// ---------------------------------------------------------------------------
// Key-switch finish for a key context at most one modulus above the ciphertext
// context. switch_down (math/src/rq/mod.rs:390-449) maps limb i to
//   q_l^-1 (x_i - [y]_i) mod q_i,  y = (x_l + floor(q_l/2)) mod q_l,
// [y]_i = lazy_reduce(y) - floor(q_l/2) mod q_i,
// which is linear in x_i. Only the dropped limb l must be in power basis
// (`special`: ks_mac_kernel, then an inverse NTT); the other limbs stay in
// the NTT domain:
//   out_i = NTT(add_pb_i - q_l^-1 [y]_i) + q_l^-1 x_i  (+ NTT-domain addend),
// where x_i = sum_d Y[d][i] * key[d][i] is the key inner product. With no
// extra modulus, out_i = NTT(add_pb_i) + x_i. Two kernels: ks_mac_kernel
// writes q_l^-1 x_i (+ addend) into `out`, elementwise; ks_finish_kernel
// adds the transformed power-basis part.
// ---------------------------------------------------------------------------

// One thread per coefficient of key limb i, both parts, for KS_MAC_INST
// (2, 4 or 8, chosen by the autotuner) instances (key_switching_key.rs:227-239): the key values of a digit are
// loaded once and used for every instance of the group. The switch-down
// factor q_l^-1 is folded into the key rows of the limbs below the dropped
// one (fin0, fin1), Y < 2q and the key rows < q, so seven products stay below
// 2^128; their exact sum (column accumulator) gets one reduction.
// Limb i == L (the dropped one, when extra) goes to `special`.
// grid = (N / 256, L + extra, ceil(instances / KS_MAC_INST))
template <u32 KS_MAC_INST>
__global__ __launch_bounds__(256) void ks_mac_kernel(KsFinishArgs f, u32 logn, u32 inst) {
    const SdNttArgs& a = f.sd;
    const u32 j = blockIdx.x * 256 + threadIdx.x;
    if (j >> logn) return;
    const u32 i = blockIdx.y, b0 = blockIdx.z * KS_MAC_INST;
    const u32 nb = inst - b0 < KS_MAC_INST ? inst - b0 : KS_MAC_INST;
    const bool dropped = i == a.L;
    const ModInfo mi = dropped ? *f.special_mod : a.mods[i];
    const u64 p = mi.p;
    const size_t ystride = (size_t)f.key_limbs << logn, kstride = (size_t)(a.L + a.extra) << logn;
    const size_t yinst = (size_t)f.digits * f.key_limbs << logn;
    const u64* yb = f.y + (((size_t)b0 * f.digits * f.key_limbs + i) << logn) + j;
    const u64* k0 = f.fin0 + ((size_t)i << logn) + j;
    const u64* k1 = f.fin1 + ((size_t)i << logn) + j;
    u64 m0[KS_MAC_INST], m1[KS_MAC_INST];
#pragma unroll
    for (u32 q = 0; q < KS_MAC_INST; q++) m0[q] = m1[q] = 0;
    // Groups of up to 7 digits: each product is below 2q * q < 2^125, so a
    // group's exact sum stays below 2^128 (two-column accumulator) and gets
    // one exact reduction.
    for (u32 d0 = 0; d0 < f.digits; d0 += 7) {
        const u32 dn = f.digits - d0 < 7 ? f.digits - d0 : 7;
        u64 kv0[7], kv1[7];
#pragma unroll
        for (u32 d = 0; d < 7; d++) {
            kv0[d] = d < dn ? k0[(d0 + d) * kstride] : 0;
            kv1[d] = d < dn ? k1[(d0 + d) * kstride] : 0;
        }
#pragma unroll
        for (u32 q = 0; q < KS_MAC_INST; q++) {
            if (q < nb) {
                Acc2 s0, s1;
                const u64 y0 = yb[q * yinst + d0 * ystride];
                acc2_first<false>(s0, y0, kv0[0]);
                acc2_first<false>(s1, y0, kv1[0]);
#pragma unroll
                for (u32 d = 1; d < 7; d++) {
                    if (d < dn) {
                        const u64 yv = yb[q * yinst + (d0 + d) * ystride];
                        acc2_mul<false>(s0, yv, kv0[d]);
                        acc2_mul<false>(s1, yv, kv1[d]);
                    }
                }
                u64 l0, h0, l1, h1;
                acc2_value(s0, l0, h0);
                acc2_value(s1, l1, h1);
                m0[q] = add_mod(m0[q], reduce128_shoup<true>(l0, h0, mi.r64, mi.r64_shoup, mi.w1s, p), p);
                m1[q] = add_mod(m1[q], reduce128_shoup<true>(l1, h1, mi.r64, mi.r64_shoup, mi.w1s, p), p);
            }
        }
    }
    const u32 src = a.perm ? a.perm[j] : j;
#pragma unroll
    for (u32 q = 0; q < KS_MAC_INST; q++) {
        if (q >= nb) break;
        const u32 b = b0 + q;
        if (dropped) {
            f.special[(((size_t)b * 2) << logn) + j] = m0[q];
            f.special[(((size_t)b * 2 + 1) << logn) + j] = m1[q];
            continue;
        }
        u64* out = a.out + (size_t)b * a.out_inst_stride + ((size_t)i << logn) + j;
        u64 m[2] = {m0[q], m1[q]};
#pragma unroll
        for (u32 part = 0; part < 2; part++) {
            if (a.add_ntt && part < a.add_ntt_parts)
                m[part] = add_mod(m[part], a.add_ntt[(size_t)b * a.add_ntt_inst_stride +
                                                     (size_t)part * a.add_ntt_part_stride + ((size_t)i << logn) + src],
                                  p);
            out[(size_t)part * a.out_part_stride] = m[part];
        }
    }
}

// out_i += NTT(add_pb_i - q_l^-1 [y]_i) (or NTT(add_pb_i) without an extra
// modulus). 1-D grid over (instance, part, limb), limb fastest: the limbs of
// one (instance, part) read the same dropped-limb row and run on one XCD.
template <int LOGN, int R>
__global__ __launch_bounds__((NttShape<LOGN, R>::T)) void ks_finish_kernel(KsFinishArgs f) {
    using S = NttShape<LOGN, R>;
    __shared__ u64 lds[S::LDS_WORDS];
    const SdNttArgs& a = f.sd;
    const int t = threadIdx.x;
    const u32 l = f.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const u32 i = l % a.L, part = (l / a.L) & 1, b = l / (2 * a.L);
    const ModInfo mi = a.mods[i];
    const u64 p = mi.p, p2 = 2 * p;
    const u64* ad = a.add_pb ? a.add_pb + (size_t)b * a.add_inst_stride + (size_t)part * a.add_part_stride +
                                   ((size_t)i << LOGN)
                             : nullptr;
    u64 x[S::E];
    if (a.extra) {
        const SdConst cst = a.sd[i];
        const u64 ql = a.extra_mods[0];
        const u64 half = ql >> 1;
        const u64* lrow = f.special + (((size_t)b * 2 + part) << LOGN);
        auto body = [&](auto small) {
#pragma unroll
            for (int e = 0; e < S::E; e++) {
                const int j = fwd_index<LOGN, 0, R>(t, e);
                u64 yv = lrow[j] + half;
                yv = yv >= ql ? yv - ql : yv;
                // [y]_i only has to be < 2 q_i and congruent to y (the rest
                // is exact modulo q_i and the output is reduced): y < q_l <=
                // 2 q_i already is.
                const u64 yr = decltype(small)::value ? yv : lazy_reduce_u64(yv, mi.barrett_lo, mi.barrett_hi, p);
                const u64 tmp = yr + cst.half_mod;
                u64 z = mul_shoup_u(3 * p - tmp, cst.inv, cst.inv_shoup, p);
                if (ad) z = add_mod(z, ad[j], p);
                x[e] = z;
            }
        };
        if (ql <= 2 * p)
            body(std::true_type{});
        else
            body(std::false_type{});
    } else {
#pragma unroll
        for (int e = 0; e < S::E; e++) x[e] = ad ? ad[fwd_index<LOGN, 0, R>(t, e)] : 0;
    }
    ntt_fwd_regs<LOGN, R>(lds, x, t, a.fwd_tw + ((size_t)i << LOGN), p);
    u64* dst = a.out + (size_t)b * a.out_inst_stride + (size_t)part * a.out_part_stride + ((size_t)i << LOGN);
    store_fwd_order<LOGN, R, 16>(lds, x, t, [&](int j, u64 v) { dst[j] = add_mod(reduce1(reduce1(v, p2), p), dst[j], p); });
}

// The same finish with rows transformed as two halves of 2^LOGH: the
// workgroup of half h forms the power-basis part z (as ks_finish_kernel)
// at both coefficients j and j + 2^LOGH of each pair it owns, takes half h of
// the stage-0 butterfly on load (native.rs:256-269), runs the remaining 14
// stages with the half tables and adds the reduced result into `out`. 1-D
// grid over (instance, part, limb, half), half fastest.
//
// pre != null: the rows after their first stage come ready from
// ks_z_pair_kernel as pre [inst][2][L][N], and are only loaded.
template <int LOGH>
__global__ __launch_bounds__(NttShape<LOGH>::T) void ks_finish_half_kernel(KsFinishArgs f,
                                                                        const ulonglong2* __restrict__ half_tw,
                                                                        const ulonglong2* __restrict__ split,
                                                                        const u64* __restrict__ pre) {
    using S = NttShape<LOGH>;
    __shared__ u64 lds[S::LDS_WORDS];
    const SdNttArgs& a = f.sd;
    const int t = threadIdx.x;
    const u32 l = xcd_block(blockIdx.x, gridDim.x);
    const u32 h = l & 1, r = l >> 1;
    const u32 i = r % a.L, part = (r / a.L) & 1, b = r / (2 * a.L);
    const ModInfo mi = a.mods[i];
    const u64 p = mi.p, p2 = 2 * p;
    const ulonglong2 w = split[3 * i];
    const u64* ad = a.add_pb ? a.add_pb + (size_t)b * a.add_inst_stride + (size_t)part * a.add_part_stride +
                                   ((size_t)i << (LOGH + 1))
                             : nullptr;
    const u64* lrow = a.extra ? f.special + (((size_t)b * 2 + part) << (LOGH + 1)) : nullptr;
    const SdConst cst = a.sd[i];
    const u64 ql = a.extra ? a.extra_mods[0] : 0, half = ql >> 1;
    // z = add_pb - q_l^-1 [y]_i (add_pb alone without an extra modulus); see
    // ks_finish_kernel for the representatives.
    auto z_at = [&](int k) {
        u64 z = 0;
        if (lrow) {
            u64 yv = lrow[k] + half;
            yv = yv >= ql ? yv - ql : yv;
            const u64 yr = ql <= 2 * p ? yv : lazy_reduce_u64(yv, mi.barrett_lo, mi.barrett_hi, p);
            z = mul_shoup_u(3 * p - (yr + cst.half_mod), cst.inv, cst.inv_shoup, p);
        }
        return ad ? add_mod(z, ad[k], p) : z;
    };
    u64 x[S::E];
    if (pre) {
        const u64* xr = pre + ((((size_t)b * 2 + part) * a.L + i) << (LOGH + 1)) + ((size_t)h << LOGH);
#pragma unroll
        for (int e = 0; e < S::E; e++) x[e] = xr[fwd_index<LOGH, 0>(t, e)];
    } else {
#pragma unroll
        for (int e = 0; e < S::E; e++) {
            const int k = fwd_index<LOGH, 0>(t, e);
            x[e] = split_half(z_at(k), z_at(k + (1 << LOGH)), h, w.x, w.y, p, p2);
        }
    }
    ntt_fwd_regs<LOGH>(lds, x, t, half_tw + ((size_t)(2 * i + h) << LOGH), p);
    u64* dst = a.out + (size_t)b * a.out_inst_stride + (size_t)part * a.out_part_stride + ((size_t)i << (LOGH + 1)) +
               ((size_t)h << LOGH);
    store_fwd_order<LOGH, 4, 16>(lds, x, t,
                               [&](int k, u64 v) { dst[k] = add_mod(reduce1(reduce1(v, p2), p), dst[k], p); });
}

// ---------------------------------------------------------------------------
// Ciphertext addition (fhe/src/bfv/ops/mod.rs:15-69 -> Poly AddAssign,
// math/src/rq/ops.rs:92-118): out = a + b mod q_limb, elementwise over
// [rows][N] where row r has modulus r % L.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void add_rows_kernel(const u64* __restrict__ a, const u64* __restrict__ b,
                                                       u64* __restrict__ out, const ModInfo* __restrict__ mods,
                                                       u32 L, u32 logn, size_t total) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const u64 p = mods[(i >> logn) % L].p;
    out[i] = add_mod(a[i], b[i], p);
}

void launch_add_rows(const u64* a, const u64* b, u64* out, const ModInfo* mods, u32 L, int logn, size_t rows,
                     hipStream_t s) {
    const size_t total = rows << logn;
    if (!total) return;
    ProfScope ps("add_rows", logn, s);
    hipLaunchKernelGGL(add_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a, b, out, mods, L,
                       (u32)logn, total);
    HIPCHK(hipGetLastError());
}

// ---------------------------------------------------------------------------
#define FHEHIP_LOGN_DISPATCH(logn, KERNEL, grid, block, stream, ...)                         \
    switch (logn) {                                                                          \
        case 3: hipLaunchKernelGGL(KERNEL<3>, grid, dim3(NttShape<3>::T), 0, stream, __VA_ARGS__); break; \
        case 4: hipLaunchKernelGGL(KERNEL<4>, grid, dim3(NttShape<4>::T), 0, stream, __VA_ARGS__); break; \
        case 5: hipLaunchKernelGGL(KERNEL<5>, grid, dim3(NttShape<5>::T), 0, stream, __VA_ARGS__); break; \
        case 6: hipLaunchKernelGGL(KERNEL<6>, grid, dim3(NttShape<6>::T), 0, stream, __VA_ARGS__); break; \
        case 7: hipLaunchKernelGGL(KERNEL<7>, grid, dim3(NttShape<7>::T), 0, stream, __VA_ARGS__); break; \
        case 8: hipLaunchKernelGGL(KERNEL<8>, grid, dim3(NttShape<8>::T), 0, stream, __VA_ARGS__); break; \
        case 9: hipLaunchKernelGGL(KERNEL<9>, grid, dim3(NttShape<9>::T), 0, stream, __VA_ARGS__); break; \
        case 10: hipLaunchKernelGGL(KERNEL<10>, grid, dim3(NttShape<10>::T), 0, stream, __VA_ARGS__); break; \
        case 11: hipLaunchKernelGGL(KERNEL<11>, grid, dim3(NttShape<11>::T), 0, stream, __VA_ARGS__); break; \
        case 12: hipLaunchKernelGGL(KERNEL<12>, grid, dim3(NttShape<12>::T), 0, stream, __VA_ARGS__); break; \
        case 13: hipLaunchKernelGGL(KERNEL<13>, grid, dim3(NttShape<13>::T), 0, stream, __VA_ARGS__); break; \
        case 14: hipLaunchKernelGGL(KERNEL<14>, grid, dim3(NttShape<14>::T), 0, stream, __VA_ARGS__); break; \
        default: throw Error(-2, "degree not supported");                                    \
    }                                                                                        \
    HIPCHK(hipGetLastError());

// The fused kernels with a tunable number of values per thread: RE = 5 (32
// values, N / 32 threads) from degree 32 up, else RE = 4.
#define FHEHIP_LOGN_R_CASE(L, R, KERNEL, grid, stream, ...) \
    case L: hipLaunchKernelGGL((KERNEL<L, R>), grid, dim3(NttShape<L, R>::T), 0, stream, __VA_ARGS__); break;
#define FHEHIP_LOGN_R_DISPATCH(logn, re, KERNEL, grid, stream, ...)                                   \
    if ((re) == 5 && (logn) >= 5) {                                                                   \
        switch (logn) {                                                                               \
            FHEHIP_LOGN_R_CASE(5, 5, KERNEL, grid, stream, __VA_ARGS__)                               \
            FHEHIP_LOGN_R_CASE(6, 5, KERNEL, grid, stream, __VA_ARGS__)                               \
            FHEHIP_LOGN_R_CASE(7, 5, KERNEL, grid, stream, __VA_ARGS__)                               \
            FHEHIP_LOGN_R_CASE(8, 5, KERNEL, grid, stream, __VA_ARGS__)                               \
            FHEHIP_LOGN_R_CASE(9, 5, KERNEL, grid, stream, __VA_ARGS__)                               \
            FHEHIP_LOGN_R_CASE(10, 5, KERNEL, grid, stream, __VA_ARGS__)                              \
            FHEHIP_LOGN_R_CASE(11, 5, KERNEL, grid, stream, __VA_ARGS__)                              \
            FHEHIP_LOGN_R_CASE(12, 5, KERNEL, grid, stream, __VA_ARGS__)                              \
            FHEHIP_LOGN_R_CASE(13, 5, KERNEL, grid, stream, __VA_ARGS__)                              \
            FHEHIP_LOGN_R_CASE(14, 5, KERNEL, grid, stream, __VA_ARGS__)                              \
            default: throw Error(-2, "degree not supported");                                         \
        }                                                                                             \
    } else {                                                                                          \
        switch (logn) {                                                                               \
            FHEHIP_LOGN_R_CASE(3, 4, KERNEL, grid, stream, __VA_ARGS__)                               \
            FHEHIP_LOGN_R_CASE(4, 4, KERNEL, grid, stream, __VA_ARGS__)                               \
            FHEHIP_LOGN_R_CASE(5, 4, KERNEL, grid, stream, __VA_ARGS__)                               \
            FHEHIP_LOGN_R_CASE(6, 4, KERNEL, grid, stream, __VA_ARGS__)                               \
            FHEHIP_LOGN_R_CASE(7, 4, KERNEL, grid, stream, __VA_ARGS__)                               \
            FHEHIP_LOGN_R_CASE(8, 4, KERNEL, grid, stream, __VA_ARGS__)                               \
            FHEHIP_LOGN_R_CASE(9, 4, KERNEL, grid, stream, __VA_ARGS__)                               \
            FHEHIP_LOGN_R_CASE(10, 4, KERNEL, grid, stream, __VA_ARGS__)                              \
            FHEHIP_LOGN_R_CASE(11, 4, KERNEL, grid, stream, __VA_ARGS__)                              \
            FHEHIP_LOGN_R_CASE(12, 4, KERNEL, grid, stream, __VA_ARGS__)                              \
            FHEHIP_LOGN_R_CASE(13, 4, KERNEL, grid, stream, __VA_ARGS__)                              \
            FHEHIP_LOGN_R_CASE(14, 4, KERNEL, grid, stream, __VA_ARGS__)                              \
            default: throw Error(-2, "degree not supported");                                         \
        }                                                                                             \
    }                                                                                                 \
    HIPCHK(hipGetLastError());

void launch_intt_rows(int logn, const InttRowsArgs& a, u32 limbs, u32 parts, u32 inst, hipStream_t s) {
    if (!limbs || !parts || !inst) return;
    ProfScope ps(a.perm ? "intt_rows_gather" : "intt_rows", logn, s);
    dim3 grid(limbs, parts, inst);
    if (a.perm) {
        FHEHIP_LOGN_DISPATCH(logn, intt_rows_gather_kernel, grid, 0, s, a);
    } else {
        FHEHIP_LOGN_DISPATCH(logn, intt_rows_kernel, grid, 0, s, a);
    }
}

#define FHEHIP_LOGH_DISPATCH(logh, F, ...)                                     \
    switch (logh) {                                                             \
        case 11: F<11>(__VA_ARGS__); break;                                     \
        case 12: F<12>(__VA_ARGS__); break;                                     \
        case 13: F<13>(__VA_ARGS__); break;                                     \
        case 14: F<14>(__VA_ARGS__); break;                                     \
        default: throw Error(-2, "half-row transforms need a degree of 4096 .. 32768"); \
    }

template <int LOGH>
static void intt_rows_half_go(const InttRowsArgs& h, dim3 grid, hipStream_t s) {
    if (h.perm)
        hipLaunchKernelGGL((intt_rows_half_kernel<LOGH, true>), grid, dim3(NttShape<LOGH>::T), 0, s, h);
    else
        hipLaunchKernelGGL((intt_rows_half_kernel<LOGH, false>), grid, dim3(NttShape<LOGH>::T), 0, s, h);
}

void launch_intt_rows_half(const DeviceContext& c, const InttRowsArgs& a, u32 limbs, u32 parts, u32 inst,
                           hipStream_t s) {
    if (!limbs || !parts || !inst) return;
    if (!c.half) throw Error(-2, "this context has no half-row tables");
    ProfScope ps(a.perm ? "intt_rows_gather" : "intt_rows", c.logn, s);
    InttRowsArgs h = a;
    h.inv_tw = c.half->inv_tw.as<ulonglong2>();
    h.mods = c.half->mod.as<ModInfo>();
    FHEHIP_LOGH_DISPATCH(c.logn - 1, intt_rows_half_go, h, dim3(2 * limbs, parts, inst), s);
    HIPCHK(hipGetLastError());
}

template <int LOGH>
static void tensor_half_go(const TensorArgs& b, dim3 grid, hipStream_t s) {
    hipLaunchKernelGGL(tensor_intt_half_kernel<LOGH>, grid, dim3(NttShape<LOGH>::T), 0, s, b);
}

void launch_tensor_intt_half(const DeviceContext& mul, const TensorArgs& a, u32 M, u32 inst, hipStream_t s) {
    if (!M || !inst) return;
    if (!mul.half) throw Error(-2, "this context has no half-row tables");
    ProfScope ps("tensor_intt", mul.logn, s);
    TensorArgs b = a;
    b.M = M;
    b.inv_tw = mul.half->inv_tw.as<ulonglong2>();
    b.mods = mul.half->mod.as<ModInfo>();
    FHEHIP_LOGH_DISPATCH(mul.logn - 1, tensor_half_go, b, dim3(M * 6 * inst), s);
    HIPCHK(hipGetLastError());
}

void launch_tensor_intt(int logn, const TensorArgs& a, u32 M, u32 inst, hipStream_t s) {
    if (!M || !inst) return;
    ProfScope ps("tensor_intt", logn, s);
    const FusedTuning tu = fused_tuning(logn);
    TensorArgs b = a;
    b.M = M;
    b.xcd = tu.xcd;
    dim3 grid(M * 3 * inst);
    FHEHIP_LOGN_R_DISPATCH(logn, tu.re_tensor, tensor_intt_kernel, grid, s, b);
}

void launch_lift_ntt(int logn, const KeySwitchArgs& a, u64* y, u32 inst, hipStream_t s) {
    if (!a.digits || !inst) return;
    ProfScope ps("lift_ntt", logn, s);
    const FusedTuning tu = fused_tuning(logn);
    KeySwitchArgs b = a;
    b.xcd = tu.xcd;
    dim3 grid(a.key_limbs * a.digits * inst);
    FHEHIP_LOGN_R_DISPATCH(logn, tu.re_lift, lift_ntt_kernel, grid, s, b, y);
}

void launch_ks_mac(int logn, const KsFinishArgs& f, u32 inst, hipStream_t s) {
    if (!f.sd.L || !inst) return;
    ProfScope ps("ks_mac", logn, s);
    const u32 k = (u32)fused_tuning(logn).ks_mac_inst;
    const dim3 grid(((1u << logn) + 255) / 256, f.sd.L + f.sd.extra, (inst + k - 1) / k);
    if (k == 2)
        hipLaunchKernelGGL(ks_mac_kernel<2>, grid, dim3(256), 0, s, f, (u32)logn, inst);
    else if (k == 8)
        hipLaunchKernelGGL(ks_mac_kernel<8>, grid, dim3(256), 0, s, f, (u32)logn, inst);
    else
        hipLaunchKernelGGL(ks_mac_kernel<4>, grid, dim3(256), 0, s, f, (u32)logn, inst);
    HIPCHK(hipGetLastError());
}

void launch_ks_finish(int logn, const KsFinishArgs& f, u32 inst, hipStream_t s) {
    if (!f.sd.L || !inst) return;
    ProfScope ps("ks_finish", logn, s);
    const FusedTuning tu = fused_tuning(logn);
    KsFinishArgs g = f;
    g.xcd = tu.xcd;
    dim3 grid(inst * 2 * f.sd.L);
    FHEHIP_LOGN_R_DISPATCH(logn, tu.re_finish, ks_finish_kernel, grid, s, g);
}

template <int LOGH>
static void ks_finish_half_go(const KsFinishArgs& f, const DeviceContext& ct, const u64* pre, dim3 grid,
                              hipStream_t s) {
    hipLaunchKernelGGL(ks_finish_half_kernel<LOGH>, grid, dim3(NttShape<LOGH>::T), 0, s, f,
                       ct.half->fwd_tw.as<ulonglong2>(), ct.split.as<ulonglong2>(), pre);
}

void launch_ks_finish_half(const KsFinishArgs& f, const DeviceContext& ct, u32 inst, hipStream_t s, const u64* pre) {
    if (!f.sd.L || !inst) return;
    if (!ct.half) throw Error(-2, "this context has no half-row tables");
    ProfScope ps("ks_finish", ct.logn, s);
    FHEHIP_LOGH_DISPATCH(ct.logn - 1, ks_finish_half_go, f, ct, pre, dim3(inst * 2 * f.sd.L * 2), s);
    HIPCHK(hipGetLastError());
}

// The first half of the key-switch finish of rows transformed as halves, as
// one elementwise pass over column pairs (c, c + N/2): the last inverse
// stage of the dropped limb's two half transforms (as ntt_split_inv_kernel,
// with the key context's constants of that limb), then for every limb i of
// the ciphertext context z = add_pb - q_l^-1 [y]_i at both columns (as
// ks_finish_kernel) and the first forward stage of the pair (Harvey
// butterfly, native.rs:256-269, outputs < 4p) into pre [inst][2][L][N].
// grid = (N / 512, 2, instances)
__global__ __launch_bounds__(256) void ks_z_pair_kernel(KsFinishArgs f, const ulonglong2* __restrict__ ct_split,
                                                        const ulonglong2* __restrict__ drop_split,
                                                        u64* __restrict__ pre, u32 logn) {
    const SdNttArgs& a = f.sd;
    const u32 half = 1u << (logn - 1);
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    const u32 part = blockIdx.y, b = blockIdx.z;
    const u64 ql = a.extra ? a.extra_mods[0] : 0, hq = ql >> 1;
    u64 y0 = 0, y1 = 0;  // (y + floor(q_l / 2)) mod q_l at c and c + N/2
    if (a.extra) {
        const u64* lrow = f.special + (((size_t)b * 2 + part) << logn);
        const u64 s0 = lrow[c], s1 = lrow[c + half];
        const ulonglong2 z = drop_split[1], ni = drop_split[2];
        y0 = mul_shoup_u(s0 + s1, ni.x, ni.y, ql) + hq;
        y1 = mul_shoup_u(s0 + 2 * ql - s1, z.x, z.y, ql) + hq;
        y0 = y0 >= ql ? y0 - ql : y0;
        y1 = y1 >= ql ? y1 - ql : y1;
    }
    u64* xb = pre + (((size_t)b * 2 + part) * a.L << logn);
    const u64* ad = a.add_pb ? a.add_pb + (size_t)b * a.add_inst_stride + (size_t)part * a.add_part_stride : nullptr;
    for (u32 i = 0; i < a.L; i++) {
        const ModInfo mi = a.mods[i];
        const u64 p = mi.p;
        const SdConst cst = a.sd[i];
        auto z_of = [&](u64 yv) {
            const u64 yr = ql <= 2 * p ? yv : lazy_reduce_u64(yv, mi.barrett_lo, mi.barrett_hi, p);
            return mul_shoup_u(3 * p - (yr + cst.half_mod), cst.inv, cst.inv_shoup, p);
        };
        u64 z0 = a.extra ? z_of(y0) : 0, z1 = a.extra ? z_of(y1) : 0;
        if (ad) {
            z0 = add_mod(z0, ad[((size_t)i << logn) + c], p);
            z1 = add_mod(z1, ad[((size_t)i << logn) + c + half], p);
        }
        const ulonglong2 w = ct_split[3 * i];
        const u64 wc = lazy_mul_shoup_u(z1, w.x, w.y, p);
        xb[((size_t)i << logn) + c] = z0 + wc;
        xb[((size_t)i << logn) + c + half] = z0 + 2 * p - wc;
    }
}

void launch_ks_z_pair(const KsFinishArgs& f, const DeviceContext& ct, const DeviceContext& key, u64* pre, u32 inst,
                      hipStream_t s) {
    if (!f.sd.L || !inst) return;
    if (ct.N < 512 || key.N != ct.N) throw Error(-2, "paired key-switch finish needs one degree of at least 512");
    ProfScope ps("ks_z_pair", ct.logn, s);
    hipLaunchKernelGGL(ks_z_pair_kernel, dim3(ct.N / 512, 2, inst), dim3(256), 0, s, f, ct.split.as<ulonglong2>(),
                       key.split.as<ulonglong2>() + 3 * (size_t)(key.L - 1), pre, (u32)ct.logn);
    HIPCHK(hipGetLastError());
}

void launch_sd_add_ntt(int logn, const SdNttArgs& a, u32 L, u32 parts, u32 inst, hipStream_t s) {
    if (!L || !parts || !inst) return;
    ProfScope ps("sd_add_ntt", logn, s);
    dim3 grid(L, parts, inst);
    FHEHIP_LOGN_DISPATCH(logn, sd_add_ntt_kernel, grid, 0, s, a);
}

}  // namespace fhehip
