// Kernels of the BFV evaluation path: RNS scaling columns, INTT with optional
// Galois gather, fused tensor + INTT, key switching (digit lift + lazy NTT +
// Shoup multiply-accumulate), and fused switch-down + add + NTT.
#include "bfv_launch.h"
#include "ntt_device.h"
#include "profiler.h"
#include "rns_device.h"

namespace fhehip {

// ---------------------------------------------------------------------------
// RnsScaler::scale over the columns of a batch (rq::Scaler::scale conversion
// loop, math/src/rq/scaler.rs:84-93). Power-basis rows of part k of instance b
// start at in + b * in_inst + k * in_part; outputs likewise.
// grid = (N / 256, parts, instances)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void rns_scale_kernel(const u64* __restrict__ in, size_t in_inst, size_t in_part,
                                                        u64* __restrict__ out, size_t out_inst, size_t out_part,
                                                        const u64* __restrict__ sc, u32 n) {
    const u32 col = blockIdx.x * 256 + threadIdx.x;
    if (col >= n) return;
    const u64* src = in + blockIdx.z * in_inst + blockIdx.y * in_part + col;
    u64* dst = out + blockIdx.z * out_inst + blockIdx.y * out_part + col;
    rns_scale_column(
        sc, [&](int j) { return src[(size_t)j * n]; }, [&](int i, u64 v) { dst[(size_t)i * n] = v; });
}

void launch_rns_scale(const u64* in, size_t in_inst, size_t in_part, u64* out, size_t out_inst, size_t out_part,
                      const u64* d_sc, u32 n, u32 parts, u32 inst, hipStream_t s) {
    if (!parts || !inst) return;
    ProfScope ps("rns_scale", 31 - __builtin_clz(n), s);
    dim3 grid((n + 255) / 256, parts, inst);
    hipLaunchKernelGGL(rns_scale_kernel, grid, dim3(256), 0, s, in, in_inst, in_part, out, out_inst, out_part, d_sc,
                       n);
    HIPCHK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Out-of-place inverse NTT of rows, with an optional NTT-domain gather
// (Poly::substitute in the Ntt representation, math/src/rq/mod.rs:324-346,
// as q[m] = p[perm[m]]). grid = (limbs, parts, instances).
// ---------------------------------------------------------------------------
template <int LOGN>
__global__ __launch_bounds__(NttShape<LOGN>::T) void intt_rows_kernel(InttRowsArgs a) {
    using S = NttShape<LOGN>;
    __shared__ u64 lds[S::LDS_WORDS];
    const int t = threadIdx.x;
    const u32 l = blockIdx.x, k = blockIdx.y, b = blockIdx.z;
    const u64* src = a.src[k] + (size_t)b * a.src_inst_stride + ((size_t)l << LOGN);
    u64* dst = a.dst + (size_t)b * a.dst_inst_stride + (size_t)k * a.dst_part_stride + ((size_t)l << LOGN);
    const u32 limb = l + a.table_offset;
    const ulonglong2* tw = a.inv_tw + ((size_t)limb << LOGN);
    const ModInfo mi = a.mods[limb];
    u64 x[S::E];
    if (a.perm) {
#pragma unroll
        for (int e = 0; e < S::E; e++) x[e] = src[a.perm[(t << S::RE) + e]];
    } else {
        const ulonglong2* in = reinterpret_cast<const ulonglong2*>(src + ((size_t)t << S::RE));
#pragma unroll
        for (int e = 0; e < S::E / 2; e++) {
            ulonglong2 v = in[e];
            x[2 * e] = v.x;
            x[2 * e + 1] = v.y;
        }
    }
    ntt_inv_regs<LOGN>(lds, x, t, tw, mi.p);
#pragma unroll
    for (int e = 0; e < S::E; e++)
        dst[inv_index<LOGN, last_start<LOGN>()>(t, e)] = mul_shoup(x[e], mi.n_inv, mi.n_inv_shoup, mi.p);
}

// ---------------------------------------------------------------------------
// Tensor product fused with the inverse NTT (Multiplicator::multiply,
// fhe/src/bfv/ops/mul.rs:181-185 then the INTT of the down-scaler, rq/scaler.rs
// :68-78): c0 = a0 b0, c1 = a0 b1 + a1 b0, c2 = a1 b1 over the M limbs of the
// multiplication context. grid = (M, 3, instances).
// ---------------------------------------------------------------------------
template <int LOGN>
__global__ __launch_bounds__(NttShape<LOGN>::T) void tensor_intt_kernel(TensorArgs a) {
    using S = NttShape<LOGN>;
    __shared__ u64 lds[S::LDS_WORDS];
    const int t = threadIdx.x;
    const u32 m = blockIdx.x, c = blockIdx.y, b = blockIdx.z;
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
    u64 x[S::E];
    const size_t off = (size_t)t << S::RE;
#pragma unroll
    for (int e = 0; e < S::E; e += 2) {
        ulonglong2 a0 = *reinterpret_cast<const ulonglong2*>(rows[0] + off + e);
        ulonglong2 a1 = *reinterpret_cast<const ulonglong2*>(rows[1] + off + e);
        ulonglong2 b0 = *reinterpret_cast<const ulonglong2*>(rows[2] + off + e);
        ulonglong2 b1 = *reinterpret_cast<const ulonglong2*>(rows[3] + off + e);
        if (c == 0) {
            x[e] = mul_mod(a0.x, b0.x, mi.barrett_lo, mi.barrett_hi, mi.p);
            x[e + 1] = mul_mod(a0.y, b0.y, mi.barrett_lo, mi.barrett_hi, mi.p);
        } else if (c == 1) {
            x[e] = add_mod(mul_mod(a0.x, b1.x, mi.barrett_lo, mi.barrett_hi, mi.p),
                           mul_mod(a1.x, b0.x, mi.barrett_lo, mi.barrett_hi, mi.p), mi.p);
            x[e + 1] = add_mod(mul_mod(a0.y, b1.y, mi.barrett_lo, mi.barrett_hi, mi.p),
                               mul_mod(a1.y, b0.y, mi.barrett_lo, mi.barrett_hi, mi.p), mi.p);
        } else {
            x[e] = mul_mod(a1.x, b1.x, mi.barrett_lo, mi.barrett_hi, mi.p);
            x[e + 1] = mul_mod(a1.y, b1.y, mi.barrett_lo, mi.barrett_hi, mi.p);
        }
    }
    const ulonglong2* tw = a.inv_tw + ((size_t)m << LOGN);
    ntt_inv_regs<LOGN>(lds, x, t, tw, mi.p);
    u64* dst = a.out + (size_t)b * a.out_inst_stride + (size_t)c * a.out_part_stride + ((size_t)m << LOGN);
#pragma unroll
    for (int e = 0; e < S::E; e++)
        dst[inv_index<LOGN, last_start<LOGN>()>(t, e)] = mul_shoup(x[e], mi.n_inv, mi.n_inv_shoup, mi.p);
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
template <int LOGN>
__global__ __launch_bounds__(NttShape<LOGN>::T) void keyswitch_kernel(KeySwitchArgs a) {
    using S = NttShape<LOGN>;
    __shared__ u64 lds[S::LDS_WORDS];
    const int t = threadIdx.x;
    const u32 j = blockIdx.x, b = blockIdx.y;
    const ModInfo mi = a.key_mods[j];
    const u64 p = mi.p;
    const ulonglong2* ftw = a.key_fwd_tw + ((size_t)j << LOGN);
    u64* acc0 = a.acc + (size_t)b * a.acc_inst_stride + ((size_t)j << LOGN);
    u64* acc1 = acc0 + a.acc_part_stride;
    for (u32 i = 0; i < a.digits; i++) {
        const u64* src = a.src + (size_t)b * a.src_inst_stride + ((size_t)i << LOGN);
        u64 x[S::E];
#pragma unroll
        for (int e = 0; e < S::E; e++)
            x[e] = lazy_reduce_u64(src[fwd_index<LOGN, 0>(t, e)], mi.barrett_lo, mi.barrett_hi, p);
        // Launder the twiddle pointer so that the per-thread twiddles of the
        // later passes are reloaded per digit instead of being hoisted out of
        // the loop (which would keep ~170 extra registers live).
        u64 twp = (u64)ftw;
        asm volatile("" : "+s"(twp));
        ntt_fwd_regs<LOGN>(lds, x, t, (const ulonglong2*)twp, p);
        // Stage the lazily transformed digit through LDS so that the key rows
        // and the accumulators are walked with coalesced accesses.
#pragma unroll
        for (int e = 0; e < S::E; e++) lds[lds_slot((t << S::RE) + e)] = x[e];
        __syncthreads();
        const ulonglong2* k0 = a.key0 + (((size_t)i * a.key_limbs + j) << LOGN);
        const ulonglong2* k1 = a.key1 + (((size_t)i * a.key_limbs + j) << LOGN);
#pragma unroll 4
        for (int r = 0; r < S::E; r++) {
            const int jj = r * S::T + t;
            const u64 v = lds[lds_slot(jj)];
            const ulonglong2 w0 = k0[jj], w1 = k1[jj];
            u64 v0 = mul_shoup(v, w0.x, w0.y, p);
            u64 v1 = mul_shoup(v, w1.x, w1.y, p);
            if (i) {
                v0 = add_mod(v0, acc0[jj], p);
                v1 = add_mod(v1, acc1[jj], p);
            }
            acc0[jj] = v0;
            acc1[jj] = v1;
        }
        __syncthreads();
    }
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
                y[e] = mul_shoup(y[e] + 3 * cst.q - tmp, cst.inv, cst.inv_shoup, cst.q);
            }
        }
        const SdConst cst = a.sd[s * a.sd_stride + i];
#pragma unroll
        for (int e = 0; e < S::E; e++) {
            u64 z = y[e] + half;
            z = z >= ql ? z - ql : z;
            u64 tmp = lazy_reduce_u64(z, mi.barrett_lo, mi.barrett_hi, p) + cst.half_mod;
            x[e] = mul_shoup(x[e] + 3 * p - tmp, cst.inv, cst.inv_shoup, p);
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
    const size_t off = (size_t)t << S::RE;
    if (a.add_ntt && part < a.add_ntt_parts) {
        const u64* g = a.add_ntt + (size_t)b * a.add_ntt_inst_stride + (size_t)part * a.add_ntt_part_stride +
                       ((size_t)i << LOGN);
#pragma unroll
        for (int e = 0; e < S::E; e++) {
            u64 v = reduce1(reduce1(x[e], p2), p);
            x[e] = add_mod(v, g[a.perm ? a.perm[off + e] : (u32)(off + e)], p);
        }
    } else {
#pragma unroll
        for (int e = 0; e < S::E; e++) x[e] = reduce1(reduce1(x[e], p2), p);
    }
    u64* dst = a.out + (size_t)b * a.out_inst_stride + (size_t)part * a.out_part_stride + ((size_t)i << LOGN);
    ulonglong2* o = reinterpret_cast<ulonglong2*>(dst + off);
#pragma unroll
    for (int e = 0; e < S::E / 2; e++) o[e] = make_ulonglong2(x[2 * e], x[2 * e + 1]);
}

// ---------------------------------------------------------------------------
#define FHEHIP_LOGN_DISPATCH(logn, KERNEL, grid, block, stream, arg)                         \
    switch (logn) {                                                                          \
        case 3: hipLaunchKernelGGL(KERNEL<3>, grid, dim3(NttShape<3>::T), 0, stream, arg); break;   \
        case 4: hipLaunchKernelGGL(KERNEL<4>, grid, dim3(NttShape<4>::T), 0, stream, arg); break;   \
        case 5: hipLaunchKernelGGL(KERNEL<5>, grid, dim3(NttShape<5>::T), 0, stream, arg); break;   \
        case 6: hipLaunchKernelGGL(KERNEL<6>, grid, dim3(NttShape<6>::T), 0, stream, arg); break;   \
        case 7: hipLaunchKernelGGL(KERNEL<7>, grid, dim3(NttShape<7>::T), 0, stream, arg); break;   \
        case 8: hipLaunchKernelGGL(KERNEL<8>, grid, dim3(NttShape<8>::T), 0, stream, arg); break;   \
        case 9: hipLaunchKernelGGL(KERNEL<9>, grid, dim3(NttShape<9>::T), 0, stream, arg); break;   \
        case 10: hipLaunchKernelGGL(KERNEL<10>, grid, dim3(NttShape<10>::T), 0, stream, arg); break; \
        case 11: hipLaunchKernelGGL(KERNEL<11>, grid, dim3(NttShape<11>::T), 0, stream, arg); break; \
        case 12: hipLaunchKernelGGL(KERNEL<12>, grid, dim3(NttShape<12>::T), 0, stream, arg); break; \
        case 13: hipLaunchKernelGGL(KERNEL<13>, grid, dim3(NttShape<13>::T), 0, stream, arg); break; \
        case 14: hipLaunchKernelGGL(KERNEL<14>, grid, dim3(NttShape<14>::T), 0, stream, arg); break; \
        default: throw Error(-2, "degree not supported");                                    \
    }                                                                                        \
    HIPCHK(hipGetLastError());

void launch_intt_rows(int logn, const InttRowsArgs& a, u32 limbs, u32 parts, u32 inst, hipStream_t s) {
    if (!limbs || !parts || !inst) return;
    ProfScope ps(a.perm ? "intt_rows_gather" : "intt_rows", logn, s);
    dim3 grid(limbs, parts, inst);
    FHEHIP_LOGN_DISPATCH(logn, intt_rows_kernel, grid, 0, s, a);
}

void launch_tensor_intt(int logn, const TensorArgs& a, u32 M, u32 inst, hipStream_t s) {
    if (!M || !inst) return;
    ProfScope ps("tensor_intt", logn, s);
    dim3 grid(M, 3, inst);
    FHEHIP_LOGN_DISPATCH(logn, tensor_intt_kernel, grid, 0, s, a);
}

void launch_keyswitch(int logn, const KeySwitchArgs& a, u32 key_limbs, u32 inst, hipStream_t s) {
    if (!key_limbs || !inst) return;
    ProfScope ps("keyswitch", logn, s);
    dim3 grid(key_limbs, inst);
    FHEHIP_LOGN_DISPATCH(logn, keyswitch_kernel, grid, 0, s, a);
}

void launch_sd_add_ntt(int logn, const SdNttArgs& a, u32 L, u32 parts, u32 inst, hipStream_t s) {
    if (!L || !parts || !inst) return;
    ProfScope ps("sd_add_ntt", logn, s);
    dim3 grid(L, parts, inst);
    FHEHIP_LOGN_DISPATCH(logn, sd_add_ntt_kernel, grid, 0, s, a);
}

}  // namespace fhehip
