// Kernels of the PIR server path: plaintext-ciphertext dot products over a
// database resident in HBM, and the elementwise step of the oblivious
// expansion of a query ciphertext.
#include "bfv_launch.h"
#include "ntt_device.h"
#include "profiler.h"
#include "rns_device.h"

// Most columns one thread accumulates (4 or 2).
#ifndef FHEHIP_DOT_CG_MAX
#define FHEHIP_DOT_CG_MAX 4
#endif

namespace fhehip {

// dot_product_scalar (fhe/src/bfv/ops/dot_product.rs:56-143) for CG columns
// per thread: out[c][p][k] = sum_r ct[r][p][k] * pt[r][c][k] mod q, k = one
// word of [L][N]. The sums are exact (column accumulators) and reduced once,
// so the result is the reference's whichever of its two accumulation
// schedules it takes. Each thread loads the ciphertext words of a row once
// for its CG columns; blocks that share a coefficient chunk get consecutive
// logical indices on one XCD, so the other column groups read those words
// from L2. 1-D grid of (64-word chunk, column group) blocks of 256 threads.
// The plaintext (database) words are read once: non-temporal loads keep the
// ciphertext rows, which every column group re-reads, in L2.
template <int PARTS, int CG>
__device__ __forceinline__ void dot_load(const u64* ct, const u64* pt, const size_t (&poff)[PARTS],
                                         const size_t (&coff)[CG], u64 (&x)[PARTS], u64 (&y)[CG]) {
#pragma unroll
    for (int p = 0; p < PARTS; p++) x[p] = ct[poff[p]];
#pragma unroll
    for (int c = 0; c < CG; c++) y[c] = __builtin_nontemporal_load(pt + coff[c]);
}

template <int PARTS, int CG>
__device__ __forceinline__ void dot_mac(Acc3 (&acc)[CG][PARTS], const u64 (&x)[PARTS], const u64 (&y)[CG]) {
#pragma unroll
    for (int c = 0; c < CG; c++)
#pragma unroll
        for (int p = 0; p < PARTS; p++) acc3_mul<false>(acc[c][p], x[p], y[c]);
}

// Rows are split over the four waves of a workgroup (wave g takes rows g,
// g + 4, ...), each lane one word of a 64-word chunk: four times the waves of
// one thread per (word, column group), so the grid is several rounds of
// resident waves instead of one and a fraction. Each wave reduces its sums
// to residues; the four residues of an output are added through LDS.
constexpr u32 kDotGroups = 4;

template <int PARTS, int CG>
__global__ __launch_bounds__(256) void dot_product_kernel(DotArgs a) {
    __shared__ u64 part[kDotGroups][CG * PARTS][64];
    // Consecutive logical blocks (one XCD) share a coefficient chunk: part
    // groups vary fastest (they share the database words), then column groups
    // (they share the ciphertext words).
    const u32 w = xcd_block(blockIdx.x, gridDim.x);
    const u32 pg = w % a.partgroups, rest = w / a.partgroups;
    const u32 chunk = rest / a.colgroups, cg = rest % a.colgroups;
    const u32 lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const u32 k = chunk * 64 + lane;
    const bool live = k < a.words;
    const u32 kk = live ? k : a.words - 1;   // idle lanes read a valid word and store nothing
    const u32 c0 = cg * CG;
    const u32 ncol = a.columns - c0 < (u32)CG ? a.columns - c0 : (u32)CG;
    // Columns past the last one re-read the last column (in bounds, no
    // branch in the loop); their sums are not stored.
    size_t coff[CG];
#pragma unroll
    for (int c = 0; c < CG; c++) coff[c] = (size_t)((u32)c < ncol ? (u32)c : ncol - 1) * a.pt_col_stride;
    // Likewise for the parts past the last one of a part group.
    const u32 p0 = pg * PARTS;
    const u32 npart = a.parts - p0 < (u32)PARTS ? a.parts - p0 : (u32)PARTS;
    size_t poff[PARTS];
#pragma unroll
    for (int p = 0; p < PARTS; p++) poff[p] = (size_t)(p0 + ((u32)p < npart ? (u32)p : npart - 1)) * a.words;
    Acc3 acc[CG][PARTS];
#pragma unroll
    for (int c = 0; c < CG; c++)
#pragma unroll
        for (int p = 0; p < PARTS; p++) acc3_init(acc[c][p], 0);
    if (g < a.count) {
        const u64* ct = a.ct + (size_t)g * a.ct_stride + kk;
        const u64* pt = a.pt + (size_t)g * a.pt_row_stride + (size_t)c0 * a.pt_col_stride + kk;
        const size_t cstep = kDotGroups * a.ct_stride, pstep = kDotGroups * a.pt_row_stride;
        u64 x[PARTS], y[CG];
        dot_load<PARTS, CG>(ct, pt, poff, coff, x, y);
        // The next row's words are in flight while this row's products accumulate.
        for (u32 r = g + kDotGroups; r < a.count; r += kDotGroups) {
            ct += cstep;
            pt += pstep;
            u64 xn[PARTS], yn[CG];
            dot_load<PARTS, CG>(ct, pt, poff, coff, xn, yn);
            dot_mac<PARTS, CG>(acc, x, y);
#pragma unroll
            for (int p = 0; p < PARTS; p++) x[p] = xn[p];
#pragma unroll
            for (int c = 0; c < CG; c++) y[c] = yn[c];
        }
        dot_mac<PARTS, CG>(acc, x, y);
    }
    const ModInfo mi = a.mods[kk >> a.logn];
#pragma unroll
    for (int c = 0; c < CG; c++)
#pragma unroll
        for (int p = 0; p < PARTS; p++) {
            // lo + 2^64 hi + 2^128 top: (top 2^64 + hi) mod q, then (that 2^64 + lo) mod q;
            // top < count < 2^31 and q < 2^62 keep both Barrett inputs below 2^127.
            u64 lo, hi, top;
            acc3_value(acc[c][p], lo, hi, top);
            const u64 r1 = reduce_u128(hi, top, mi.barrett_lo, mi.barrett_hi, mi.p);
            part[g][c * PARTS + p][lane] = reduce_u128(lo, r1, mi.barrett_lo, mi.barrett_hi, mi.p);
        }
    __syncthreads();
    for (u32 j = g; j < (u32)(CG * PARTS); j += kDotGroups) {
        const u32 c = j / PARTS, p = j % PARTS;
        u64 v = part[0][j][lane];
#pragma unroll
        for (u32 h = 1; h < kDotGroups; h++) v = add_mod(v, part[h][j][lane], mi.p);
        if (live && c < ncol && p < npart)
            a.out[(size_t)(c0 + c) * a.out_col_stride + (size_t)(p0 + p) * a.words + k] = v;
    }
}

template <int PARTS, int CG>
static void launch_dot(const DotArgs& a, u32 blocks, hipStream_t s) {
    hipLaunchKernelGGL((dot_product_kernel<PARTS, CG>), dim3(blocks), dim3(256), 0, s, a);
}

void launch_dot_product(DotArgs a, u32 parts, hipStream_t s) {
    if (!a.count || !a.columns || !a.words || !parts) return;
    // Parts and columns per thread: the ciphertext words of a row are loaded
    // once for CG columns and the database words once for PARTS parts; the
    // accumulators take 9 VGPRs per (column, part), at most 8 of them.
    int pt, cg;
    if (parts <= 2) {
        pt = (int)parts;
        cg = a.columns >= 4 && FHEHIP_DOT_CG_MAX >= 4 ? 4 : (a.columns >= 2 ? 2 : 1);
    } else {
        pt = parts == 3 ? 3 : 4;
        cg = a.columns >= 2 ? 2 : 1;
    }
    a.parts = parts;
    a.partgroups = (parts + pt - 1) / pt;
    a.colgroups = (a.columns + cg - 1) / cg;
    const u32 chunks = (a.words + 63) / 64;
    const u64 blocks = (u64)chunks * a.colgroups * a.partgroups;
    if (blocks > 0x7fffffffu) throw Error(-1, "dot product grid too large");
    ProfScope ps("dot_product", a.logn, s);
    const u32 nb = (u32)blocks;
    if (pt == 1)
        cg == 4 ? launch_dot<1, 4>(a, nb, s) : cg == 2 ? launch_dot<1, 2>(a, nb, s) : launch_dot<1, 1>(a, nb, s);
    else if (pt == 2)
        cg == 4 ? launch_dot<2, 4>(a, nb, s) : cg == 2 ? launch_dot<2, 2>(a, nb, s) : launch_dot<2, 1>(a, nb, s);
    else if (pt == 3)
        cg == 2 ? launch_dot<3, 2>(a, nb, s) : launch_dot<3, 1>(a, nb, s);
    else
        cg == 2 ? launch_dot<4, 2>(a, nb, s) : launch_dot<4, 1>(a, nb, s);
    HIPCHK(hipGetLastError());
}

// One level of EvaluationKey::expands (fhe/src/bfv/keys/evaluation_key.rs:
// 168-185) for i < step, with sub[i] = GaloisKey::relinearize(out[i]):
//   out[step + i] = (out[i] - sub[i]) * monomial   (when step + i < size)
//   out[i]       += sub[i]
// All NTT-domain and exactly reduced. grid = (N/256, 2 L, step)
__global__ __launch_bounds__(256) void expand_step_kernel(u64* __restrict__ out, const u64* __restrict__ sub,
                                                          const ulonglong2* __restrict__ mono,
                                                          const ModInfo* __restrict__ mods, u32 L, u32 logn,
                                                          u32 step, u32 size) {
    const u32 e = blockIdx.x * 256 + threadIdx.x;
    if (e >> logn) return;
    const u32 m = blockIdx.y % L, i = blockIdx.z;
    const size_t sz = (size_t)2 * L << logn;
    const size_t w = ((size_t)blockIdx.y << logn) + e;
    const u64 q = mods[m].p;
    const u64 lo = out[i * sz + w], sb = sub[i * sz + w];
    if (step + i < size) {
        const ulonglong2 mw = mono[((size_t)m << logn) + e];
        out[(step + i) * sz + w] = mul_shoup(sub_mod(lo, sb, q), mw.x, mw.y, q);
    }
    out[i * sz + w] = add_mod(lo, sb, q);
}

void launch_expand_step(u64* out, const u64* sub, const ulonglong2* mono, const ModInfo* mods, u32 L, int logn,
                        u32 step, u32 size, hipStream_t s) {
    if (!step) return;
    ProfScope ps("expand_step", logn, s);
    hipLaunchKernelGGL(expand_step_kernel, dim3(((1u << logn) + 255) / 256, 2 * L, step), dim3(256), 0, s, out, sub,
                       mono, mods, L, (u32)logn, step, size);
    HIPCHK(hipGetLastError());
}

}  // namespace fhehip
