// RnsScaler::scale on rows transformed as two halves (scale_pair_kernel):
// the columns c and c + N/2 of a pair run on two lanes of one wave, so that
// the stage pairing the halves runs inside the scaler. On load, the last inverse stage of two
// half inverse transforms (as ntt_split_inv_kernel); on store, the first
// forward stage of the transform that follows (as ntt_split_fwd_kernel).
// Each of those kernels is one pass over the rows, so the scaler between an
// inverse and a forward half-row transform saves two.
#include "bfv_launch.h"
#include "ntt_device.h"
#include "profiler.h"
#include "rns_device.h"

namespace fhehip {

// grid = (N / 256, parts, instances); 256 threads, one column each. Lanes
// l and l + 32 of a wave take the two columns c and c + N/2 of a pair (32
// pairs per wave): each loads both residues of every limb for the inverse
// stage, and the two lanes exchange one value per output for the forward
// stage. LDS: the constant block, the split entries of the source limbs and
// their moduli, and the split entries of the output limbs are first copied to
// LDS (as rns_scale_kernel's LDS form).
template <int NF, bool LDS>
__global__ __launch_bounds__(256) void scale_pair_kernel(ScalePairArgs a) {
    constexpr bool SW = !LDS;
    extern __shared__ u64 pair_lds[];
    const u32 n = a.n, half = n >> 1;
    const u32 lane = threadIdx.x & 63, side = lane >> 5;
    const u32 c = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 32 + (lane & 31);
    const u32 col = c + side * half, other = c + (side ^ 1) * half;
    const u64* src = a.in + blockIdx.z * a.in_inst + blockIdx.y * a.in_part;
    u64* dst = a.out + blockIdx.z * a.out_inst + blockIdx.y * a.out_part + col;
    u64 r[NF], q[NF];
    const u64* isp = (const u64*)a.in_split;    // limb j: words 6 j .. 6 j + 5 = (w, w', z, z', n^-1, n^-1')
    const u64* osp = (const u64*)a.out_split;   // output i: words 6 i, 6 i + 1 = (w, w')
#pragma unroll
    for (int j = 0; j < NF; j++) {
        r[j] = src[(size_t)j * n + col];
        if (isp) q[j] = src[(size_t)j * n + other];
    }
    const u64* sc = a.sc;
    const u64* imod = nullptr;  // LDS: modulus of source limb j
    if constexpr (LDS) {
        const u32 w_sc = a.sc_words, w_in = isp ? 7 * NF : 0, w_out = osp ? 6 * a.count : 0;
        for (u32 k = threadIdx.x; k < w_sc + w_in + w_out; k += 256) {
            u64 v;
            if (k < w_sc) v = a.sc[k];
            else if (k < w_sc + 6 * NF && isp) v = isp[k - w_sc];
            else if (k < w_sc + w_in) v = a.in_mods[k - w_sc - 6 * NF].p;
            else v = osp[k - w_sc - w_in];
            pair_lds[k] = v;
        }
        __syncthreads();
        sc = pair_lds;
        if (isp) {
            isp = pair_lds + w_sc;
            imod = pair_lds + w_sc + 6 * NF;
        }
        if (osp) osp = pair_lds + w_sc + w_in;
    }
    if (isp) {
        // X = (x + y) n^-1 at c, Y = (x - y) zeta n^-1 at c + N/2, with x at
        // c and y at c + N/2 (native.rs:288-300, 130-131)
#pragma unroll
        for (int j = 0; j < NF; j++) {
            const u64 p = LDS ? imod[j] : a.in_mods[j].p;
            const u64 x = side ? q[j] : r[j], y = side ? r[j] : q[j];
            const u64 w = side ? isp[6 * j + 2] : isp[6 * j + 4];
            const u64 ws = side ? isp[6 * j + 3] : isp[6 * j + 5];
            r[j] = mul_shoup(side ? x + 2 * p - y : x + y, w, ws, p);
        }
    }
    const ScHeader<SW> h(sc);
    auto rests = [&](int j) { return r[j]; };
    const ScColumn cv = sc_column<NF, SW>(sc, rests, h);
    const u64* outc = sc + sc_out_base(NF);
    for (int i = 0; i < h.count; i++) {
        const u64* o = outc + (size_t)i * sc_out_stride(NF);
        u64 y = sc_output<NF, SW>(o, rests, cv, h);
        if (osp) {
            // Harvey forward butterfly of reduced inputs (native.rs:256-269),
            // outputs < 4q: the lane at c + N/2 forms t = w y and sends it,
            // the lane at c sends its y.
            const u64 qi = o[0];
            const u64 t = lazy_mul_shoup(y, osp[6 * i], osp[6 * i + 1], qi);
            const u64 send = side ? t : y;
            const u32 lo = (u32)__shfl_xor((int)(u32)send, 32), hi = (u32)__shfl_xor((int)(u32)(send >> 32), 32);
            const u64 got = ((u64)hi << 32) | lo;
            y = side ? got + 2 * qi - t : y + got;
        }
        dst[(size_t)i * n] = y;
    }
}

// Large launches: one thread takes both columns c and c + N/2 of a pair, so
// every residue is loaded once and the two stages pairing the halves need
// no exchange between lanes; the constants (scalar cache) serve both
// columns. The columns are scaled one after the other: while column c is in
// registers, the residues of column c + N/2 (and, with the forward stage,
// column c's outputs) wait in the thread's own LDS slots, so the kernel holds
// the registers of one column, as rns_scale_kernel does.
// grid = (N / 512, parts, instances), 256 threads; LDS: (NF + count) * 256
// words with the forward stage, NF * 256 without.
// The pointers are separate __restrict__ parameters (not fields of the
// argument struct), so that the constants are known not to be written by the
// kernel and load through the scalar cache.
template <int NF>
__global__ __launch_bounds__(256) void scale_pair2_kernel(const u64* __restrict__ in, size_t in_inst, size_t in_part,
                                                          u64* __restrict__ out, size_t out_inst, size_t out_part,
                                                          const u64* __restrict__ sc, const u64* __restrict__ isp,
                                                          const ModInfo* __restrict__ in_mods,
                                                          const u64* __restrict__ osp, u32 n) {
    // isp: limb j at words 6 j .. 6 j + 5 = (w, w', z, z', n^-1, n^-1');
    // osp: output i at words 6 i, 6 i + 1 = (w, w')
    extern __shared__ u64 stage[];
    const u32 half = n >> 1, tid = threadIdx.x;
    const u32 c = blockIdx.x * 256 + tid;
    const u64* src = in + blockIdx.z * in_inst + blockIdx.y * in_part + c;
    u64* dst = out + blockIdx.z * out_inst + blockIdx.y * out_part + c;
    u64 r[NF];
#pragma unroll
    for (int j = 0; j < NF; j++) {
        const u64 x = src[(size_t)j * n], y = src[(size_t)j * n + half];
        if (isp) {
            // X = (x + y) n^-1 at c, Y = (x - y) zeta n^-1 at c + N/2
            // (native.rs:288-300, 130-131)
            const u64 p = in_mods[j].p;
            r[j] = mul_shoup_u(x + y, isp[6 * j + 4], isp[6 * j + 5], p);
            stage[j * 256 + tid] = mul_shoup_u(x + 2 * p - y, isp[6 * j + 2], isp[6 * j + 3], p);
        } else {
            r[j] = x;
            stage[j * 256 + tid] = y;
        }
    }
    const ScHeader<true> h(sc);
    auto rests = [&](int j) { return r[j]; };
    const u64* outc = sc + sc_out_base(NF);
    {
        const ScColumn cv = sc_column<NF, true>(sc, rests, h);
        for (int i = 0; i < h.count; i++) {
            const u64 y = sc_output<NF, true>(outc + (size_t)i * sc_out_stride(NF), rests, cv, h);
            if (osp) stage[(NF + i) * 256 + tid] = y;
            else dst[(size_t)i * n] = y;
        }
    }
#pragma unroll
    for (int j = 0; j < NF; j++) r[j] = stage[j * 256 + tid];
    const ScColumn cv = sc_column<NF, true>(sc, rests, h);
    for (int i = 0; i < h.count; i++) {
        const u64* o = outc + (size_t)i * sc_out_stride(NF);
        u64 y1 = sc_output<NF, true>(o, rests, cv, h);
        if (osp) {
            // Harvey forward butterfly of reduced inputs (native.rs:256-269),
            // outputs < 4q.
            const u64 qi = o[0];
            const u64 y0 = stage[(NF + i) * 256 + tid];
            const u64 t = lazy_mul_shoup_u(y1, osp[6 * i], osp[6 * i + 1], qi);
            dst[(size_t)i * n] = y0 + t;
            y1 = y0 + 2 * qi - t;
        }
        dst[(size_t)i * n + half] = y1;
    }
}

template <int NF>
static void scale_pair_go(const ScalePairArgs& a, dim3 grid, bool lds, hipStream_t s) {
    if (lds) {
        const size_t words = a.sc_words + (a.in_split ? 7 * (size_t)NF : 0) + (a.out_split ? 6 * (size_t)a.count : 0);
        hipLaunchKernelGGL((scale_pair_kernel<NF, true>), grid, dim3(256), words * 8, s, a);
    } else {
        const dim3 g2(a.n / 512, grid.y, grid.z);
        const size_t words = (size_t)(NF + (a.out_split ? a.count : 0)) * 256;
        hipLaunchKernelGGL((scale_pair2_kernel<NF>), g2, dim3(256), words * 8, s, a.in, a.in_inst, a.in_part, a.out,
                           a.out_inst, a.out_part, a.sc, (const u64*)a.in_split, a.in_mods, (const u64*)a.out_split,
                           a.n);
    }
}

void launch_scale_pair(const ScalePairArgs& a, u32 nf, u32 parts, u32 inst, hipStream_t s, const char* name) {
    if (!parts || !inst) return;
    if (a.n < 512 || (a.n & (a.n - 1))) throw Error(-2, "paired scaling needs a power-of-two degree of at least 512");
    if (a.in_split && !a.in_mods) throw Error(-1, "paired scaling: the inverse stage needs the source moduli");
    ProfScope ps(name, 31 - __builtin_clz(a.n), s);
    const dim3 grid(a.n / 256, parts, inst);
    const bool lds = (size_t)grid.x * parts * inst <= kScaleLdsMaxBlocks;
#define FHEHIP_PAIR_CASE(K) \
    case K: scale_pair_go<K>(a, grid, lds, s); break;
    switch (nf) {
        FHEHIP_PAIR_CASE(1) FHEHIP_PAIR_CASE(2) FHEHIP_PAIR_CASE(3) FHEHIP_PAIR_CASE(4) FHEHIP_PAIR_CASE(5)
        FHEHIP_PAIR_CASE(6) FHEHIP_PAIR_CASE(7) FHEHIP_PAIR_CASE(8) FHEHIP_PAIR_CASE(9) FHEHIP_PAIR_CASE(10)
        FHEHIP_PAIR_CASE(11) FHEHIP_PAIR_CASE(12) FHEHIP_PAIR_CASE(13) FHEHIP_PAIR_CASE(14) FHEHIP_PAIR_CASE(15)
        FHEHIP_PAIR_CASE(16) FHEHIP_PAIR_CASE(17) FHEHIP_PAIR_CASE(18) FHEHIP_PAIR_CASE(19) FHEHIP_PAIR_CASE(20)
        FHEHIP_PAIR_CASE(21) FHEHIP_PAIR_CASE(22) FHEHIP_PAIR_CASE(23) FHEHIP_PAIR_CASE(24) FHEHIP_PAIR_CASE(25)
        FHEHIP_PAIR_CASE(26) FHEHIP_PAIR_CASE(27) FHEHIP_PAIR_CASE(28) FHEHIP_PAIR_CASE(29) FHEHIP_PAIR_CASE(30)
        FHEHIP_PAIR_CASE(31) FHEHIP_PAIR_CASE(32)
        default: throw Error(-2, "more than 32 source moduli in one scaling");
    }
#undef FHEHIP_PAIR_CASE
    HIPCHK(hipGetLastError());
}

}  // namespace fhehip
