// Batched negacyclic NTT kernels: one workgroup per polynomial row.
#include <atomic>

#include "fhehip_internal.h"
#include "ntt_device.h"
#include "profiler.h"

#include <cstdlib>

namespace fhehip {

__device__ __forceinline__ u64* row_ptr(const RowMap& m, u32 r, int logn, u32& limb) {
    u32 poly = r / m.rows_per_poly;
    u32 l = m.first_limb + r % m.rows_per_poly;
    limb = l - m.first_limb + m.table_offset;
    return m.data + (((size_t)poly * m.limbs_per_poly + l) << logn);
}

// Forward NTT (NttOperator::forward / forward_vt_lazy, native.rs:77-189).
template <int LOGN, bool LAZY, int R>
__global__ __launch_bounds__((NttShape<LOGN, R>::T)) void ntt_fwd_kernel(RowMap m, const ulonglong2* __restrict__ tw_all,
                                                                      const ModInfo* __restrict__ mods) {
    using S = NttShape<LOGN, R>;
    __shared__ u64 lds[S::LDS_WORDS];
    const int t = threadIdx.x;
    u32 limb;
    u64* row = row_ptr(m, blockIdx.x, LOGN, limb);
    const ulonglong2* tw = tw_all + ((size_t)limb << LOGN);
    const u64 p = mods[limb].p;
    u64 x[S::E];
#pragma unroll
    for (int k = 0; k < S::E; k++) x[k] = row[fwd_index<LOGN, 0, R>(t, k)];
    ntt_fwd_regs<LOGN, R>(lds, x, t, tw, p);
    const u64 p2 = 2 * p;
    store_fwd_order<LOGN, R>(lds, x, t, [&](int j, u64 v) { row[j] = LAZY ? v : reduce1(reduce1(v, p2), p); });
}

// Inverse NTT (NttOperator::backward, native.rs:106-132).
template <int LOGN, int R>
__global__ __launch_bounds__((NttShape<LOGN, R>::T)) void ntt_inv_kernel(RowMap m, const ulonglong2* __restrict__ tw_all,
                                                                      const ModInfo* __restrict__ mods) {
    using S = NttShape<LOGN, R>;
    __shared__ u64 lds[S::LDS_WORDS];
    const int t = threadIdx.x;
    u32 limb;
    u64* row = row_ptr(m, blockIdx.x, LOGN, limb);
    const ulonglong2* tw = tw_all + ((size_t)limb << LOGN);
    const ModInfo mi = mods[limb];
    u64 x[S::E];
    load_inv_order<LOGN, 16, R>(lds, x, t, [&](int j) { return row[j]; });
    ntt_inv_regs<LOGN, R>(lds, x, t, tw, mi.p, inv_scale_n(mi));
    store_inv_order<LOGN, R>(row, x, t);
}

// Values per thread (log2) of the plain NTT kernels, per degree and
// direction: 4 (16 values, N/16 threads) until fhehip_ntt_autotune picks the
// faster of 4 and 5 by timing both; the environment variable FHEHIP_NTT_RE=5
// forces 5.
static std::atomic<int> g_re_fwd[16], g_re_inv[16];

static int forced_re() {
    static int re = [] {
        const char* e = getenv("FHEHIP_NTT_RE");
        return (e && atoi(e) == 5) ? 5 : 0;
    }();
    return re;
}
static int ntt_re(int logn, bool forward) {
    if (forced_re()) return forced_re();
    const int r = (forward ? g_re_fwd : g_re_inv)[logn].load(std::memory_order_relaxed);
    return r ? r : 4;
}

template <int LOGN>
static void fwd_launch(const DeviceContext& c, RowMap m, u32 rows, bool lazy, hipStream_t s) {
    dim3 grid(rows);
    const ulonglong2* tw = c.fwd_tw.as<ulonglong2>();
    const ModInfo* md = c.mod.as<ModInfo>();
    if (ntt_re(LOGN, true) == 5) {
        dim3 block(NttShape<LOGN, 5>::T);
        if (lazy) hipLaunchKernelGGL((ntt_fwd_kernel<LOGN, true, 5>), grid, block, 0, s, m, tw, md);
        else hipLaunchKernelGGL((ntt_fwd_kernel<LOGN, false, 5>), grid, block, 0, s, m, tw, md);
    } else {
        dim3 block(NttShape<LOGN, 4>::T);
        if (lazy) hipLaunchKernelGGL((ntt_fwd_kernel<LOGN, true, 4>), grid, block, 0, s, m, tw, md);
        else hipLaunchKernelGGL((ntt_fwd_kernel<LOGN, false, 4>), grid, block, 0, s, m, tw, md);
    }
}

template <int LOGN>
static void inv_launch(const DeviceContext& c, RowMap m, u32 rows, hipStream_t s) {
    dim3 grid(rows);
    const ulonglong2* tw = c.inv_tw.as<ulonglong2>();
    const ModInfo* md = c.mod.as<ModInfo>();
    if (ntt_re(LOGN, false) == 5)
        hipLaunchKernelGGL((ntt_inv_kernel<LOGN, 5>), grid, dim3(NttShape<LOGN, 5>::T), 0, s, m, tw, md);
    else
        hipLaunchKernelGGL((ntt_inv_kernel<LOGN, 4>), grid, dim3(NttShape<LOGN, 4>::T), 0, s, m, tw, md);
}

#define FHEHIP_LOGN_SWITCH(logn, F, ...)                                    \
    switch (logn) {                                                         \
        case 3: F<3>(__VA_ARGS__); break;                                   \
        case 4: F<4>(__VA_ARGS__); break;                                   \
        case 5: F<5>(__VA_ARGS__); break;                                   \
        case 6: F<6>(__VA_ARGS__); break;                                   \
        case 7: F<7>(__VA_ARGS__); break;                                   \
        case 8: F<8>(__VA_ARGS__); break;                                   \
        case 9: F<9>(__VA_ARGS__); break;                                   \
        case 10: F<10>(__VA_ARGS__); break;                                 \
        case 11: F<11>(__VA_ARGS__); break;                                 \
        case 12: F<12>(__VA_ARGS__); break;                                 \
        case 13: F<13>(__VA_ARGS__); break;                                 \
        case 14: F<14>(__VA_ARGS__); break;                                 \
        default: throw Error(-2, "degree not supported by the NTT kernels"); \
    }

// Rows transformed as two halves: the first forward stage (pairs j,
// j + N/2 with omega_1; Harvey butterfly of native.rs:256-269, outputs < 4p).
// grid = (N / 512, rows)
__global__ __launch_bounds__(256) void ntt_split_fwd_kernel(RowMap m, const ulonglong2* __restrict__ split,
                                                            const ModInfo* __restrict__ mods, u32 logn) {
    u32 limb;
    u64* row = row_ptr(m, blockIdx.y, logn, limb);
    const u32 j = blockIdx.x * 256 + threadIdx.x, half = 1u << (logn - 1);
    const ulonglong2 w = split[3 * limb];
    const u64 p = mods[limb].p;
    u64 x = row[j], y = row[j + half];
    fwd_bfly<true>(x, y, w.x, w.y, p, 2 * p);
    row[j] = x;
    row[j + half] = y;
}

// The last inverse stage on the two reduced half transforms, with the N^-1
// factor folded in: X = (x + y) N^-1, Y = (x - y) zeta N^-1 (native.rs
// :288-300 and :130-131). grid = (N / 512, rows)
__global__ __launch_bounds__(256) void ntt_split_inv_kernel(RowMap m, const ulonglong2* __restrict__ split,
                                                            const ModInfo* __restrict__ mods, u32 logn) {
    u32 limb;
    u64* row = row_ptr(m, blockIdx.y, logn, limb);
    const u32 j = blockIdx.x * 256 + threadIdx.x, half = 1u << (logn - 1);
    const ulonglong2 z = split[3 * limb + 1], ni = split[3 * limb + 2];
    const u64 p = mods[limb].p;
    const u64 x = row[j], y = row[j + half];
    row[j] = mul_shoup_u(x + y, ni.x, ni.y, p);
    row[j + half] = mul_shoup_u(x + 2 * p - y, z.x, z.y, p);
}

static std::atomic<u32> g_split_rows{128};

u32 split_row_limit() { return g_split_rows.load(std::memory_order_relaxed); }
void set_split_row_limit(u32 rows) { g_split_rows.store(rows, std::memory_order_relaxed); }

bool split_rows(const DeviceContext& c, size_t rows) {
    if (c.logn == 15) return true;
    return c.logn >= 12 && c.logn <= 14 && c.half && rows <= split_row_limit();
}

// N = 32768 digit lift (rq/mod.rs:518-540) with its whole lazy forward
// transform in one kernel: the workgroup of half h of (instance b, digit i,
// key limb j) reads both halves of lazy_reduce_{q'_j}(src_i), forms half h of
// the stage-0 butterfly on load and runs the remaining 14 stages on it with
// the half tables; outputs < 2 q'_j, as lift_ntt_kernel's. 1-D grid over
// (instance, digit, key limb, half), half fastest: the two halves of one row
// and the key limbs of one digit read the same source row and run on one XCD.
//
// The last 2 fr.rows workgroups of the grid transform the rows of `fr`
// instead (fwd_rows_half_body), so that independent rows share the launch.
template <int LOGH>
__device__ __forceinline__ void fwd_rows_half_body(u64* lds, const HalfFwdRows& fr, u32 l, int t) {
    using S = NttShape<LOGH>;
    const u32 h = l & 1, r = l >> 1, i = r % fr.L;
    const u64 p = fr.mods[i].p, p2 = 2 * p;
    u64* row = fr.x + ((size_t)r << (LOGH + 1)) + ((size_t)h << LOGH);
    u64 x[S::E];
#pragma unroll
    for (int e = 0; e < S::E; e++) x[e] = row[fwd_index<LOGH, 0>(t, e)];
    ntt_fwd_regs<LOGH>(lds, x, t, fr.half_tw + ((size_t)(2 * i + h) << LOGH), p);
    store_fwd_order<LOGH>(lds, x, t, [&](int k, u64 v) { row[k] = reduce1(reduce1(v, p2), p); });
}

template <int LOGH>
__global__ __launch_bounds__(NttShape<LOGH>::T) void lift_ntt_half_kernel(const u64* __restrict__ src, size_t src_inst,
                                                                       u64* __restrict__ y,
                                                                       const ulonglong2* __restrict__ half_tw,
                                                                       const ModInfo* __restrict__ mods,
                                                                       const ulonglong2* __restrict__ split,
                                                                       u32 digits, u32 key_limbs, u32 log_base,
                                                                       HalfFwdRows fr) {
    using S = NttShape<LOGH>;
    __shared__ u64 lds[S::LDS_WORDS];
    const int t = threadIdx.x;
    const u32 l = xcd_block(blockIdx.x, gridDim.x);
    const u32 lift_blocks = gridDim.x - 2 * fr.rows;
    if (l >= lift_blocks) {
        fwd_rows_half_body<LOGH>(lds, fr, l - lift_blocks, t);
        return;
    }
    const u32 h = l & 1, r = l >> 1;
    const u32 j = r % key_limbs, i = (r / key_limbs) % digits, b = r / (key_limbs * digits);
    const ModInfo mi = mods[j];
    const u64 p = mi.p, p2 = 2 * p;
    const ulonglong2 w = split[3 * j];
    const u64* row = src + b * src_inst + ((size_t)(log_base ? 0 : i) << (LOGH + 1));
    u64 x[S::E];
    // Digit i is below q_i (a prefix of the key moduli); below 4 q'_j it is
    // already a valid lazy transform input (as in lift_ntt_kernel), and so is
    // a power-of-two decomposition digit.
    if (log_base) {
        const u32 sh = i * log_base;
        const u64 mask = ((u64)1 << log_base) - 1;
#pragma unroll
        for (int e = 0; e < S::E; e++) {
            const int k = fwd_index<LOGH, 0>(t, e);
            x[e] = split_half((row[k] >> sh) & mask, (row[k + (1 << LOGH)] >> sh) & mask, h, w.x, w.y, p, p2);
        }
    } else if (mods[i].p <= 4 * p) {
#pragma unroll
        for (int e = 0; e < S::E; e++) {
            const int k = fwd_index<LOGH, 0>(t, e);
            x[e] = split_half(row[k], row[k + (1 << LOGH)], h, w.x, w.y, p, p2);
        }
    } else {
#pragma unroll
        for (int e = 0; e < S::E; e++) {
            const int k = fwd_index<LOGH, 0>(t, e);
            const u64 a = lazy_reduce_u64(row[k], mi.barrett_lo, mi.barrett_hi, p);
            const u64 c = lazy_reduce_u64(row[k + (1 << LOGH)], mi.barrett_lo, mi.barrett_hi, p);
            x[e] = split_half(a, c, h, w.x, w.y, p, p2);
        }
    }
    ntt_fwd_regs<LOGH>(lds, x, t, half_tw + ((size_t)(2 * j + h) << LOGH), p);
    u64* dst = y + ((((size_t)b * digits + i) * key_limbs + j) << (LOGH + 1)) + ((size_t)h << LOGH);
    store_fwd_order<LOGH>(lds, x, t, [&](int k, u64 v) { dst[k] = reduce1(v, p2); });
}

template <int LOGH>
static void lift_half_go(dim3 grid, hipStream_t s, const u64* src, size_t src_inst, u64* y, const ulonglong2* tw,
                         const ModInfo* mods, const ulonglong2* split, u32 digits, u32 key_limbs, u32 log_base,
                         const HalfFwdRows& fr) {
    hipLaunchKernelGGL(lift_ntt_half_kernel<LOGH>, grid, dim3(NttShape<LOGH>::T), 0, s, src, src_inst, y, tw, mods,
                       split, digits, key_limbs, log_base, fr);
}

void launch_lift_ntt_split(const DeviceContext& key, const u64* src, size_t src_inst, u64* y, u32 digits, u32 inst,
                           u32 log_base, hipStream_t s, const HalfFwdRows* rows) {
    if (!key.half) throw Error(-2, "this context has no half-row tables");
    if (!digits || !inst) return;
    HalfFwdRows fr{};
    if (rows && rows->rows) {
        if (rows->L == 0) throw Error(-1, "rows to transform without a context");
        fr = *rows;
    }
    ProfScope ps("lift_ntt", key.logn, s);
    const dim3 grid(2 * (inst * digits * key.L + fr.rows));
    const ulonglong2* tw = key.half->fwd_tw.as<ulonglong2>();
    const ModInfo* md = key.mod.as<ModInfo>();
    const ulonglong2* sp = key.split.as<ulonglong2>();
    switch (key.logn - 1) {
        case 11: lift_half_go<11>(grid, s, src, src_inst, y, tw, md, sp, digits, key.L, log_base, fr); break;
        case 12: lift_half_go<12>(grid, s, src, src_inst, y, tw, md, sp, digits, key.L, log_base, fr); break;
        case 13: lift_half_go<13>(grid, s, src, src_inst, y, tw, md, sp, digits, key.L, log_base, fr); break;
        case 14: lift_half_go<14>(grid, s, src, src_inst, y, tw, md, sp, digits, key.L, log_base, fr); break;
        default: throw Error(-2, "half-row transforms need a degree of 4096 .. 32768");
    }
    HIPCHK(hipGetLastError());
}

static RowMap half_rows(RowMap m) {
    return RowMap{m.data, 2 * m.rows_per_poly, 2 * m.limbs_per_poly, 2 * m.first_limb, 2 * m.table_offset};
}

// The transforms of rows longer than one workgroup holds (degree 32768 and
// up), or of small launches at 4096 .. 16384: the stage pairing the halves as
// one elementwise pass, and the halves as rows of the half context, split
// again while they are 32768 or more coefficients long.
static void ntt_forward_rows(const DeviceContext& c, RowMap m, u32 rows, bool lazy, hipStream_t s) {
    if (c.logn >= 15 || split_rows(c, rows)) {
        hipLaunchKernelGGL(ntt_split_fwd_kernel, dim3((c.N / 2) / 256, rows), dim3(256), 0, s, m,
                           c.split.as<ulonglong2>(), c.mod.as<ModInfo>(), (u32)c.logn);
        HIPCHK(hipGetLastError());
        ntt_forward_rows(*c.half, half_rows(m), 2 * rows, lazy, s);
    } else {
        FHEHIP_LOGN_SWITCH(c.logn, fwd_launch, c, m, rows, lazy, s);
    }
    HIPCHK(hipGetLastError());
}

static void ntt_backward_rows(const DeviceContext& c, RowMap m, u32 rows, hipStream_t s) {
    if (c.logn >= 15 || split_rows(c, rows)) {
        ntt_backward_rows(*c.half, half_rows(m), 2 * rows, s);
        hipLaunchKernelGGL(ntt_split_inv_kernel, dim3((c.N / 2) / 256, rows), dim3(256), 0, s, m,
                           c.split.as<ulonglong2>(), c.mod.as<ModInfo>(), (u32)c.logn);
    } else {
        FHEHIP_LOGN_SWITCH(c.logn, inv_launch, c, m, rows, s);
    }
    HIPCHK(hipGetLastError());
}

void launch_ntt_forward(const DeviceContext& c, RowMap m, u32 rows, bool lazy, hipStream_t s) {
    if (rows == 0) return;
    ProfScope ps(lazy ? "ntt_forward_lazy" : "ntt_forward", c.logn, s);
    ntt_forward_rows(c, m, rows, lazy, s);
}

void launch_ntt_forward_halves(const DeviceContext& c, RowMap m, u32 rows, bool lazy, hipStream_t s) {
    if (rows == 0) return;
    if (!c.half) throw Error(-2, "this context has no half-row tables");
    ProfScope ps(lazy ? "ntt_forward_lazy" : "ntt_forward", c.logn, s);
    ntt_forward_rows(*c.half, half_rows(m), 2 * rows, lazy, s);
}

void launch_ntt_backward_halves(const DeviceContext& c, RowMap m, u32 rows, hipStream_t s) {
    if (rows == 0) return;
    if (!c.half) throw Error(-2, "this context has no half-row tables");
    ProfScope ps("ntt_backward", c.logn, s);
    ntt_backward_rows(*c.half, half_rows(m), 2 * rows, s);
}

void launch_ntt_split_inv(const DeviceContext& c, RowMap m, u32 rows, hipStream_t s) {
    if (rows == 0) return;
    if (!c.half) throw Error(-2, "this context has no half-row tables");
    ProfScope ps("ntt_split_inv", c.logn, s);
    hipLaunchKernelGGL(ntt_split_inv_kernel, dim3((c.N / 2) / 256, rows), dim3(256), 0, s, m,
                       c.split.as<ulonglong2>(), c.mod.as<ModInfo>(), (u32)c.logn);
    HIPCHK(hipGetLastError());
}

void launch_ntt_backward(const DeviceContext& c, RowMap m, u32 rows, hipStream_t s) {
    if (rows == 0) return;
    ProfScope ps("ntt_backward", c.logn, s);
    ntt_backward_rows(c, m, rows, s);
}

// Times forward and inverse transforms of `rows` rows of `c` (the half
// transforms at degree 32768) with 16 and 32 values per thread and keeps the
// faster shape per direction for this degree. Returns the chosen log2 values
// per thread and the measured milliseconds per launch [fwd16, fwd32, inv16, inv32].
void autotune_ntt(const DeviceContext& c, u32 rows, int& re_fwd, int& re_inv, float (&ms)[4]) {
    // rows longer than 16384 coefficients: the shape of their pieces
    const DeviceContext* tp = &c;
    u32 trows = rows;
    while (tp->logn > 14) {
        tp = tp->half.get();
        trows *= 2;
    }
    const DeviceContext& t = *tp;
    const int logn = t.logn;
    if (logn < 5) {  // one shape only
        re_fwd = re_inv = 4;
        for (float& m : ms) m = 0;
        return;
    }
    DevBuf buf((size_t)trows * t.N * 8);
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    try {
        HIPCHK(hipMemsetAsync(buf.ptr, 0, buf.bytes, s));  // zero rows: valid residues
        const RowMap m{buf.as<u64>(), t.L, t.L, 0, 0};
        const int reps = 5;
        for (int dir = 0; dir < 2; dir++) {
            for (int k = 0; k < 2; k++) {
                const int re = 4 + k;
                auto& slot = (dir == 0 ? g_re_fwd : g_re_inv)[logn];
                const int saved = slot.load();
                slot.store(re);
                auto run = [&] {
                    if (dir == 0) {
                        FHEHIP_LOGN_SWITCH(logn, fwd_launch, t, m, trows, false, s);
                    } else {
                        FHEHIP_LOGN_SWITCH(logn, inv_launch, t, m, trows, s);
                    }
                };
                run();
                run();
                HIPCHK(hipEventRecord(e0, s));
                for (int r = 0; r < reps; r++) run();
                HIPCHK(hipEventRecord(e1, s));
                HIPCHK(hipEventSynchronize(e1));
                HIPCHK(hipGetLastError());
                float el = 0;
                HIPCHK(hipEventElapsedTime(&el, e0, e1));
                ms[2 * dir + k] = el / reps;
                slot.store(saved);
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
        throw;
    }
    HIPCHK(hipStreamSynchronize(s));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    re_fwd = ms[1] < ms[0] ? 5 : 4;
    re_inv = ms[3] < ms[2] ? 5 : 4;
    g_re_fwd[logn].store(re_fwd);
    g_re_inv[logn].store(re_inv);
}

void set_ntt_shape(int logn, int re_fwd, int re_inv) {
    if (logn > 14 && logn <= 17) logn = 14;  // the pieces of longer rows
    if (logn < 3 || logn > 14 || (re_fwd != 4 && re_fwd != 5) || (re_inv != 4 && re_inv != 5))
        throw Error(-1, "NTT shape: degree 8..131072, 4 or 5 log2 values per thread");
    g_re_fwd[logn].store(re_fwd);
    g_re_inv[logn].store(re_inv);
}

}  // namespace fhehip
