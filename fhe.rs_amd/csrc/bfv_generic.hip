// Elementwise kernels of the BFV path: products of any part counts, the key
// inner product and switch-down steps of deep keys, and job staging. One
// thread per coefficient; grid.y / grid.z index rows.
#include "bfv_launch.h"
#include "profiler.h"

namespace fhehip {

static inline dim3 ew_grid(u32 n, u32 y, u32 z) { return dim3((n + 255) / 256, y, z); }

// grid = (N/256, M, num); each thread forms every output part of its
// coefficient. At most 8 products meet in one part (lp + rp - 1 <= 16), each
// below q^2 < 2^124, so the 128-bit sum is exact.
__global__ __launch_bounds__(256) void tensor_parts_kernel(const u64* __restrict__ a, u32 lp, const u64* __restrict__ b,
                                                           u32 rp, u64* __restrict__ c,
                                                           const ModInfo* __restrict__ mods, u32 M, u32 logn) {
    const u32 e = blockIdx.x * 256 + threadIdx.x;
    if (e >> logn) return;
    const u32 m = blockIdx.y, inst = blockIdx.z;
    const ModInfo mi = mods[m];
    const size_t row = (size_t)M << logn, off = ((size_t)m << logn) + e;
    const u64* ab = a + (size_t)inst * lp * row + off;
    const u64* bb = b + (size_t)inst * rp * row + off;
    u64* cb = c + (size_t)inst * (lp + rp - 1) * row + off;
    for (u32 k = 0; k + 1 < lp + rp; k++) {
        u64 lo = 0, hi = 0;
        for (u32 i = k + 1 > rp ? k + 1 - rp : 0; i < lp && i <= k; i++) {
            u64 l2, h2;
            mul_wide(ab[i * row], bb[(k - i) * row], l2, h2);
            add128(lo, hi, l2, h2);
        }
        cb[k * row] = reduce_u128(lo, hi, mi.barrett_lo, mi.barrett_hi, mi.p);
    }
}

void launch_tensor_parts(const u64* a, u32 lp, const u64* b, u32 rp, u64* c, const ModInfo* mods, u32 M, int logn,
                         u32 num, hipStream_t s) {
    if (!num || !M) return;
    ProfScope ps("tensor_parts", logn, s);
    hipLaunchKernelGGL(tensor_parts_kernel, ew_grid(1u << logn, M, num), dim3(256), 0, s, a, lp, b, rp, c, mods, M,
                       (u32)logn);
    HIPCHK(hipGetLastError());
}

// Key inner product (key_switching_key.rs:227-239) of key limbs first ..
// first + count - 1: acc[b][p][j - first] = sum_i Y[b][i][j] * key_p[i][j]
// (Shoup, exact). grid = (N/256, count * 2, instances)
__global__ __launch_bounds__(256) void mac_ew_kernel(const u64* __restrict__ y, u32 digits, u32 key_limbs, u32 first,
                                                     u32 count, const ulonglong2* __restrict__ key0,
                                                     const ulonglong2* __restrict__ key1,
                                                     const ModInfo* __restrict__ key_mods, u64* __restrict__ acc,
                                                     u32 logn) {
    const u32 e = blockIdx.x * 256 + threadIdx.x;
    if (e >> logn) return;
    const u32 jj = blockIdx.y % count, part = blockIdx.y / count, b = blockIdx.z;
    const u32 j = first + jj;
    const u64 p = key_mods[j].p;
    const ulonglong2* key = part ? key1 : key0;
    u64 r = 0;
    for (u32 i = 0; i < digits; i++) {
        const u64 v = y[((((size_t)b * digits + i) * key_limbs + j) << logn) + e];
        const ulonglong2 w = key[(((size_t)i * key_limbs + j) << logn) + e];
        r = add_mod(r, mul_shoup(v, w.x, w.y, p), p);
    }
    acc[((((size_t)b * 2 + part) * count + jj) << logn) + e] = r;
}

void launch_mac_ew(const u64* y, u32 digits, u32 key_limbs, u32 first, u32 count, const ulonglong2* key0,
                   const ulonglong2* key1, const ModInfo* key_mods, u64* acc, int logn, u32 inst, hipStream_t s) {
    if (!count || !inst) return;
    ProfScope ps("mac_ew", logn, s);
    hipLaunchKernelGGL(mac_ew_kernel, ew_grid(1u << logn, 2 * count, inst), dim3(256), 0, s, y, digits, key_limbs,
                       first, count, key0, key1, key_mods, acc, (u32)logn);
    HIPCHK(hipGetLastError());
}

// Poly::switch_down (math/src/rq/mod.rs:390-449) in power basis: with
// y = (x_l + floor(q_l / 2)) mod q_l, x_i <- (x_i - y + floor(q_l / 2)) q_l^-1
// mod q_i for each remaining limb i < L - 1. Polynomial b reads limbs at
// in + b in_stride rows and writes at out + b out_stride rows; out may be in
// (each thread rewrites only the limb it read, and limb L - 1 is not
// written). grid = (N / 256, L - 1, polys)
__global__ __launch_bounds__(256) void switch_down_pb_kernel(const u64* in, u64* out, const SdConst* __restrict__ consts,
                                                             u64 ql, u32 L, u32 in_stride, u32 out_stride, u32 logn) {
    const u32 e = blockIdx.x * 256 + threadIdx.x;
    if (e >> logn) return;
    const u32 i = blockIdx.y, b = blockIdx.z;
    const u64* src = in + ((size_t)b * in_stride << logn);
    u64 y = src[((size_t)(L - 1) << logn) + e] + (ql >> 1);
    y = y >= ql ? y - ql : y;
    const SdConst cst = consts[i];
    const u64 q = cst.q;
    // [y]_i only has to be < 2 q_i and congruent to y (ks_finish_kernel)
    const u64 yr = ql <= 2 * q ? y : lazy_reduce_u64(y, cst.bl, cst.bh, q);
    const u64 tmp = yr + cst.half_mod;
    out[(((size_t)b * out_stride + i) << logn) + e] =
        mul_shoup_u(src[((size_t)i << logn) + e] + 3 * q - tmp, cst.inv, cst.inv_shoup, q);
}

void launch_switch_down_pb(const u64* in, u64* out, const SdConst* consts, u64 ql, u32 L, u32 in_stride,
                           u32 out_stride, int logn, u32 polys, hipStream_t s) {
    if (L < 2 || !polys) return;
    ProfScope ps("switch_down", logn, s);
    hipLaunchKernelGGL(switch_down_pb_kernel, ew_grid(1u << logn, L - 1, polys), dim3(256), 0, s, in, out, consts,
                       ql, L, in_stride, out_stride, (u32)logn);
    HIPCHK(hipGetLastError());
}

// The two elementwise ends of the key-switch finish after a switch-down
// chain of any length (key context any number of moduli above the
// ciphertext context, relinearization_key.rs:84-91, galois_key.rs:70-77):
// before the forward NTT, x[b][p][i] = ks[b][p][i] (the switched-down limbs,
// in place in the key-switch rows) + add_pb; after it, out[b][p][i] =
// x[b][p][i] + add_ntt[b][p][i][perm] for parts below add_ntt_parts.
// x is contiguous [inst][2][L][N]. grid = (N/256, 2 L, instances)
__global__ __launch_bounds__(256) void sd_tail_pre_kernel(SdNttArgs a, u64* __restrict__ x, u32 logn) {
    const u32 e = blockIdx.x * 256 + threadIdx.x;
    if (e >> logn) return;
    const u32 i = blockIdx.y % a.L, part = blockIdx.y / a.L, b = blockIdx.z;
    u64 z = a.ks[(size_t)b * a.ks_inst_stride + (size_t)part * a.ks_part_stride + ((size_t)i << logn) + e];
    if (a.add_pb)
        z = add_mod(z,
                    a.add_pb[(size_t)b * a.add_inst_stride + (size_t)part * a.add_part_stride + ((size_t)i << logn) + e],
                    a.mods[i].p);
    x[((((size_t)b * 2 + part) * a.L + i) << logn) + e] = z;
}

__global__ __launch_bounds__(256) void sd_tail_post_kernel(SdNttArgs a, const u64* __restrict__ x, u32 logn) {
    const u32 e = blockIdx.x * 256 + threadIdx.x;
    if (e >> logn) return;
    const u32 i = blockIdx.y % a.L, part = blockIdx.y / a.L, b = blockIdx.z;
    u64 z = x[((((size_t)b * 2 + part) * a.L + i) << logn) + e];
    if (a.add_ntt && part < a.add_ntt_parts) {
        const u32 src = a.perm ? a.perm[e] : e;
        z = add_mod(z,
                    a.add_ntt[(size_t)b * a.add_ntt_inst_stride + (size_t)part * a.add_ntt_part_stride +
                              ((size_t)i << logn) + src],
                    a.mods[i].p);
    }
    a.out[(size_t)b * a.out_inst_stride + (size_t)part * a.out_part_stride + ((size_t)i << logn) + e] = z;
}

void launch_sd_tail(int logn, const SdNttArgs& a, u64* x, bool post, u32 inst, hipStream_t s) {
    if (!inst || !a.L) return;
    ProfScope ps(post ? "sd_tail_post" : "sd_tail_pre", logn, s);
    if (post)
        hipLaunchKernelGGL(sd_tail_post_kernel, ew_grid(1u << logn, 2 * a.L, inst), dim3(256), 0, s, a, x, (u32)logn);
    else
        hipLaunchKernelGGL(sd_tail_pre_kernel, ew_grid(1u << logn, 2 * a.L, inst), dim3(256), 0, s, a, x, (u32)logn);
    HIPCHK(hipGetLastError());
}

}  // namespace fhehip

namespace fhehip {

// dst[k][e] = src[k][e] for `count` (src, dst) pairs of `elems` words each,
// read from a device table of pointer pairs. grid = (ceil(elems / 1024), count)
__global__ __launch_bounds__(256) void copy_slots_kernel(const u64* const* __restrict__ table, size_t elems) {
    const u64* src = table[2 * blockIdx.y];
    u64* dst = (u64*)table[2 * blockIdx.y + 1];
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const size_t e = base + (size_t)k * 256;
        if (e < elems) dst[e] = src[e];
    }
}

// x[e] = splitmix64(seed + e) >> shift: pseudorandom words below 2^(64 - shift),
// residues of every modulus of at least 65 - shift bits.
__global__ __launch_bounds__(256) void fill_words_kernel(u64* __restrict__ x, size_t elems, u32 shift, u64 seed) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= elems) return;
    u64 z = seed + e * 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    x[e] = (z ^ (z >> 31)) >> shift;
}

void launch_fill_words(u64* x, size_t elems, u32 shift, u64 seed, hipStream_t s) {
    if (!elems) return;
    hipLaunchKernelGGL(fill_words_kernel, dim3((unsigned)((elems + 255) / 256)), dim3(256), 0, s, x, elems, shift,
                       seed);
    HIPCHK(hipGetLastError());
}

void launch_copy_slots(const u64* const* table, u32 count, size_t elems, hipStream_t s) {
    if (!count || !elems) return;
    ProfScope ps("copy_slots", 0, s);
    hipLaunchKernelGGL(copy_slots_kernel, dim3((unsigned)((elems + 1023) / 1024), count), dim3(256), 0, s, table,
                       elems);
    HIPCHK(hipGetLastError());
}

}  // namespace fhehip
