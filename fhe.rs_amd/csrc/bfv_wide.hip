// Elementwise kernels of the BFV path at degrees above 32768 (wide rows).
// There a row is longer than one workgroup can transform, so the work that
// the fused row kernels do on load (the Galois gather, the tensor products,
// the digit lift) runs as its own pass, followed by the plain transforms
// (ntt_kernels.hip, split into halves until the pieces fit a workgroup).
// One thread per coefficient; grid.y / grid.z index rows.
#include "bfv_launch.h"
#include "profiler.h"

namespace fhehip {

// dst rows = src rows, or their NTT-domain gather q[m] = p[perm[m]]
// (Poly::substitute in the Ntt representation, math/src/rq/mod.rs:324-346),
// ahead of an in-place inverse transform of dst (intt_rows_kernel's load).
// dst must not overlap src. grid = (N / 256, limbs * parts, instances)
__global__ __launch_bounds__(256) void rows_gather_kernel(InttRowsArgs a, u32 limbs, u32 logn) {
    const u32 e = blockIdx.x * 256 + threadIdx.x;
    if (e >> logn) return;
    const u32 l = blockIdx.y % limbs, k = blockIdx.y / limbs, b = blockIdx.z;
    const u64* src = a.src[k] + (size_t)b * a.src_inst_stride + ((size_t)l << logn);
    u64* dst = a.dst + (size_t)b * a.dst_inst_stride + (size_t)k * a.dst_part_stride + ((size_t)l << logn);
    dst[e] = src[a.perm ? a.perm[e] : e];
}

void launch_rows_gather(const InttRowsArgs& a, u32 limbs, u32 parts, u32 inst, int logn, hipStream_t s) {
    if (!limbs || !parts || !inst) return;
    ProfScope ps(a.perm ? "rows_gather" : "rows_copy", logn, s);
    hipLaunchKernelGGL(rows_gather_kernel, dim3(((1u << logn) + 255) / 256, limbs * parts, inst), dim3(256), 0, s, a,
                       limbs, (u32)logn);
    HIPCHK(hipGetLastError());
}

// Tensor product of Multiplicator::multiply (fhe/src/bfv/ops/mul.rs:181-185)
// at one coefficient of limb m: c0 = a0 b0, c1 = a0 b1 + a1 b0, c2 = a1 b1
// mod q_m, exact and reduced (a valid input of the inverse transform), with
// the inputs read as tensor_intt_kernel reads them (limbs below `common` from
// the ciphertexts, the others from the extension rows).
// grid = (N / 256, M, instances)
__global__ __launch_bounds__(256) void tensor_ew_kernel(TensorArgs a, u32 logn) {
    const u32 e = blockIdx.x * 256 + threadIdx.x;
    if (e >> logn) return;
    const u32 m = blockIdx.y, b = blockIdx.z;
    const ModInfo mi = a.mods[m];
    u64 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const u32 common = k < 2 ? a.common_lhs : a.common_rhs;
        const u64* row;
        if (m < common) {
            const u64* base = k < 2 ? a.lhs : a.rhs;
            row = base + (size_t)b * a.in_inst_stride + (size_t)(k & 1) * a.in_part_stride + ((size_t)m << logn);
        } else {
            row = a.ext + (size_t)b * a.ext_inst_stride + (size_t)k * a.ext_part_stride +
                  ((size_t)(m - common) << logn);
        }
        v[k] = row[e];
    }
    u64* out = a.out + (size_t)b * a.out_inst_stride + ((size_t)m << logn) + e;
    out[0] = mul_mod(v[0], v[2], mi.barrett_lo, mi.barrett_hi, mi.p);
    u64 lo, hi, l2, h2;
    mul_wide(v[0], v[3], lo, hi);
    mul_wide(v[1], v[2], l2, h2);
    add128(lo, hi, l2, h2);  // two products below q^2 < 2^124: exact
    out[a.out_part_stride] = reduce_u128(lo, hi, mi.barrett_lo, mi.barrett_hi, mi.p);
    out[2 * a.out_part_stride] = mul_mod(v[1], v[3], mi.barrett_lo, mi.barrett_hi, mi.p);
}

void launch_tensor_ew(const TensorArgs& a, u32 M, u32 inst, int logn, hipStream_t s) {
    if (!M || !inst) return;
    ProfScope ps("tensor_ew", logn, s);
    hipLaunchKernelGGL(tensor_ew_kernel, dim3(((1u << logn) + 255) / 256, M, inst), dim3(256), 0, s, a, (u32)logn);
    HIPCHK(hipGetLastError());
}

// Digit lift of key switching (rq/mod.rs:518-540 inside key_switching_key.rs
// :227-239): Y[b][i][j] = digit i of the input reduced modulo key limb j, as
// lift_ntt_kernel forms it before its transform (below 4 q'_j; a
// power-of-two decomposition digit when log_base is set).
// grid = (N / 256, digits * key_limbs, instances)
__global__ __launch_bounds__(256) void lift_ew_kernel(KeySwitchArgs a, u64* __restrict__ y, u32 logn) {
    const u32 e = blockIdx.x * 256 + threadIdx.x;
    if (e >> logn) return;
    const u32 j = blockIdx.y % a.key_limbs, i = blockIdx.y / a.key_limbs, b = blockIdx.z;
    const ModInfo mi = a.key_mods[j];
    const u64 v = a.src[(size_t)b * a.src_inst_stride + ((size_t)(a.log_base ? 0 : i) << logn) + e];
    u64 x;
    if (a.log_base) x = (v >> (i * a.log_base)) & (((u64)1 << a.log_base) - 1);
    else if (a.key_mods[i].p <= 4 * mi.p) x = v;
    else x = lazy_reduce_u64(v, mi.barrett_lo, mi.barrett_hi, mi.p);
    y[((((size_t)b * a.digits + i) * a.key_limbs + j) << logn) + e] = x;
}

void launch_lift_ew(const KeySwitchArgs& a, u64* y, u32 inst, int logn, hipStream_t s) {
    if (!a.digits || !a.key_limbs || !inst) return;
    ProfScope ps("lift_ew", logn, s);
    hipLaunchKernelGGL(lift_ew_kernel, dim3(((1u << logn) + 255) / 256, a.digits * a.key_limbs, inst), dim3(256), 0,
                       s, a, y, (u32)logn);
    HIPCHK(hipGetLastError());
}

}  // namespace fhehip
