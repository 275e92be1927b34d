// Kernels on either side of the evaluation path: plaintext encoding and
// scaling, secret-key encryption and decryption. One thread per coefficient;
// grid.y indexes limbs, grid.z polynomials. All results are exact residues.
#include "bfv_launch.h"
#include "profiler.h"

namespace fhehip {

static inline dim3 cgrid(int logn, u32 y, u32 z) { return dim3(((1u << logn) + 255) / 256, y, z); }

__device__ __forceinline__ u64 reduce_u64(u64 a, const CodecConst& c) {
    return reduce1(lazy_reduce_u64(a, c.bl, c.bh, c.q), c.q);
}

// Values [num][N] into rows of the plaintext-modulus context: SIMD slots go
// to row[index_map[i]] (plaintext_vec.rs:63-67), coefficients are copied.
__global__ __launch_bounds__(256) void encode_gather_kernel(const u64* __restrict__ v, const u32* __restrict__ map,
                                                            u64* __restrict__ rows, u32 logn) {
    const u32 i = blockIdx.x * 256 + threadIdx.x;
    if (i >> logn) return;
    const size_t b = (size_t)blockIdx.z << logn;
    rows[b + (map ? map[i] : i)] = v[b + i];
}

// Row of values m (< t) -> limb i: with `scaled`, Plaintext::to_poly before
// its NTT (plaintext.rs:71-98): ([m q_mod_t]_t reduced mod q_i) * delta_i,
// the NTT-domain constant delta moved before the (linear) transform; else
// Plaintext::poly_ntt before its NTT: m mod q_i.
__global__ __launch_bounds__(256) void pt_lift_kernel(const u64* __restrict__ m, u64* __restrict__ out,
                                                      const CodecConst* __restrict__ cc, u64 t, u64 qmt,
                                                      u64 qmt_shoup, u32 L, u32 logn, u32 scaled) {
    const u32 e = blockIdx.x * 256 + threadIdx.x;
    if (e >> logn) return;
    const u32 i = blockIdx.y, b = blockIdx.z;
    const CodecConst c = cc[i];
    const u64 x = m[((size_t)b << logn) + e];
    u64 r;
    if (scaled)
        r = mul_shoup(reduce_u64(mul_shoup(x, qmt, qmt_shoup, t), c), c.delta, c.delta_shoup, c.q);
    else
        r = reduce_u64(x, c);
    out[(((size_t)b * L + i) << logn) + e] = r;
}

// Signed coefficients [num][N] -> [num][L][N] residues (Poly::try_convert_from
// of i64 values, rq/convert.rs:296-318).
__global__ __launch_bounds__(256) void small_lift_kernel(const int64_t* __restrict__ e, u64* __restrict__ out,
                                                         const CodecConst* __restrict__ cc, u32 L, u32 logn) {
    const u32 j = blockIdx.x * 256 + threadIdx.x;
    if (j >> logn) return;
    const u32 i = blockIdx.y, b = blockIdx.z;
    const CodecConst c = cc[i];
    const int64_t v = e[((size_t)b << logn) + j];
    const u64 mag = v < 0 ? 0 - (u64)v : (u64)v;
    const u64 r = reduce_u64(mag, c);
    out[(((size_t)b * L + i) << logn) + j] = v < 0 && r ? c.q - r : r;
}

// SecretKey::encrypt_poly (secret_key.rs:102-137): ct = [e - a s + pt, a],
// NTT domain. grid = (N/256, L, num)
__global__ __launch_bounds__(256) void encrypt_sk_kernel(const u64* __restrict__ pt, const u64* __restrict__ a,
                                                         const u64* __restrict__ e, const u64* __restrict__ s,
                                                         u64* __restrict__ ct, const CodecConst* __restrict__ cc,
                                                         u32 L, u32 logn) {
    const u32 j = blockIdx.x * 256 + threadIdx.x;
    if (j >> logn) return;
    const u32 i = blockIdx.y, b = blockIdx.z;
    const CodecConst c = cc[i];
    const size_t w = (((size_t)b * L + i) << logn) + j;
    const u64 av = a[w];
    const u64 as = mul_mod(av, s[((size_t)i << logn) + j], c.bl, c.bh, c.q);
    const size_t o = ((((size_t)b * 2) * L + i) << logn) + j;
    ct[o] = add_mod(sub_mod(e[w], as, c.q), pt[w], c.q);
    ct[o + ((size_t)L << logn)] = av;
}

// PublicKey::try_encrypt (public_key.rs:47-87): ct = [u pk0 + e1 + pt,
// u pk1 + e2] with u, e1, e2 = sm[0..3][b] in the NTT domain. grid = (N/256, L, num)
__global__ __launch_bounds__(256) void encrypt_pk_kernel(const u64* __restrict__ pt, const u64* __restrict__ pk,
                                                         const u64* __restrict__ sm, size_t sm_stride,
                                                         u64* __restrict__ ct, const CodecConst* __restrict__ cc,
                                                         u32 L, u32 logn) {
    const u32 j = blockIdx.x * 256 + threadIdx.x;
    if (j >> logn) return;
    const u32 i = blockIdx.y, b = blockIdx.z;
    const CodecConst c = cc[i];
    const size_t w = (((size_t)b * L + i) << logn) + j;
    const size_t k = ((size_t)i << logn) + j, row = (size_t)L << logn;
    const u64 u = sm[w];
    const u64 c0 = add_mod(add_mod(mul_mod(u, pk[k], c.bl, c.bh, c.q), sm[sm_stride + w], c.q), pt[w], c.q);
    const u64 c1 = add_mod(mul_mod(u, pk[row + k], c.bl, c.bh, c.q), sm[2 * sm_stride + w], c.q);
    const size_t o = ((((size_t)b * 2) * L + i) << logn) + j;
    ct[o] = c0;
    ct[o + row] = c1;
}

void launch_encrypt_pk(const u64* pt, const u64* pk, const u64* sm, size_t sm_stride, u64* ct, const CodecConst* cc,
                       u32 L, int logn, u32 num, hipStream_t s) {
    if (!num) return;
    ProfScope ps("encrypt_pk", logn, s);
    hipLaunchKernelGGL(encrypt_pk_kernel, cgrid(logn, L, num), dim3(256), 0, s, pt, pk, sm, sm_stride, ct, cc, L,
                       (u32)logn);
    HIPCHK(hipGetLastError());
}

// The phase c0 + c1 s + c2 s^2 + ... of try_decrypt (secret_key.rs:211-223),
// NTT domain. grid = (N/256, L, num)
__global__ __launch_bounds__(256) void phase_kernel(const u64* __restrict__ ct, u32 parts, const u64* __restrict__ s,
                                                    u64* __restrict__ out, const CodecConst* __restrict__ cc, u32 L,
                                                    u32 logn) {
    const u32 j = blockIdx.x * 256 + threadIdx.x;
    if (j >> logn) return;
    const u32 i = blockIdx.y, b = blockIdx.z;
    const CodecConst c = cc[i];
    const size_t row = (size_t)L << logn;
    const u64* cb = ct + (size_t)b * parts * row + ((size_t)i << logn) + j;
    const u64 sv = s[((size_t)i << logn) + j];
    u64 acc = cb[0], si = sv;
    for (u32 k = 1; k < parts; k++) {
        acc = add_mod(acc, mul_mod(cb[k * row], si, c.bl, c.bh, c.q), c.q);
        if (k + 1 < parts) si = mul_mod(si, sv, c.bl, c.bh, c.q);
    }
    out[(((size_t)b * L + i) << logn) + j] = acc;
}

// try_decrypt's last step (secret_key.rs:235-245): w = ((d + t) mod q_0) mod t
// from the first row of the scaled polynomial [num][k][N]. grid = (N/256, 1, num)
__global__ __launch_bounds__(256) void decode_reduce_kernel(const u64* __restrict__ d, u32 k, u64* __restrict__ w,
                                                            CodecConst q0, CodecConst tc, u32 logn) {
    const u32 j = blockIdx.x * 256 + threadIdx.x;
    if (j >> logn) return;
    const u32 b = blockIdx.z;
    const u64 x = d[(((size_t)b * k) << logn) + j] + tc.q;
    w[((size_t)b << logn) + j] = reduce_u64(reduce_u64(x, q0), tc);
}

// SIMD decode (plaintext_vec.rs decode): out[i] = row[index_map[i]] after the
// forward NTT modulo t. grid = (N/256, 1, num)
__global__ __launch_bounds__(256) void decode_gather_kernel(const u64* __restrict__ rows, const u32* __restrict__ map,
                                                            u64* __restrict__ out, u32 logn) {
    const u32 i = blockIdx.x * 256 + threadIdx.x;
    if (i >> logn) return;
    const size_t b = (size_t)blockIdx.z << logn;
    out[b + i] = rows[b + map[i]];
}

// Sub / Neg for &Ciphertext (fhe/src/bfv/ops/mod.rs:110-227 over Poly
// SubAssign / Neg, math/src/rq/ops.rs): out = a - b (op 0) or -a (op 1),
// elementwise over rows where row r has modulus r % L.
__global__ __launch_bounds__(256) void rows_sub_neg_kernel(const u64* a, const u64* b, u64* out,
                                                           const ModInfo* __restrict__ mods, u32 L, u32 logn,
                                                           size_t total, u32 op) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const u64 p = mods[(i >> logn) % L].p;
    out[i] = op == 0 ? sub_mod(a[i], b[i], p) : (a[i] ? p - a[i] : 0);
}

// MulAssign<&Plaintext> for Ciphertext (ops/mod.rs:229-258): every part of
// ciphertext b times the NTT-domain plaintext polynomial at pt + b pt_stride.
// grid = (N/256, parts L, num)
__global__ __launch_bounds__(256) void mul_plain_kernel(const u64* ct, u32 parts, const u64* __restrict__ pt,
                                                        size_t pt_stride, u64* out, const ModInfo* __restrict__ mods,
                                                        u32 L, u32 logn) {
    const u32 e = blockIdx.x * 256 + threadIdx.x;
    if (e >> logn) return;
    const u32 i = blockIdx.y % L, b = blockIdx.z;
    const ModInfo mi = mods[i];
    const size_t w = (((size_t)b * parts * L + blockIdx.y) << logn) + e;
    out[w] = mul_mod(ct[w], pt[(size_t)b * pt_stride + ((size_t)i << logn) + e], mi.barrett_lo, mi.barrett_hi, mi.p);
}

void launch_rows_sub_neg(const u64* a, const u64* b, u64* out, const ModInfo* mods, u32 L, int logn, size_t rows,
                         bool neg, hipStream_t s) {
    const size_t total = rows << logn;
    if (!total) return;
    ProfScope ps(neg ? "neg_rows" : "sub_rows", logn, s);
    hipLaunchKernelGGL(rows_sub_neg_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a, b, out, mods,
                       L, (u32)logn, total, neg ? 1u : 0u);
    HIPCHK(hipGetLastError());
}

void launch_mul_plain(const u64* ct, u32 parts, const u64* pt, size_t pt_stride, u64* out, const ModInfo* mods, u32 L,
                      int logn, u32 num, hipStream_t s) {
    if (!num || !parts) return;
    ProfScope ps("mul_plain", logn, s);
    hipLaunchKernelGGL(mul_plain_kernel, cgrid(logn, parts * L, num), dim3(256), 0, s, ct, parts, pt, pt_stride, out,
                       mods, L, (u32)logn);
    HIPCHK(hipGetLastError());
}

void launch_encode_gather(const u64* v, const u32* map, u64* rows, int logn, u32 num, hipStream_t s) {
    if (!num) return;
    ProfScope ps("encode_gather", logn, s);
    hipLaunchKernelGGL(encode_gather_kernel, cgrid(logn, 1, num), dim3(256), 0, s, v, map, rows, (u32)logn);
    HIPCHK(hipGetLastError());
}

void launch_pt_lift(const u64* m, u64* out, const CodecConst* cc, u64 t, u64 qmt, u64 qmt_shoup, u32 L, int logn,
                    u32 num, bool scaled, hipStream_t s) {
    if (!num) return;
    ProfScope ps("pt_lift", logn, s);
    hipLaunchKernelGGL(pt_lift_kernel, cgrid(logn, L, num), dim3(256), 0, s, m, out, cc, t, qmt, qmt_shoup, L,
                       (u32)logn, scaled ? 1u : 0u);
    HIPCHK(hipGetLastError());
}

void launch_small_lift(const int64_t* e, u64* out, const CodecConst* cc, u32 L, int logn, u32 num, hipStream_t s) {
    if (!num) return;
    ProfScope ps("small_lift", logn, s);
    hipLaunchKernelGGL(small_lift_kernel, cgrid(logn, L, num), dim3(256), 0, s, e, out, cc, L, (u32)logn);
    HIPCHK(hipGetLastError());
}

void launch_encrypt_sk(const u64* pt, const u64* a, const u64* e, const u64* sk, u64* ct, const CodecConst* cc,
                       u32 L, int logn, u32 num, hipStream_t s) {
    if (!num) return;
    ProfScope ps("encrypt_sk", logn, s);
    hipLaunchKernelGGL(encrypt_sk_kernel, cgrid(logn, L, num), dim3(256), 0, s, pt, a, e, sk, ct, cc, L, (u32)logn);
    HIPCHK(hipGetLastError());
}

void launch_phase(const u64* ct, u32 parts, const u64* sk, u64* out, const CodecConst* cc, u32 L, int logn, u32 num,
                  hipStream_t s) {
    if (!num) return;
    ProfScope ps("phase", logn, s);
    hipLaunchKernelGGL(phase_kernel, cgrid(logn, L, num), dim3(256), 0, s, ct, parts, sk, out, cc, L, (u32)logn);
    HIPCHK(hipGetLastError());
}

void launch_decode_reduce(const u64* d, u32 k, u64* w, const CodecConst& q0, const CodecConst& tc, int logn,
                          u32 num, hipStream_t s) {
    if (!num) return;
    ProfScope ps("decode_reduce", logn, s);
    hipLaunchKernelGGL(decode_reduce_kernel, cgrid(logn, 1, num), dim3(256), 0, s, d, k, w, q0, tc, (u32)logn);
    HIPCHK(hipGetLastError());
}

void launch_decode_gather(const u64* rows, const u32* map, u64* out, int logn, u32 num, hipStream_t s) {
    if (!num) return;
    ProfScope ps("decode_gather", logn, s);
    hipLaunchKernelGGL(decode_gather_kernel, cgrid(logn, 1, num), dim3(256), 0, s, rows, map, out, (u32)logn);
    HIPCHK(hipGetLastError());
}

}  // namespace fhehip
