// The Rq wire format (math/src/rq/convert.rs:15-87): each limb's
// power-basis coefficients packed least significant bit first, with the
// bit width of q_i - 1 (transcode_to_bytes / transcode_from_bytes,
// fhe-util/src/lib.rs:60-133), limbs back to back. Byte work, HBM-bound.
#include "bfv_launch.h"
#include "profiler.h"

namespace fhehip {

// One thread per output word W (uint8_t, or uint64_t when every limb's
// length N nbits / 8 is a multiple of 8, i.e. N >= 64): the coefficients
// overlapping its bits (a limb's bit length is a multiple of the word, so
// they lie in the limb). grid = (ceil(per_poly / sizeof(W) / 256), num)
template <class W>
__global__ __launch_bounds__(256) void pack_kernel(const u64* __restrict__ in, uint8_t* __restrict__ out, WireTable t,
                                                   u32 logn, u32 per_poly) {
    constexpr u32 BITS = 8 * sizeof(W);
    const u32 wi = blockIdx.x * 256 + threadIdx.x;
    const u32 b = wi * (u32)sizeof(W);
    if (b >= per_poly) return;
    const u32 poly = blockIdx.y;
    u32 i = 0;
    while (i + 1 < t.L && t.offset[i + 1] <= b) i++;
    const u32 nbits = t.nbits[i];
    const u64 mask = ~0ULL >> (64 - nbits);
    const u32 bit = (b - t.offset[i]) * 8;
    u32 k = bit / nbits, have = nbits - bit % nbits;
    const u64* row = in + (((size_t)poly * t.L + i) << logn);
    u64 acc = (row[k] & mask) >> (bit % nbits);
    while (have < BITS) {
        k++;
        acc |= (row[k] & mask) << have;
        have += nbits;
    }
    *(W*)(out + (size_t)poly * per_poly + b) = (W)acc;
}

// One thread per coefficient: its nbits bits from at most 9 bytes of the
// limb. grid = (N/256, L, num)
__global__ __launch_bounds__(256) void unpack_kernel(const uint8_t* __restrict__ in, u64* __restrict__ out,
                                                     WireTable t, u32 logn, u32 per_poly) {
    const u32 k = blockIdx.x * 256 + threadIdx.x;
    if (k >> logn) return;
    const u32 i = blockIdx.y, poly = blockIdx.z;
    const u32 nbits = t.nbits[i];
    const u32 bit = k * nbits, sh = bit % 8, nb = (sh + nbits + 7) / 8;
    const uint8_t* src = in + (size_t)poly * per_poly + t.offset[i] + bit / 8;
    u64 lo = 0;
    for (u32 j = 0; j < nb && j < 8; j++) lo |= (u64)src[j] << (8 * j);
    u64 v = lo >> sh;
    if (nb == 9) v |= (u64)src[8] << (64 - sh);
    out[(((size_t)poly * t.L + i) << logn) + k] = v & (~0ULL >> (64 - nbits));
}

void launch_pack(const u64* in, uint8_t* out, const WireTable& t, int logn, u32 per_poly, u32 num, hipStream_t s) {
    if (!num || !per_poly) return;
    ProfScope ps("pack", logn, s);
    if (logn >= 6 && ((uintptr_t)out & 7) == 0)
        hipLaunchKernelGGL(pack_kernel<uint64_t>, dim3((per_poly / 8 + 255) / 256, num), dim3(256), 0, s, in, out, t,
                           (u32)logn, per_poly);
    else
        hipLaunchKernelGGL(pack_kernel<uint8_t>, dim3((per_poly + 255) / 256, num), dim3(256), 0, s, in, out, t,
                           (u32)logn, per_poly);
    HIPCHK(hipGetLastError());
}

void launch_unpack(const uint8_t* in, u64* out, const WireTable& t, int logn, u32 per_poly, u32 num, hipStream_t s) {
    if (!num) return;
    ProfScope ps("unpack", logn, s);
    hipLaunchKernelGGL(unpack_kernel, dim3(((1u << logn) + 255) / 256, t.L, num), dim3(256), 0, s, in, out, t,
                       (u32)logn, per_poly);
    HIPCHK(hipGetLastError());
}

}  // namespace fhehip
