// Internal host-side declarations of the fhehip library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "device_math.h"
#include "host_math.h"

namespace fhehip {

// Error with a status code from include/fhehip.h.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);
#define HIPCHK(x) ::fhehip::hip_check((x), #x)

// Device allocation owned by a handle.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() {}
    explicit DevBuf(size_t n);
    ~DevBuf();
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void reset(size_t n);
    template <class T>
    T* as() const { return (T*)ptr; }
};

// Largest supported degree. Rows of up to 16384 coefficients are transformed
// in one workgroup; longer rows are split into halves, one elementwise stage
// pass per halving (DeviceContext::half, recursively).
constexpr u32 kMaxDegree = 131072;

// Resident per-context data (rq::Context, math/src/rq/context.rs:9-19).
struct DeviceContext {
    std::vector<u64> moduli;
    u32 L = 0, N = 0;
    int logn = 0;
    std::vector<u64> psi;
    std::vector<ModInfo> h_mod;
    DevBuf fwd_tw;   // [L][N] ulonglong2 (omega, omega_shoup)
    DevBuf inv_tw;   // [L][N] ulonglong2 (zeta_inv, zeta_inv_shoup)
    DevBuf mod;      // [L] ModInfo
    RnsBasis rns;    // host-side CRT constants
    // N >= 4096: a row can be transformed as two halves. The forward
    // transform's first stage and the inverse's last stage pair the two halves
    // of a row; each half is then an independent N/2-point transform. `half`
    // is a context of 2L rows of degree N/2 whose tables are the full tables
    // re-indexed for half h of limb l at row 2l + h (its inverse leaves out the
    // N^-1 factor). Degrees from 32768 always run this way (a row does not fit
    // one workgroup's LDS), and a half of 32768 or more coefficients has halves
    // of its own; degrees 4096 .. 16384 do for small launches (split_rows).
    std::shared_ptr<DeviceContext> half;
    DevBuf split;    // [L] ulonglong2 pairs: (omega_1, shoup), (zeta_last * N^-1, shoup), (N^-1, shoup)

    const ulonglong2* fwd(u32 limb) const { return fwd_tw.as<ulonglong2>() + (size_t)limb * N; }
    const ulonglong2* inv(u32 limb) const { return inv_tw.as<ulonglong2>() + (size_t)limb * N; }
};

std::shared_ptr<DeviceContext> make_context(const std::vector<u64>& moduli, u32 n, const u64* psi);

// Rows of a batch of polynomials laid out [poly][limbs_per_poly][N]. Row r of a
// launch covers limb first_limb + r % rows_per_poly of poly r / rows_per_poly.
struct RowMap {
    u64* data;
    u32 rows_per_poly;
    u32 limbs_per_poly;
    u32 first_limb;
    u32 table_offset;  // context limb index of limb first_limb
};

// Whether a launch of `rows` rows of context c runs the half-row kernels
// (each row as two halves of at most 16384 coefficients, the stage pairing
// the halves formed on load or left to the consumer): always at degree 32768;
// at 4096 .. 16384 when the launch has at most split_row_limit() rows (fewer
// workgroups than the chip has CUs, so that two workgroups per row shorten the
// launch). Never above 32768, where rows take the wide path (wide()).
bool split_rows(const DeviceContext& c, size_t rows);
// Degrees above 32768: the BFV operations run as elementwise kernels around
// the plain (recursively split) transforms.
inline bool wide(const DeviceContext& c) { return c.logn > 15; }
u32 split_row_limit();
// Sets the limit (0: only degree 32768 splits); process-wide.
void set_split_row_limit(u32 rows);

// NTT launchers (ntt_kernels.hip). `rows` = total rows in the launch.
void launch_ntt_forward(const DeviceContext& c, RowMap m, u32 rows, bool lazy, hipStream_t s);
void launch_ntt_backward(const DeviceContext& c, RowMap m, u32 rows, hipStream_t s);
// Picks the faster of 16 and 32 values per thread for this degree's plain
// NTT kernels by timing `rows` rows each way (ms: fwd16, fwd32, inv16, inv32).
void autotune_ntt(const DeviceContext& c, u32 rows, int& re_fwd, int& re_inv, float (&ms)[4]);
void set_ntt_shape(int logn, int re_fwd, int re_inv);
// Rows of one context transformed in the same launch as the digit lift:
// rows [rows][N] after their first forward stage (< 4p), row r of limb r % L,
// transformed in place to reduced NTT values with the half tables half_tw.
struct HalfFwdRows {
    u64* x;
    u32 rows, L;
    const ulonglong2* half_tw;
    const ModInfo* mods;
};
// Half-row transforms: y[b][i][j] = forward_vt_lazy(lazy_reduce_{q'_j}(src[b][i]))
// for every digit i < digits and key limb j of `key`, and the `rows`.
// log_base != 0: power-of-two decomposition digits of limb 0 (one-modulus keys).
void launch_lift_ntt_split(const DeviceContext& key, const u64* src, size_t src_inst, u64* y, u32 digits, u32 inst,
                           u32 log_base, hipStream_t s, const HalfFwdRows* rows = nullptr);
// The last inverse stage of `rows` rows whose two halves went through the
// half-row inverse transforms (N^-1 folded in).
void launch_ntt_split_inv(const DeviceContext& c, RowMap m, u32 rows, hipStream_t s);
// The forward transform of rows whose first stage (pairing the halves) the
// producer already formed: the two half transforms of every row only.
void launch_ntt_forward_halves(const DeviceContext& c, RowMap m, u32 rows, bool lazy, hipStream_t s);
// The inverse transform of rows as two half transforms each, without the
// stage pairing the halves (left to the consumer, as launch_ntt_split_inv).
void launch_ntt_backward_halves(const DeviceContext& c, RowMap m, u32 rows, hipStream_t s);

}  // namespace fhehip
