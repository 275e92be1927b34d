// Argument blocks and launchers of the BFV kernels (bfv_kernels.hip).
#pragma once

#include "fhehip_internal.h"

namespace fhehip {

struct InttRowsArgs {
    const u64* src[4];        // per part: base of instance 0, row 0
    size_t src_inst_stride;   // elements between instances (same for all parts)
    u64* dst;
    size_t dst_inst_stride;
    size_t dst_part_stride;
    const ulonglong2* inv_tw;  // context tables
    const ModInfo* mods;
    u32 table_offset;         // context limb of row 0
    const u32* perm;          // optional NTT-domain gather: q[m] = p[perm[m]]
};

struct TensorArgs {
    const u64* lhs;  // [inst][2][L][N] NTT
    const u64* rhs;
    size_t in_inst_stride, in_part_stride;
    const u64* ext;  // [inst][4][M - common][N] NTT (extension limbs of a0, a1, b0, b1)
    size_t ext_inst_stride, ext_part_stride;
    u32 common_lhs, common_rhs;
    u64* out;  // [inst][3][M][N] power basis
    size_t out_inst_stride, out_part_stride;
    const ulonglong2* inv_tw;  // multiplication context
    const ModInfo* mods;
    u32 M;                     // limbs of the multiplication context (set by the launcher)
    u32 xcd;                   // XCD-aware block order (set by the launcher)
};

struct KeySwitchArgs {
    const u64* src;  // [inst][digits][N] power basis (digit i = ciphertext limb i)
    size_t src_inst_stride;
    u32 digits, key_limbs;
    const ulonglong2* key0;  // [digits][key_limbs][N] (value, shoup)
    const ulonglong2* key1;
    const ulonglong2* key_fwd_tw;
    const ModInfo* key_mods;
    u32 xcd;       // XCD-aware block order (set by the launcher)
    u32 log_base;  // power-of-two decomposition: digit i is bits [i log_base, (i + 1) log_base) of limb 0; 0: RNS digits
};

// Constants of one switch-down step applied to one remaining limb.
struct SdConst {
    u64 q, bl, bh;    // modulus of the remaining limb and its Barrett constant
    u64 half_mod;     // q - (floor(q_last / 2) mod q)
    u64 inv, inv_shoup;  // q_last^-1 mod q
};

struct SdNttArgs {
    const u64* ks;  // [inst][parts][L + extra][N] power basis
    size_t ks_inst_stride, ks_part_stride;
    u32 L, extra;
    const u64* extra_mods;  // [extra] moduli of the dropped limbs
    const SdConst* sd;      // [extra][sd_stride]
    u32 sd_stride;
    const u64* add_pb;  // optional power-basis addend [inst][parts][L][N]
    size_t add_inst_stride, add_part_stride;
    const u64* add_ntt;  // optional NTT-domain addend for parts < add_ntt_parts
    size_t add_ntt_inst_stride, add_ntt_part_stride;
    u32 add_ntt_parts;
    const u32* perm;     // optional gather of the NTT addend: value m comes from perm[m]
    u64* out;  // [inst][parts][L][N] NTT, reduced
    size_t out_inst_stride, out_part_stride;
    const ulonglong2* fwd_tw;  // ciphertext context
    const ModInfo* mods;
};

// Key-switch finish with the inner product of the key limbs below the dropped
// one computed in place (extra <= 1): sd.ks is unused, `special` holds the
// dropped limb in power basis [inst][2][N] when sd.extra == 1.
struct KsFinishArgs {
    SdNttArgs sd;
    const u64* y;     // [inst][digits][key_limbs][N], values < 2q
    const u64* fin0;  // [digits][L + extra][N]: rows i < L times q_l^-1, row L plain
    const u64* fin1;
    u32 digits, key_limbs;
    u64* special;               // [inst][2][N] dropped limb (extra == 1)
    const ModInfo* special_mod; // modulus of the dropped limb
    u32 xcd;                    // XCD-aware block order (set by the launcher)
};

// Scaling launches of at most this many workgroups (4 per CU) copy their
// constants to LDS first; larger ones read them through the scalar cache.
constexpr size_t kScaleLdsMaxBlocks = 1024;
// d_sc: the constant block of sc_words words (rns_device.h).
void launch_rns_scale(const u64* in, size_t in_inst, size_t in_part, u64* out, size_t out_inst, size_t out_part,
                      const u64* d_sc, u32 sc_words, u32 nf, u32 n, u32 parts, u32 inst, hipStream_t s,
                      const char* name = "rns_scale");
// The same on the matrix cores (scale_mfma.hip) with the digit tables d_mt
// of layout `lay`, for inputs that are reduced residues; false (nothing
// launched) when the shape or the degree has no matrix-core form.
struct ScaleMfmaLayout;
bool launch_rns_scale_mfma(const u64* in, size_t in_inst, size_t in_part, u64* out, size_t out_inst,
                           size_t out_part, const u64* d_sc, const u64* d_mt, const ScaleMfmaLayout& lay, u32 n,
                           u32 parts, u32 inst, hipStream_t s, const char* name);
// Scaling of rows transformed as two halves (scale_pair.hip): rows of part k
// of instance b at in + b in_inst + k in_part (likewise out), degree n >= 512.
struct ScalePairArgs {
    const u64* in;
    size_t in_inst, in_part;
    u64* out;
    size_t out_inst, out_part;
    const u64* sc;  // constant block (rns_device.h) of sc_words words, `count` outputs
    u32 sc_words, count, n;
    // Optional: the inputs are the two half inverse transforms of each row;
    // the split table of the source context (3 entries per limb) and its
    // moduli give the last inverse stage.
    const ulonglong2* in_split;
    const ModInfo* in_mods;
    // Optional: the split table entries of the output limbs (3 per limb,
    // from the first output limb); outputs get the first forward stage.
    const ulonglong2* out_split;
};
void launch_scale_pair(const ScalePairArgs& a, u32 nf, u32 parts, u32 inst, hipStream_t s, const char* name);
// The same on the matrix cores (scale_mfma.hip), for inputs that are
// reduced residues (or open halves below 2p, with in_split); false (nothing
// launched) when the shape or the degree has no matrix-core form.
bool launch_rns_scale_mfma_pair(const ScalePairArgs& a, const u64* d_mt, const ScaleMfmaLayout& lay, u32 parts,
                                u32 inst, hipStream_t s, const char* name);
void launch_intt_rows(int logn, const InttRowsArgs& a, u32 limbs, u32 parts, u32 inst, hipStream_t s);
void launch_tensor_intt(int logn, const TensorArgs& a, u32 M, u32 inst, hipStream_t s);
// Degree 32768 (`c` the context): the same with the two half-row inverse
// transforms of every row; the last inverse stage (ntt_split_inv) follows
// separately, on the contiguous destination rows.
void launch_intt_rows_half(const DeviceContext& c, const InttRowsArgs& a, u32 limbs, u32 parts, u32 inst,
                           hipStream_t s);
void launch_tensor_intt_half(const DeviceContext& mul, const TensorArgs& a, u32 M, u32 inst, hipStream_t s);
// Degrees above 32768 (bfv_wide.hip): the elementwise halves of intt_rows,
// tensor_intt and lift_ntt, each followed by the plain transforms.
void launch_rows_gather(const InttRowsArgs& a, u32 limbs, u32 parts, u32 inst, int logn, hipStream_t s);
void launch_tensor_ew(const TensorArgs& a, u32 M, u32 inst, int logn, hipStream_t s);
void launch_lift_ew(const KeySwitchArgs& a, u64* y, u32 inst, int logn, hipStream_t s);
// Lifted, lazily transformed digits Y [inst][digits][key_limbs][N].
void launch_lift_ntt(int logn, const KeySwitchArgs& a, u64* y, u32 inst, hipStream_t s);
void launch_sd_add_ntt(int logn, const SdNttArgs& a, u32 L, u32 parts, u32 inst, hipStream_t s);
// out += NTT(power-basis switch-down part), degree <= 16384 (after ks_mac and
// the inverse NTT of `special`).
void launch_ks_finish(int logn, const KsFinishArgs& f, u32 inst, hipStream_t s);
// The same at degree 32768 (`ct` the ciphertext context): z on load, the
// half-row transforms, out += result.
// pre != null: the rows after their first stage, from launch_ks_z_pair.
void launch_ks_finish_half(const KsFinishArgs& f, const DeviceContext& ct, u32 inst, hipStream_t s,
                           const u64* pre = nullptr);
// The elementwise half of that finish over column pairs: the last inverse
// stage of the dropped limb (`special` after launch_ntt_backward_halves in
// the key context), z of every limb and the first forward stage, into pre
// [inst][2][L][N].
void launch_ks_z_pair(const KsFinishArgs& f, const DeviceContext& ct, const DeviceContext& key, u64* pre, u32 inst,
                      hipStream_t s);
void launch_add_rows(const u64* a, const u64* b, u64* out, const ModInfo* mods, u32 L, int logn, size_t rows,
                     hipStream_t s);

// NTT-domain products of ciphertexts of any part counts in the
// multiplication context: c[b][k] = sum over i + j = k of a[b][i] * b[b][j]
// (exact), a [num][lp][M][N], b [num][rp][M][N], c [num][lp + rp - 1][M][N].
void launch_tensor_parts(const u64* a, u32 lp, const u64* b, u32 rp, u64* c, const ModInfo* mods, u32 M, int logn,
                         u32 num, hipStream_t s);
// Elementwise kernels (bfv_generic.hip).
void launch_mac_ew(const u64* y, u32 digits, u32 key_limbs, u32 first, u32 count, const ulonglong2* key0,
                   const ulonglong2* key1, const ModInfo* key_mods, u64* acc, int logn, u32 inst, hipStream_t s);
// Key inner products, any degree (ks_mac_kernel): limbs below the dropped
// one (+ NTT-domain addend) into sd.out, the dropped limb (when sd.extra) into
// `special`, still in the NTT domain.
void launch_ks_mac(int logn, const KsFinishArgs& f, u32 inst, hipStream_t s);
// Poly::switch_down (math/src/rq/mod.rs:390-449) of power-basis polynomials
// of L limbs (in_stride limbs apart) into L - 1 limbs (out_stride apart),
// dropping limb L - 1 (modulus q_l); out may equal in.
void launch_switch_down_pb(const u64* in, u64* out, const SdConst* consts, u64 ql, u32 L, u32 in_stride,
                           u32 out_stride, int logn, u32 polys, hipStream_t s);
// Dot products of ciphertexts with NTT-domain plaintext polynomials
// (pir_kernels.hip). Row r of column c: ciphertext at ct + r * ct_stride
// ([parts][L][N], part stride `words` = L N), plaintext at pt + r *
// pt_row_stride + c * pt_col_stride; out[c] at out + c * out_col_stride.
struct DotArgs {
    const u64* ct;
    size_t ct_stride;
    const u64* pt;
    size_t pt_row_stride, pt_col_stride;
    u64* out;
    size_t out_col_stride;
    const ModInfo* mods;
    u32 count, columns, colgroups, words, logn;
    u32 parts, partgroups;  // parts of a row; groups of the kernel's PARTS parts (set by the launcher)
};
void launch_dot_product(DotArgs a, u32 parts, hipStream_t s);
// One level of the oblivious expansion (EvaluationKey::expands); `mono` is
// the level's monomial [L][N] as (value, Shoup) pairs.
void launch_expand_step(u64* out, const u64* sub, const ulonglong2* mono, const ModInfo* mods, u32 L, int logn,
                        u32 step, u32 size, hipStream_t s);
// Per-modulus constants of the encoding, encryption and decryption kernels
// (codec_kernels.hip): the modulus, floor(2^128 / q) and the scaling
// constant delta_i of Plaintext::to_poly with its Shoup companion.
struct CodecConst {
    u64 q, bl, bh, delta, delta_shoup;
};
// out = a - b, or -a with neg (a, b, out may alias), over rows of modulus r % L.
void launch_rows_sub_neg(const u64* a, const u64* b, u64* out, const ModInfo* mods, u32 L, int logn, size_t rows,
                         bool neg, hipStream_t s);
// out [num][parts][L][N] = ct times the plaintext row block at pt + b pt_stride (words).
void launch_mul_plain(const u64* ct, u32 parts, const u64* pt, size_t pt_stride, u64* out, const ModInfo* mods, u32 L,
                      int logn, u32 num, hipStream_t s);
void launch_encode_gather(const u64* v, const u32* map, u64* rows, int logn, u32 num, hipStream_t s);
void launch_pt_lift(const u64* m, u64* out, const CodecConst* cc, u64 t, u64 qmt, u64 qmt_shoup, u32 L, int logn,
                    u32 num, bool scaled, hipStream_t s);
void launch_small_lift(const int64_t* e, u64* out, const CodecConst* cc, u32 L, int logn, u32 num, hipStream_t s);
void launch_encrypt_sk(const u64* pt, const u64* a, const u64* e, const u64* sk, u64* ct, const CodecConst* cc,
                       u32 L, int logn, u32 num, hipStream_t s);
void launch_encrypt_pk(const u64* pt, const u64* pk, const u64* sm, size_t sm_stride, u64* ct, const CodecConst* cc,
                       u32 L, int logn, u32 num, hipStream_t s);
void launch_phase(const u64* ct, u32 parts, const u64* sk, u64* out, const CodecConst* cc, u32 L, int logn, u32 num,
                  hipStream_t s);
void launch_decode_reduce(const u64* d, u32 k, u64* w, const CodecConst& q0, const CodecConst& tc, int logn,
                          u32 num, hipStream_t s);
void launch_decode_gather(const u64* rows, const u32* map, u64* out, int logn, u32 num, hipStream_t s);
// Byte layout of one polynomial's Rq coefficients (wire_kernels.hip): limb i
// starts at byte offset[i] and packs N values of nbits[i] bits.
constexpr u32 kWireMaxLimbs = 64;
struct WireTable {
    u32 L;
    u32 nbits[kWireMaxLimbs];
    u32 offset[kWireMaxLimbs];
};
void launch_pack(const u64* in, uint8_t* out, const WireTable& t, int logn, u32 per_poly, u32 num, hipStream_t s);
void launch_unpack(const uint8_t* in, u64* out, const WireTable& t, int logn, u32 per_poly, u32 num, hipStream_t s);
// Ends of the finish after an in-place switch-down chain (any key depth,
// unfused path): post == false writes x = ks + add_pb (x contiguous
// [inst][2][L][N]); post == true writes out = x + add_ntt[perm].
void launch_sd_tail(int logn, const SdNttArgs& a, u64* x, bool post, u32 inst, hipStream_t s);
// Copies `count` slots of `elems` words; `table` is a device array of
// (src, dst) pointer pairs.
void launch_copy_slots(const u64* const* table, u32 count, size_t elems, hipStream_t s);
// x[e] = splitmix64(seed + e golden-ratio steps) >> shift (timing inputs).
void launch_fill_words(u64* x, size_t elems, u32 shift, u64 seed, hipStream_t s);

}  // namespace fhehip
