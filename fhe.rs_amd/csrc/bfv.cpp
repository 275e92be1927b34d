// Host orchestration of the BFV evaluation path: rq::Scaler, key-switching
// keys, Multiplicator, Galois keys. Every operation is a fixed sequence of
// batched kernel launches on the caller's stream; all data stays on the device.
#include "bfv.h"

#include <algorithm>
#include <cstring>

namespace fhehip {

// ---------------------------------------------------------------------------
// Scaler constants blob (layout documented in rns_device.h).
//
// SC_F_SUM128 is set when, for inputs below 2 q_j (reduced residues, or lazy
// ones below 2 q_j), every output's exact sum
//   v0 (q - gamma) + v1 [2^64 (q - gamma)]_q (+ w0 c0 + w1 c1) + sum_j r_j omega_ij
// is below 2^128. With r_j <= 2 q_j - 1 and theta_garner_j <= 2^shift,
// v <= sum_j r_j + 1; |theta_omega_j|, |theta_gamma| < 2^127 give
// w <= sum_j r_j + v + 1; c0, c1 < q.
namespace {
struct Sat128 {
    u128 v = 0;
    bool over = false;
    void add(u128 x) {
        const u128 s = v + x;
        if (s < v) over = true;
        v = s;
    }
};
}  // namespace

std::vector<u64> scaler_blob(const ScalerConsts& c, const std::vector<ModInfo>& from_mods,
                             const std::vector<ModInfo>& to_mods, u32 start) {
    const u32 nf = (u32)c.theta_garner_lo.size();
    const u32 nt = (u32)c.gamma.size();
    const u32 count = nt - start;
    if (nf > 32) throw Error(-2, "more than 32 source moduli in one scaling");
    std::vector<u64> b(sc_out_base(nf) + (size_t)count * sc_out_stride(nf), 0);
    b[SC_NF] = nf;
    b[SC_NT] = nt;
    b[SC_START] = start;
    b[SC_COUNT] = count;
    b[SC_SHIFT] = c.theta_garner_shift;
    b[SC_IS_ONE] = c.is_one ? 1 : 0;
    b[SC_TG_LO] = c.theta_gamma_lo;
    b[SC_TG_HI] = c.theta_gamma_hi;
    b[SC_TG_SIGN] = c.theta_gamma_sign ? 1 : 0;
    u64* tg = b.data() + SC_HEADER;
    u64* tom = tg + 2 * nf;
    u64 omega_mask = 0;
    for (u32 j = 0; j < nf; j++) {
        tg[2 * j] = c.theta_garner_lo[j];
        tg[2 * j + 1] = c.theta_garner_hi[j];
        tom[3 * j] = c.theta_omega_lo[j];
        tom[3 * j + 1] = c.theta_omega_hi[j];
        tom[3 * j + 2] = c.theta_omega_sign[j];
        if (c.theta_omega_lo[j] | c.theta_omega_hi[j]) omega_mask |= (u64)1 << j;
    }
    b[SC_OMEGA_MASK] = omega_mask;
    // bounds of the inputs and of v, w (see above)
    Sat128 rsum;
    std::vector<u64> rmax(nf);
    for (u32 j = 0; j < nf; j++) {
        rmax[j] = 2 * from_mods[j].p - 1;
        rsum.add(rmax[j]);
    }
    const u128 vmax = rsum.v + 1, wmax = rsum.v + vmax + 1;
    const u64 v1max = (u64)(vmax >> 64), w1max = (u64)(wmax >> 64);
    bool sum128 = !rsum.over && vmax > rsum.v && wmax > vmax;
    for (u32 i = 0; i < count; i++) {
        u64* o = b.data() + sc_out_base(nf) + (size_t)i * sc_out_stride(nf);
        const u32 ti = start + i;
        const u64 q = to_mods[ti].p;
        const u64 gneg = (q - c.gamma[ti] % q) % q;  // -gamma mod q
        o[0] = q;
        o[1] = to_mods[ti].barrett_lo;
        o[2] = to_mods[ti].barrett_hi;
        o[3] = gneg;
        o[4] = (u64)(((u128)gneg << 64) % q);
        o[5] = (u64)(((u128)1 << 64) % q);
        o[6] = (u64)(((u128)o[5] << 64) % q);
        o[7] = (u64)(((u128)o[5] << 64) / q);  // shoup([2^64]_q)
        o[8] = (u64)(((u128)1 << 64) / q);     // shoup(1)
        Sat128 bound;
        bound.add((u128)~0ull * o[3]);
        bound.add((u128)v1max * o[4]);
        if (!c.is_one) {
            bound.add((u128)~0ull * (q - 1));
            bound.add((u128)w1max * (q - 1));
        }
        for (u32 j = 0; j < nf; j++) {
            o[9 + j] = c.omega[(size_t)ti * nf + j];
            bound.add((u128)rmax[j] * o[9 + j]);
        }
        if (bound.over) sum128 = false;
    }
    b[SC_FLAGS] = (sum128 ? SC_F_SUM128 : 0u) | ((c.theta_gamma_lo | c.theta_gamma_hi) ? 0u : SC_F_TG_ZERO);
    return b;
}

// ---------------------------------------------------------------------------
// Batched launches index instances with the third grid dimension.
static void check_batch(u32 inst) {
    if (inst > 65535) throw Error(-1, "at most 65535 instances (or polynomials) per call");
}

// ---------------------------------------------------------------------------
// Rows of up to 16384 coefficients fit one workgroup's LDS, where they are
// transformed together with the elementwise work around the transform.
// Longer rows (N = 32768) are transformed as two halves, and so are the rows
// of small launches from N = 4096 (split_rows), to spread them over more CUs.
static bool fused(const DeviceContext& c) { return c.logn <= 14; }

// Whether launches of any size at this degree run `stage` (a HALVES_* bit of
// the tuning) as half rows (N = 4096 .. 16384).
static bool halves_any(const DeviceContext& c, int stage) {
    return c.logn >= 12 && c.logn <= 14 && c.half && (fused_tuning(c.logn).halves & stage);
}

// Inverse NTT of `parts` polynomials per instance into contiguous rows.
// Rows transformed as halves end with the stage pairing the halves, unless
// `close` is false: then that stage is left to the consumer (a paired
// scaling) and the function returns true.
static bool intt_rows(const DeviceContext& c, const InttRowsArgs& a, u32 L, u32 parts, u32 inst, hipStream_t s,
                      bool close = true) {
    if (wide(c)) {
        // copy (or gather) the rows, then transform them in place
        if (a.dst_inst_stride != (size_t)parts * L * c.N || (parts > 1 && a.dst_part_stride != (size_t)L * c.N))
            throw Error(-1, "the wide-row inverse NTT needs contiguous destination rows");
        launch_rows_gather(a, L, parts, inst, c.logn, s);
        launch_ntt_backward(c, RowMap{a.dst, L, L, 0, a.table_offset}, inst * parts * L, s);
        return false;
    }
    if (!split_rows(c, (size_t)L * parts * inst) && (close || !halves_any(c, HALVES_INPUT))) {
        launch_intt_rows(c.logn, a, L, parts, inst, s);
        return false;
    }
    const size_t rows = (size_t)L * c.N;
    if (a.dst_inst_stride != parts * rows || (parts > 1 && a.dst_part_stride != rows))
        throw Error(-1, "the half-row inverse NTT needs contiguous destination rows");
    launch_intt_rows_half(c, a, L, parts, inst, s);
    if (!close) return true;
    launch_ntt_split_inv(c, RowMap{a.dst, L, L, 0, a.table_offset}, inst * parts * L, s);
    return false;
}

// Tensor product followed by the inverse NTT of the three output polynomials
// (`close` as for intt_rows).
static bool tensor_intt(const DeviceContext& mul, const TensorArgs& a, u32 M, u32 inst, hipStream_t s,
                        bool close = true) {
    if (wide(mul)) {
        if (a.out_inst_stride != 3 * (size_t)M * mul.N || a.out_part_stride != (size_t)M * mul.N)
            throw Error(-1, "the wide-row tensor product needs contiguous output rows");
        launch_tensor_ew(a, M, inst, mul.logn, s);
        launch_ntt_backward(mul, RowMap{a.out, M, M, 0, 0}, inst * 3 * M, s);
        return false;
    }
    if (!split_rows(mul, (size_t)3 * M * inst) && (close || !halves_any(mul, HALVES_TENSOR))) {
        launch_tensor_intt(mul.logn, a, M, inst, s);
        return false;
    }
    if (a.out_inst_stride != 3 * (size_t)M * mul.N || a.out_part_stride != (size_t)M * mul.N)
        throw Error(-1, "the half-row tensor product needs contiguous output rows");
    launch_tensor_intt_half(mul, a, M, inst, s);
    if (!close) return true;
    launch_ntt_split_inv(mul, RowMap{a.out, M, M, 0, 0}, inst * 3 * M, s);
    return false;
}

// ---------------------------------------------------------------------------
PolyScaler::PolyScaler(std::shared_ptr<DeviceContext> f, std::shared_ptr<DeviceContext> t, const BigUint& num,
                       const BigUint& den)
    : from(std::move(f)), to(std::move(t)) {
    if (from->N != to->N) throw Error(-4, "Incompatible degrees");
    if (den.is_zero()) throw Error(-1, "zero denominator");
    is_one = (num == den);
    common = 0;
    if (is_one) {
        while (common < from->L && common < to->L && from->moduli[common] == to->moduli[common]) ++common;
    }
    consts.build(from->rns, to->rns, num, den);
    std::vector<u64> blob = scaler_blob(consts, from->h_mod, to->h_mod, common);
    d_consts.reset(blob.size() * sizeof(u64));
    HIPCHK(hipMemcpy(d_consts.ptr, blob.data(), d_consts.bytes, hipMemcpyHostToDevice));
    if (common < to->L) {
        const u64 pmax = *std::max_element(from->moduli.begin(), from->moduli.end());
        const std::vector<u64> t = scaler_mfma_blob(blob, pmax, mfma);
        if (!t.empty()) {
            d_mfma.reset(t.size() * sizeof(u64));
            HIPCHK(hipMemcpy(d_mfma.ptr, t.data(), d_mfma.bytes, hipMemcpyHostToDevice));
        }
    }
}

void PolyScaler::scale_pair_rows(const ScalePairArgs& a, u32 parts, u32 inst, hipStream_t s,
                                 const char* name) const {
    if (d_mfma.ptr && fused_tuning(from->logn).scaler_mfma &&
        launch_rns_scale_mfma_pair(a, d_mfma.as<u64>(), mfma, parts, inst, s, name))
        return;
    launch_scale_pair(a, from->L, parts, inst, s, name);
}

void PolyScaler::scale_rows(const u64* in, size_t in_inst, size_t in_part, u64* out, size_t out_inst,
                            size_t out_part, u32 parts, u32 inst, bool reduced, hipStream_t s, const char* name) const {
    const size_t n = from->N;
    if (reduced && d_mfma.ptr && fused_tuning(from->logn).scaler_mfma &&
        launch_rns_scale_mfma(in, in_inst, in_part, out, out_inst, out_part, d_consts.as<u64>(), d_mfma.as<u64>(),
                              mfma, (u32)n, parts, inst, s, name))
        return;
    launch_rns_scale(in, in_inst, in_part, out, out_inst, out_part, d_consts.as<u64>(), words(), from->L, (u32)n,
                     parts, inst, s, name);
}

ScalePairArgs PolyScaler::pair(const u64* in, size_t in_inst, size_t in_part, u64* out, size_t out_inst,
                               size_t out_part, bool in_open, bool out_fwd) const {
    ScalePairArgs a{};
    a.in = in;
    a.in_inst = in_inst;
    a.in_part = in_part;
    a.out = out;
    a.out_inst = out_inst;
    a.out_part = out_part;
    a.sc = d_consts.as<u64>();
    a.sc_words = words();
    a.count = to->L - common;
    a.n = (u32)from->N;
    if (in_open) {
        a.in_split = from->split.as<ulonglong2>();
        a.in_mods = from->mod.as<ModInfo>();
    }
    if (out_fwd) a.out_split = to->split.as<ulonglong2>() + 3 * (size_t)common;
    return a;
}

// rq::Scaler::scale (math/src/rq/scaler.rs:52-126) on `polys` polynomials.
void PolyScaler::scale(const u64* in, u64* out, u32 polys, bool ntt_repr, hipStream_t s) {
    check_batch(polys);
    const size_t n = from->N;
    const size_t isz = (size_t)from->L * n, osz = (size_t)to->L * n;
    if (!polys) return;
    if (common > 0)
        HIPCHK(hipMemcpy2DAsync(out, osz * 8, in, isz * 8, (size_t)common * n * 8, polys, hipMemcpyDeviceToDevice, s));
    if (common >= to->L) return;
    const u64* src = in;
    if (ntt_repr) {
        u64* tmp = scratch(Scratch::Scale).get(s, (size_t)polys * isz * 8).as<u64>();
        InttRowsArgs a{};
        a.src[0] = in;
        a.src_inst_stride = isz;
        a.dst = tmp;
        a.dst_inst_stride = isz;
        a.dst_part_stride = 0;
        a.inv_tw = from->inv_tw.as<ulonglong2>();
        a.mods = from->mod.as<ModInfo>();
        intt_rows(*from, a, from->L, 1, polys, s);
        src = tmp;
    }
    // the inverse transform's outputs are reduced; power-basis inputs are
    // taken as given
    scale_rows(src, isz, 0, out + (size_t)common * n, osz, 0, 1, polys, ntt_repr, s, "rns_scale");
    if (ntt_repr) {
        RowMap m{out, to->L - common, to->L, common, common};
        launch_ntt_forward(*to, m, polys * (to->L - common), false, s);
    }
}

// ---------------------------------------------------------------------------
Ksk::Ksk(std::shared_ptr<DeviceContext> ct, std::shared_ptr<DeviceContext> key, const u64* c0, const u64* c1,
         u32 ndigits)
    : ct_ctx(std::move(ct)), key_ctx(std::move(key)), digits(ndigits) {
    if (ct_ctx->N != key_ctx->N) throw Error(-4, "Incompatible degrees");
    if (key_ctx->L < ct_ctx->L) throw Error(-4, "key context below the ciphertext context");
    for (u32 i = 0; i < ct_ctx->L; i++)
        if (ct_ctx->moduli[i] != key_ctx->moduli[i])
            throw Error(-4, "the ciphertext context is not a level of the key context");
    if (key_ctx->L == 1) {
        // KeySwitchingKey::new with a one-modulus key context (key_switching_key
        // .rs:70-86): log_base = log2(next_power_of_two(q)) / 2 bit digits.
        u32 log_modulus = 0;
        while (log_modulus < 64 && (((u64)1 << log_modulus) < key_ctx->moduli[0])) log_modulus++;
        log_base = log_modulus / 2;
        const u32 need = (log_modulus + log_base - 1) / log_base;
        if (digits != need)
            throw Error(-1, "a one-modulus key has ceil(log2(q) / log_base) rows of the power-of-two decomposition");
    } else if (digits != ct_ctx->L) {
        throw Error(-1, "one key row per ciphertext modulus is required");
    }
    extra = key_ctx->L - ct_ctx->L;
    const u32 Lk = key_ctx->L, n = key_ctx->N;
    const size_t rows = (size_t)digits * Lk;
    std::vector<ulonglong2> h0(rows * n), h1(rows * n);
    for (size_t r = 0; r < rows; r++) {
        const u64 p = key_ctx->moduli[r % Lk];
        for (u32 j = 0; j < n; j++) {
            const size_t x = r * n + j;
            if (c0[x] >= p || c1[x] >= p) throw Error(-1, "key coefficient not reduced");
            h0[x] = make_ulonglong2(c0[x], shoup64(c0[x], p));
            h1[x] = make_ulonglong2(c1[x], shoup64(c1[x], p));
        }
    }
    key0.reset(h0.size() * sizeof(ulonglong2));
    key1.reset(h1.size() * sizeof(ulonglong2));
    HIPCHK(hipMemcpy(key0.ptr, h0.data(), key0.bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(key1.ptr, h1.data(), key1.bytes, hipMemcpyHostToDevice));
    // switch_down constants for dropping limbs L'-1 ... L (rq/mod.rs:390-449)
    const u32 L = ct_ctx->L;
    std::vector<u64> em;
    std::vector<SdConst> sd;
    sd_stride = Lk;
    sd.assign((size_t)std::max<u32>(extra, 1) * sd_stride, SdConst{});
    for (u32 s = 0; s < extra; s++) {
        const u64 ql = key_ctx->moduli[L + s];
        em.push_back(ql);
        for (u32 m = 0; m < L + s; m++) {
            const u64 q = key_ctx->moduli[m];
            SdConst c;
            c.q = q;
            c.bl = key_ctx->h_mod[m].barrett_lo;
            c.bh = key_ctx->h_mod[m].barrett_hi;
            c.half_mod = q - ((ql >> 1) % q);
            c.inv = invmod64(ql % q, q);
            c.inv_shoup = shoup64(c.inv, q);
            sd[(size_t)s * sd_stride + m] = c;
        }
    }
    if (extra) {
        extra_mods.reset(em.size() * 8);
        HIPCHK(hipMemcpy(extra_mods.ptr, em.data(), extra_mods.bytes, hipMemcpyHostToDevice));
    }
    sd_consts.reset(sd.size() * sizeof(SdConst));
    HIPCHK(hipMemcpy(sd_consts.ptr, sd.data(), sd_consts.bytes, hipMemcpyHostToDevice));
    if (extra <= 1) {
        // key_d,i * q_l^-1 mod q_i: the switch-down factor of limb i applied to
        // the key instead of to every key-switch output; the dropped limb's
        // rows (i == L) stay plain.
        std::vector<u64> f0((size_t)digits * Lk * n), f1(f0.size());
        for (u32 d = 0; d < digits; d++)
            for (u32 i = 0; i < Lk; i++) {
                const u64 q = key_ctx->moduli[i];
                const u64 inv = i < L && extra ? sd[i].inv : 1;
                const size_t src = ((size_t)d * Lk + i) * n, dst = src;
                for (u32 j = 0; j < n; j++) {
                    f0[dst + j] = mulmod64(c0[src + j], inv, q);
                    f1[dst + j] = mulmod64(c1[src + j], inv, q);
                }
            }
        fin0.reset(f0.size() * 8);
        fin1.reset(f1.size() * 8);
        HIPCHK(hipMemcpy(fin0.ptr, f0.data(), fin0.bytes, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(fin1.ptr, f1.data(), fin1.bytes, hipMemcpyHostToDevice));
    }
}

KeySwitchArgs Ksk::args(const u64* src, size_t src_inst_stride) const {
    KeySwitchArgs a{};
    a.src = src;
    a.src_inst_stride = src_inst_stride;
    a.digits = digits;
    a.key_limbs = key_ctx->L;
    a.key0 = key0.as<ulonglong2>();
    a.key1 = key1.as<ulonglong2>();
    a.key_fwd_tw = key_ctx->fwd_tw.as<ulonglong2>();
    a.key_mods = key_ctx->mod.as<ModInfo>();
    a.log_base = log_base;
    return a;
}

// Digit lift + lazy forward NTT of every (digit, key limb) pair
// (rq/mod.rs:518-540) into lift_work [inst][digits][L'][N].
size_t Ksk::lift_bytes(u32 inst) const { return (size_t)inst * digits * key_ctx->L * key_ctx->N * 8; }

// [inst][2][L'][N] power-basis key-switch rows (+ [inst][2][L][N] at degree
// 32768) for a deep key; [inst][2][N] for the dropped limb otherwise, +
// [inst][2][L][N] for the paired finish.
size_t Ksk::finish_bytes(u32 inst) const {
    const size_t n = ct_ctx->N, L = ct_ctx->L, Lk = key_ctx->L;
    if (extra > 1 || wide(*ct_ctx)) return (size_t)inst * 2 * (Lk + (fused(*ct_ctx) ? 0 : L)) * n * 8;
    return (size_t)inst * 2 * (extra + (paired_finish(inst) ? L : 0)) * n * 8;
}

// Rows transformed as halves (extra <= 1) finish in two steps when the
// launch is small or the finish bit of the halves knob is on: one elementwise pass over column
// pairs forms z and the first forward stage of every row, and the half
// transforms only load it. Other half-row launches (degree 32768) form z on
// load of each half (twice per coefficient, but no extra pass over the rows).
bool Ksk::paired_finish(u32 inst) const {
    const size_t L = ct_ctx->L, rows = (size_t)2 * L * inst;
    if (extra > 1) return false;
    if (halves_any(*ct_ctx, HALVES_FINISH) && (!extra || halves_any(*key_ctx, HALVES_FINISH))) return true;
    return rows <= split_row_limit() && split_rows(*ct_ctx, rows) && (!extra || split_rows(*key_ctx, (size_t)2 * inst));
}

u64* Ksk::lift(const u64* src, size_t src_inst_stride, u32 inst, hipStream_t s, const HalfFwdRows* rows) const {
    u64* y = scratch(Scratch::Lift).get(s, lift_bytes(inst)).as<u64>();
    const KeySwitchArgs a = args(src, src_inst_stride);
    if (wide(*key_ctx)) {
        // lifted digits, then their forward transforms (reduced: < 2q)
        if (rows && rows->rows) throw Error(-2, "paired finish rows at a wide degree");
        launch_lift_ew(a, y, inst, key_ctx->logn, s);
        launch_ntt_forward(*key_ctx, RowMap{y, key_ctx->L, key_ctx->L, 0, 0}, inst * digits * key_ctx->L, false, s);
        return y;
    }
    if (!split_rows(*key_ctx, (size_t)digits * key_ctx->L * inst)) {
        launch_lift_ntt(key_ctx->logn, a, y, inst, s);
        if (rows && rows->rows)
            launch_ntt_forward_halves(*ct_ctx, RowMap{rows->x, rows->L, rows->L, 0, 0}, rows->rows, false, s);
    } else {
        launch_lift_ntt_split(*key_ctx, src, src_inst_stride, y, digits, inst, log_base, s, rows);
    }
    return y;
}

void Ksk::key_switch_pb(const u64* src, size_t src_inst_stride, u64* acc, u32 inst, bool to_power_basis,
                        hipStream_t s) const {
    check_batch(inst);
    const u32 Lk = key_ctx->L;
    const u64* y = lift(src, src_inst_stride, inst, s);
    launch_mac_ew(y, digits, Lk, 0, Lk, key0.as<ulonglong2>(), key1.as<ulonglong2>(), key_ctx->mod.as<ModInfo>(), acc,
                  key_ctx->logn, inst, s);
    if (to_power_basis) launch_ntt_backward(*key_ctx, RowMap{acc, Lk, Lk, 0, 0}, inst * 2 * Lk, s);
}

SdNttArgs Ksk::sd_args(u64* out, const u64* add_pb, size_t add_inst, size_t add_part, const u64* add_ntt,
                       size_t add_ntt_inst, size_t add_ntt_part, u32 add_ntt_parts, const u32* perm) const {
    const size_t n = ct_ctx->N;
    SdNttArgs a{};
    a.ks_part_stride = (size_t)key_ctx->L * n;
    a.ks_inst_stride = 2 * a.ks_part_stride;
    a.L = ct_ctx->L;
    a.extra = extra;
    a.extra_mods = extra ? extra_mods.as<u64>() : nullptr;
    a.sd = sd_consts.as<SdConst>();
    a.sd_stride = sd_stride;
    a.add_pb = add_pb;
    a.add_inst_stride = add_inst;
    a.add_part_stride = add_part;
    a.add_ntt = add_ntt;
    a.add_ntt_inst_stride = add_ntt_inst;
    a.add_ntt_part_stride = add_ntt_part;
    a.add_ntt_parts = add_ntt_parts;
    a.perm = perm;
    a.out = out;
    a.out_part_stride = (size_t)ct_ctx->L * n;
    a.out_inst_stride = 2 * a.out_part_stride;
    a.fwd_tw = ct_ctx->fwd_tw.as<ulonglong2>();
    a.mods = ct_ctx->mod.as<ModInfo>();
    return a;
}

void Ksk::switch_and_finish(const u64* src, size_t src_inst_stride, u64* out, u32 inst, const u64* add_pb,
                            size_t add_inst, size_t add_part, const u64* add_ntt, size_t add_ntt_inst,
                            size_t add_ntt_part, u32 add_ntt_parts, const u32* perm, hipStream_t s) const {
    check_batch(inst);
    const size_t n = ct_ctx->N, Lk = key_ctx->L, L = ct_ctx->L;
    SdNttArgs a = sd_args(out, add_pb, add_inst, add_part, add_ntt, add_ntt_inst, add_ntt_part, add_ntt_parts, perm);
    if (extra > 1 || wide(*ct_ctx)) {
        // Several dropped moduli (switch_down_to, math/src/rq/mod.rs:455-462),
        // or rows too long for the fused finish (degrees above 32768):
        // key switch to power basis, then Poly::switch_down once per dropped
        // modulus, last first. The drops of limbs L' - 1 ... L + 1 run in
        // place on the key-switch rows; the last one (limb L) is fused with
        // the addends and the forward NTT (degree <= 16384) or runs like the
        // others before two elementwise ends around the split NTT.
        const bool fuse = fused(*ct_ctx);
        u64* acc = scratch(Scratch::SwitchDown).get(s, finish_bytes(inst)).as<u64>();
        key_switch_pb(src, src_inst_stride, acc, inst, true, s);
        const u32 keep = fuse ? (u32)L + 1 : (u32)L;  // limbs left by the in-place drops
        for (u32 cur = (u32)Lk; cur > keep; cur--)
            launch_switch_down_pb(acc, acc, sd_consts.as<SdConst>() + (size_t)(cur - 1 - L) * sd_stride,
                                  key_ctx->moduli[cur - 1], cur, (u32)Lk, (u32)Lk, ct_ctx->logn, inst * 2, s);
        a.ks = acc;
        if (fuse) {
            a.extra = 1;
            launch_sd_add_ntt(ct_ctx->logn, a, ct_ctx->L, 2, inst, s);
            return;
        }
        u64* x = acc + (size_t)inst * 2 * Lk * n;
        launch_sd_tail(ct_ctx->logn, a, x, false, inst, s);
        launch_ntt_forward(*ct_ctx, RowMap{x, (u32)L, (u32)L, 0, 0}, inst * 2 * (u32)L, false, s);
        launch_sd_tail(ct_ctx->logn, a, x, true, inst, s);
        return;
    }
    const bool paired = paired_finish(inst);
    if (paired && !extra && a.add_pb && !a.add_ntt) {
        // No dropped modulus: out = NTT(add_pb) + inner products. The
        // power-basis addend gets its first forward stage in one pass, the
        // rest of its transform in the lift's launch, and ks_mac adds it as
        // an NTT-domain addend (no finish kernels).
        u64* pre = scratch(Scratch::SwitchDown).get(s, finish_bytes(inst)).as<u64>();
        KsFinishArgs z{a, nullptr, nullptr, nullptr, digits, (u32)Lk, nullptr, key_ctx->mod.as<ModInfo>()};
        launch_ks_z_pair(z, *ct_ctx, *key_ctx, pre, inst, s);
        const HalfFwdRows rows{pre, (u32)(inst * 2 * L), (u32)L, ct_ctx->half->fwd_tw.as<ulonglong2>(),
                               ct_ctx->mod.as<ModInfo>()};
        const u64* y = lift(src, src_inst_stride, inst, s, &rows);
        SdNttArgs m = a;
        m.add_pb = nullptr;
        m.add_ntt = pre;
        m.add_ntt_inst_stride = 2 * L * n;
        m.add_ntt_part_stride = L * n;
        m.add_ntt_parts = 2;
        m.perm = nullptr;
        launch_ks_mac(ct_ctx->logn, KsFinishArgs{m, y, fin0.as<u64>(), fin1.as<u64>(), digits, (u32)Lk, nullptr,
                                                 key_ctx->mod.as<ModInfo>() + (Lk - 1)},
                      inst, s);
        return;
    }
    const u64* y = lift(src, src_inst_stride, inst, s);
    // Inner products of every key limb: the limbs below the dropped one into
    // `out` (NTT domain), the dropped limb into `special`, which alone goes
    // back to power basis; then the transformed power-basis part is added in.
    u64* base = scratch(Scratch::SwitchDown).get(s, finish_bytes(inst)).as<u64>();
    u64* special = extra ? base : nullptr;
    KsFinishArgs f{a, y, fin0.as<u64>(), fin1.as<u64>(), digits, (u32)Lk, special,
                   key_ctx->mod.as<ModInfo>() + (Lk - 1)};
    launch_ks_mac(ct_ctx->logn, f, inst, s);
    if (!extra && !a.add_pb) return;  // the finish would add NTT(0)
    if (paired) {
        u64* pre = base + (size_t)inst * 2 * extra * n;
        if (extra) launch_ntt_backward_halves(*key_ctx, RowMap{special, 1, 1, 0, (u32)(Lk - 1)}, inst * 2, s);
        launch_ks_z_pair(f, *ct_ctx, *key_ctx, pre, inst, s);
        launch_ks_finish_half(f, *ct_ctx, inst, s, pre);
        return;
    }
    if (extra) launch_ntt_backward(*key_ctx, RowMap{special, 1, 1, 0, (u32)(Lk - 1)}, inst * 2, s);
    if (split_rows(*ct_ctx, (size_t)2 * L * inst))
        launch_ks_finish_half(f, *ct_ctx, inst, s);
    else
        launch_ks_finish(ct_ctx->logn, f, inst, s);
}

// KeySwitchingKey::key_switch (key_switching_key.rs:214-241): power-basis p in
// the ciphertext context -> NTT-domain (c0, c1) in the key context.
void Ksk::key_switch(const u64* p, u64* out, u32 inst, hipStream_t s) const {
    key_switch_pb(p, (size_t)ct_ctx->L * ct_ctx->N, out, inst, false, s);
}

// RelinearizationKey::relinearizes (relinearization_key.rs:71-98) on
// [inst][3][L][N] NTT ciphertexts -> [inst][2][L][N].
void Ksk::relinearize(const u64* ct3, u64* out, u32 inst, hipStream_t s) {
    check_batch(inst);
    const size_t n = ct_ctx->N, L = ct_ctx->L;
    u64* c2 = scratch(Scratch::RelinC2).get(s, (size_t)inst * L * n * 8).as<u64>();
    InttRowsArgs a{};
    a.src[0] = ct3 + 2 * L * n;
    a.src_inst_stride = 3 * L * n;
    a.dst = c2;
    a.dst_inst_stride = L * n;
    a.inv_tw = ct_ctx->inv_tw.as<ulonglong2>();
    a.mods = ct_ctx->mod.as<ModInfo>();
    intt_rows(*ct_ctx, a, (u32)L, 1, inst, s);
    // (c0, c1) + switch_down(key switch), added in the NTT domain after the transform.
    switch_and_finish(c2, L * n, out, inst, nullptr, 0, 0, ct3, 3 * L * n, L * n, 2, nullptr, s);
}

// ---------------------------------------------------------------------------
static std::vector<u64> default_extended_basis(const DeviceContext& base) {
    // Multiplicator::default (fhe/src/bfv/ops/mul.rs:101-117)
    size_t bits = 0;
    for (u64 q : base.moduli) bits += 64 - __builtin_clzll(q);
    const size_t n_moduli = (bits + 60 + 61) / 62;
    std::vector<u64> ext = base.moduli;
    u64 ub = (u64)1 << 62;
    while (ext.size() != base.moduli.size() + n_moduli) {
        ub = generate_prime64(62, 2 * (u64)base.N, ub);
        if (!ub) throw Error(-2, "not enough 62-bit primes for the extended basis");
        if (std::find(ext.begin(), ext.end(), ub) == ext.end()) ext.push_back(ub);
    }
    return ext;
}

Multiplicator::Multiplicator(std::shared_ptr<DeviceContext> b, const std::vector<u64>& extended_basis,
                             const BigUint& lhs_num, const BigUint& lhs_den, const BigUint& rhs_num,
                             const BigUint& rhs_den, const BigUint& post_num, const BigUint& post_den)
    : base(std::move(b)) {
    mul = make_context(extended_basis, base->N, nullptr);
    ext_lhs.reset(new PolyScaler(base, mul, lhs_num, lhs_den));
    ext_rhs.reset(new PolyScaler(base, mul, rhs_num, rhs_den));
    ext_same = lhs_num == rhs_num && lhs_den == rhs_den;
    down.reset(new PolyScaler(mul, base, post_num, post_den));
}

std::unique_ptr<Multiplicator> Multiplicator::make_default(std::shared_ptr<Ksk> rk, u64 plaintext) {
    auto base = rk->ct_ctx;
    std::vector<u64> ext = default_extended_basis(*base);
    std::unique_ptr<Multiplicator> m(new Multiplicator(base, ext, BigUint(1), BigUint(1), BigUint(1), BigUint(1),
                                                       BigUint(plaintext), base->rns.product));
    m->rk = std::move(rk);
    return m;
}

// ---------------------------------------------------------------------------
Switcher::Switcher(std::shared_ptr<DeviceContext> f, std::shared_ptr<DeviceContext> t)
    : from(std::move(f)), to(std::move(t)) {
    if (from->N != to->N) throw Error(-4, "Incompatible degrees");
    if (to->L > from->L || !std::equal(to->moduli.begin(), to->moduli.end(), from->moduli.begin()))
        throw Error(-4, "the target context must be the source context without its last moduli");
    // switch_down constants of each step (rq/mod.rs:390-449): step s drops
    // q_l = modulus L_from - 1 - s from the limbs below it
    const u32 L = from->L, k = L - to->L;
    if (!k) return;
    std::vector<SdConst> sd((size_t)k * (L - 1));
    for (u32 st = 0; st < k; st++) {
        const u64 ql = from->moduli[L - 1 - st];
        for (u32 m = 0; m + 1 + st < L; m++) {
            const u64 q = from->moduli[m];
            SdConst& c = sd[(size_t)st * (L - 1) + m];
            c.q = q;
            c.bl = from->h_mod[m].barrett_lo;
            c.bh = from->h_mod[m].barrett_hi;
            c.half_mod = q - ((ql >> 1) % q);
            c.inv = invmod64(ql % q, q);
            c.inv_shoup = shoup64(c.inv, q);
        }
    }
    consts.reset(sd.size() * sizeof(SdConst));
    HIPCHK(hipMemcpy(consts.ptr, sd.data(), consts.bytes, hipMemcpyHostToDevice));
}

// Ciphertext::switch_down, once per dropped modulus (ciphertext.rs:86-119):
// every limb to power basis, the steps, the remaining limbs back to the NTT
// domain. Steps before the last run in place on the power-basis rows.
void Switcher::switch_down(const u64* in, u64* out, u32 polys, hipStream_t s) const {
    check_batch(polys);
    if (!polys) return;
    const size_t n = from->N, L = from->L, k = L - to->L;
    if (!k) {
        HIPCHK(hipMemcpyAsync(out, in, (size_t)polys * L * n * 8, hipMemcpyDeviceToDevice, s));
        return;
    }
    u64* pb = scratch(Scratch::ModSwitch).get(s, scratch_bytes(polys)).as<u64>();
    InttRowsArgs ia{};
    ia.src[0] = in;
    ia.src_inst_stride = L * n;
    ia.dst = pb;
    ia.dst_inst_stride = L * n;
    ia.dst_part_stride = L * n;
    ia.inv_tw = from->inv_tw.as<ulonglong2>();
    ia.mods = from->mod.as<ModInfo>();
    intt_rows(*from, ia, (u32)L, 1, polys, s);
    for (size_t st = 0; st < k; st++) {
        const u32 cur = (u32)(L - st);
        const bool last = st + 1 == k;
        launch_switch_down_pb(pb, last ? out : pb, consts.as<SdConst>() + st * (L - 1), from->moduli[cur - 1], cur,
                              (u32)L, last ? (u32)to->L : (u32)L, from->logn, polys, s);
    }
    launch_ntt_forward(*to, RowMap{out, to->L, to->L, 0, 0}, polys * to->L, false, s);
}

size_t Switcher::scratch_bytes(u32 polys) const {
    return from->L == to->L ? 0 : (size_t)polys * from->L * from->N * 8;
}

// [inst][4][L][N] inverse-transformed inputs, [inst][4][xr][N] extension
// rows, [inst][3][M][N] tensor rows and [inst][3][L][N] down-scaled rows.
size_t Multiplicator::mul_bytes(u32 inst) const {
    const size_t n = base->N, L = base->L, M = mul->L;
    const size_t xr = M - std::min(ext_lhs->common, ext_rhs->common);
    return (size_t)inst * (4 * L + 4 * xr + 3 * M + 3 * L) * n * 8;
}

size_t Multiplicator::scratch_bytes(u32 inst) const {
    size_t b = FixedArena::aligned(mul_bytes(inst));
    if (rk) b += FixedArena::aligned(rk->lift_bytes(inst)) + FixedArena::aligned(rk->finish_bytes(inst));
    if (mod_switch) b += FixedArena::aligned(mod_switch->scratch_bytes(inst * (rk ? 2 : 3)));
    return b;
}

// Multiplicator::multiply (fhe/src/bfv/ops/mul.rs:158-227) on a batch.
void Multiplicator::multiply(const u64* lhs, const u64* rhs, u64* out_final, u32 inst, hipStream_t s) {
    check_batch(inst);
    // the final switch_down runs on every part of every product
    if (mod_switch) check_batch(inst * (rk ? 2 : 3));
    const size_t n = base->N, L = base->L, M = mul->L;
    const u32 cl = ext_lhs->common, cr = ext_rhs->common;
    const size_t xr = M - std::min(cl, cr);  // extension rows per input polynomial
    const size_t ct_inst = 2 * L * n;
    const size_t P = (size_t)inst * 4 * L * n;
    const size_t X = (size_t)inst * 4 * xr * n;
    const size_t T = (size_t)inst * 3 * M * n;
    u64* wP = scratch(Scratch::Mul).get(s, mul_bytes(inst)).as<u64>();
    u64* wX = wP + P;
    u64* wT = wX + X;
    u64* wD = wT + T;
    // With modulus switching the product lands in wP (free after the tensor
    // product; [inst][3][L][N] fits in its [inst][4][L][N]) and is switched
    // down into the caller's buffer at the end (mul.rs:221-223).
    u64* out = mod_switch ? wP : out_final;
    // 1. inverse NTT of the four input polynomials (rq/scaler.rs:68-78)
    InttRowsArgs ia{};
    ia.src[0] = lhs;
    ia.src[1] = lhs + L * n;
    ia.src[2] = rhs;
    ia.src[3] = rhs + L * n;
    ia.src_inst_stride = ct_inst;
    ia.dst = wP;
    ia.dst_inst_stride = 4 * L * n;
    ia.dst_part_stride = L * n;
    ia.inv_tw = base->inv_tw.as<ulonglong2>();
    ia.mods = base->mod.as<ModInfo>();
    // Rows transformed as halves leave the stage pairing the halves to the
    // scalers, which also form the first stage of the forward transform of
    // the new limbs when those are transformed as halves.
    const bool in_open = intt_rows(*base, ia, (u32)L, 4, inst, s, false);
    const bool out_fwd =
        cl == cr && cl < M && (split_rows(*mul, (size_t)inst * 4 * (M - cl)) || halves_any(*mul, HALVES_INPUT));
    // 2. base extension of the power-basis columns (rns/scaler.rs:249-352)
    auto extend = [&](const PolyScaler& sc, u32 k0, u32 parts) {
        if (sc.common >= M) return;
        const u64* in = wP + k0 * L * n;
        u64* o = wX + k0 * xr * n;
        if (in_open || out_fwd)
            sc.scale_pair_rows(sc.pair(in, 4 * L * n, L * n, o, 4 * xr * n, xr * n, in_open, out_fwd), parts, inst, s,
                               "rns_extend");
        else  // inverse transform outputs: reduced
            sc.scale_rows(in, 4 * L * n, L * n, o, 4 * xr * n, xr * n, parts, inst, true, s, "rns_extend");
    };
    if (ext_same) {
        extend(*ext_lhs, 0, 4);
    } else {
        extend(*ext_lhs, 0, 2);
        extend(*ext_rhs, 2, 2);
    }
    // 3. forward NTT of the new limbs (rq/scaler.rs:96-114)
    if (cl == cr) {
        RowMap m{wX, (u32)(M - cl), (u32)xr, 0, cl};
        if (out_fwd) launch_ntt_forward_halves(*mul, m, inst * 4 * (u32)(M - cl), false, s);
        else launch_ntt_forward(*mul, m, inst * 4 * (u32)(M - cl), false, s);
    } else {
        for (int k = 0; k < 4; k++) {
            const u32 c = k < 2 ? cl : cr;
            RowMap m{wX + k * xr * n, (u32)(M - c), (u32)(4 * xr), 0, c};
            launch_ntt_forward(*mul, m, inst * (u32)(M - c), false, s);
        }
    }
    // 4. tensor product + inverse NTT (mul.rs:181-185, rq/scaler.rs:68-78)
    TensorArgs ta{};
    ta.lhs = lhs;
    ta.rhs = rhs;
    ta.in_inst_stride = ct_inst;
    ta.in_part_stride = L * n;
    ta.ext = wX;
    ta.ext_inst_stride = 4 * xr * n;
    ta.ext_part_stride = xr * n;
    ta.common_lhs = cl;
    ta.common_rhs = cr;
    ta.out = wT;
    ta.out_inst_stride = 3 * M * n;
    ta.out_part_stride = M * n;
    ta.inv_tw = mul->inv_tw.as<ulonglong2>();
    ta.mods = mul->mod.as<ModInfo>();
    const bool t_open = tensor_intt(*mul, ta, (u32)M, inst, s, false);
    // 5. t/Q down-scaling of the three polynomials (mul.rs:187-190)
    u64* dst = rk ? wD : out;
    const bool d_fwd = !rk && split_rows(*base, (size_t)inst * 3 * L);
    if (t_open || d_fwd)
        down->scale_pair_rows(down->pair(wT, 3 * M * n, M * n, dst, 3 * L * n, L * n, t_open, d_fwd), 3, inst, s,
                              "rns_down");
    else  // inverse transform outputs: reduced
        down->scale_rows(wT, 3 * M * n, M * n, dst, 3 * L * n, L * n, 3, inst, true, s, "rns_down");
    if (!rk) {
        RowMap m{out, (u32)L, (u32)L, 0, 0};
        if (d_fwd) launch_ntt_forward_halves(*base, m, inst * 3 * (u32)L, false, s);
        else launch_ntt_forward(*base, m, inst * 3 * (u32)L, false, s);
        if (mod_switch) mod_switch->switch_down(out, out_final, inst * 3, s);
        return;
    }
    // 6. key switch of c2 (power basis already) (relinearization_key.rs:101-106), then
    // 7. switch down to the ciphertext context, add c0/c1, forward NTT (mul.rs:194-210)
    rk->switch_and_finish(wD + 2 * L * n, 3 * L * n, out, inst, wD, 3 * L * n, L * n, nullptr, 0, 0, 0, nullptr,
                          s);
    // 8. Ciphertext::switch_down when modulus switching is enabled (mul.rs:221-223)
    if (mod_switch) mod_switch->switch_down(out, out_final, inst * 2, s);
}

// ---------------------------------------------------------------------------
FusedTuning autotune_multiply(Multiplicator& m, u32 inst, float (&ms)[19]) {
    check_batch(inst);
    if (!inst) throw Error(-1, "the autotuner needs at least one instance");
    if (wide(*m.base)) throw Error(-2, "the autotuner covers degrees 8 .. 32768 (the fused row kernels)");
    const int logn = m.base->logn;
    const size_t n = m.base->N, L = m.base->L;
    const size_t words_in = (size_t)inst * 2 * L * n;
    const size_t words_out = (size_t)inst * (m.rk ? 2 : 3) * (m.mod_switch ? L - 1 : L) * n;
    DevBuf in(words_in * 8), out(words_out * 8);
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const FusedTuning start = fused_tuning(logn);
    FusedTuning best = start;
    try {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        // Pseudorandom residues (below the smallest modulus), so that the
        // timed work draws the power of real operands: all-zero operands let
        // the chip clock higher than uniform residues do.
        u64 qmin = ~0ull;
        for (u64 q : m.base->moduli) qmin = std::min(qmin, q);
        launch_fill_words(in.as<u64>(), words_in, (u32)__builtin_clzll(qmin) + 1, 0x5eed, s);
        auto time_it = [&](const FusedTuning& t) {
            set_fused_tuning(logn, t);
            const int reps = 3;
            m.multiply(in.as<u64>(), in.as<u64>(), out.as<u64>(), inst, s);
            m.multiply(in.as<u64>(), in.as<u64>(), out.as<u64>(), inst, s);
            HIPCHK(hipEventRecord(e0, s));
            for (int r = 0; r < reps; r++) m.multiply(in.as<u64>(), in.as<u64>(), out.as<u64>(), inst, s);
            HIPCHK(hipEventRecord(e1, s));
            HIPCHK(hipEventSynchronize(e1));
            float el = 0;
            HIPCHK(hipEventElapsedTime(&el, e0, e1));
            return el / reps;
        };
        int slot = 0;
        auto knob = [&](int FusedTuning::*field, std::initializer_list<int> values) {
            int pick = best.*field;
            float fastest = 0;
            for (int v : values) {
                FusedTuning t = best;
                t.*field = v;
                const float el = time_it(t);
                ms[slot++] = el;
                if (fastest == 0 || el < fastest) {
                    fastest = el;
                    pick = v;
                }
            }
            best.*field = pick;
        };
        knob(&FusedTuning::xcd, {0, 1});
        knob(&FusedTuning::ks_mac_inst, {2, 4, 8});
        knob(&FusedTuning::re_tensor, {4, 5});
        knob(&FusedTuning::re_lift, {4, 5});
        knob(&FusedTuning::re_finish, {4, 5});
        knob(&FusedTuning::scaler_mfma, {0, 1});
        // each stage's half-row choice in turn
        for (int bit : {HALVES_FINISH, HALVES_INPUT, HALVES_TENSOR}) {
            float t[2];
            for (int on = 0; on < 2; on++) {
                FusedTuning c = best;
                c.halves = on ? (best.halves | bit) : (best.halves & ~bit);
                t[on] = time_it(c);
                ms[slot++] = t[on];
            }
            best.halves = t[1] < t[0] ? (best.halves | bit) : (best.halves & ~bit);
        }
        set_fused_tuning(logn, best);
    } catch (...) {
        set_fused_tuning(logn, start);
        (void)hipStreamSynchronize(s);
        release_scratch(s);
        (void)hipStreamDestroy(s);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        throw;
    }
    HIPCHK(hipStreamSynchronize(s));
    release_scratch(s);
    (void)hipStreamDestroy(s);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

// ---------------------------------------------------------------------------
// BfvParametersBuilder::build (fhe/src/bfv/parameters.rs:431-641).
Parameters::Parameters(u32 degree, u64 plaintext, const std::vector<u64>& m) : N(degree), t(plaintext), moduli(m) {
    if (degree < 8 || (degree & (degree - 1))) throw Error(-1, "the degree must be a power of two, at least 8");
    if (moduli.empty()) throw Error(-1, "no ciphertext moduli");
    if (t < 2 || (t >> 62)) throw Error(-2, "plaintext modulus out of range (2 .. 2^62)");
    for (u64 q : moduli)
        if (t % q == 0) throw Error(-1, "the plaintext modulus is not invertible modulo a ciphertext modulus");
    // plaintext context: the first moduli whose sizes reach log2(t) + 60 bits (parameters.rs:479-494)
    const u32 t_bits = 64 - __builtin_clzll(t);
    u32 count = 0, acc = 0;
    for (u64 q : moduli) {
        acc += 64 - __builtin_clzll(q);
        count++;
        if (acc >= t_bits + 60) break;
    }
    plain = make_context(std::vector<u64>(moduli.begin(), moduli.begin() + count), N, nullptr);
    // L + 1 62-bit primes = 1 mod 2N, none of the moduli (parameters.rs:585-593)
    u64 ub = (u64)1 << 62;
    while (extension.size() != moduli.size() + 1) {
        ub = generate_prime64(62, 2 * (u64)N, ub);
        if (!ub) throw Error(-2, "not enough 62-bit primes for the multiplication basis");
        if (std::find(extension.begin(), extension.end(), ub) == extension.end() &&
            std::find(moduli.begin(), moduli.end(), ub) == moduli.end())
            extension.push_back(ub);
    }
    for (size_t lvl = 0; lvl < moduli.size(); lvl++) {
        std::vector<u64> lm(moduli.begin(), moduli.end() - lvl);
        levels.push_back(make_context(lm, N, nullptr));
        size_t bits = 0;
        for (u64 q : lm) bits += 64 - __builtin_clzll(q);
        std::vector<u64> mm = lm;
        mm.insert(mm.end(), extension.begin(), extension.begin() + (bits + 60 + 61) / 62);
        mul.emplace_back(new Multiplicator(levels.back(), mm, BigUint(1), BigUint(1), BigUint(1), BigUint(1),
                                           BigUint(t), levels.back()->rns.product));
    }
}

const Multiplicator& Parameters::at(u32 level) const {
    if (level >= mul.size())
        throw Error(-3, "invalid level " + std::to_string(level) + " (the maximum is " +
                            std::to_string(mul.size() - 1) + ")");
    return *mul[level];
}

void Parameters::ct_mul(u32 level, const u64* lhs, u32 lp, const u64* rhs, u32 rp, u64* out, u32 num,
                        hipStream_t s) {
    at(level);
    Multiplicator& m = *mul[level];
    if (!lp || !rp) throw Error(-1, "empty ciphertext");
    if (lp + rp - 1 > 16) throw Error(-2, "products of more than 16 parts are not supported");
    const u32 parts = lp + rp - 1;
    check_batch(num * std::max(parts, std::max(lp, rp)));
    if (!num) return;
    if (lp == 2 && rp == 2) {
        m.multiply(lhs, rhs, out, num, s);
        return;
    }
    // every part extended (Poly::scale in the Ntt representation), products in
    // the multiplication context, every product part scaled down
    const size_t n = N, M = m.mul->L;
    u64* wa = scratch(Scratch::Mul).get(s, (size_t)num * (lp + rp + parts) * M * n * 8).as<u64>();
    u64* wb = wa + (size_t)num * lp * M * n;
    u64* wc = wb + (size_t)num * rp * M * n;
    m.ext_lhs->scale(lhs, wa, num * lp, true, s);
    m.ext_rhs->scale(rhs, wb, num * rp, true, s);
    launch_tensor_parts(wa, lp, wb, rp, wc, m.mul->mod.as<ModInfo>(), (u32)M, m.mul->logn, num, s);
    m.down->scale(wc, out, num * parts, true, s);
}

// ---------------------------------------------------------------------------
Galois::Galois(std::shared_ptr<Ksk> k, u32 exp) : ksk(std::move(k)) {
    const u32 n = ksk->ct_ctx->N;
    exponent = exp % (2 * n);
    if ((exponent & 1) == 0) throw Error(-1, "The exponent should be odd modulo 2 * degree");
    // SubstitutionExponent::new (math/src/rq/mod.rs:97-120) and the NTT-domain
    // gather q[bitrev[j]] = p[power_bitrev[j]] (rq/mod.rs:324-334) as q[m] = p[perm[m]].
    const int logn = ksk->ct_ctx->logn;
    std::vector<u32> perm(n);
    u64 power = (exponent - 1) / 2;
    for (u32 j = 0; j < n; j++) {
        u32 pb = bit_reverse((u32)(power & (n - 1)), logn);
        perm[bit_reverse(j, logn)] = pb;
        power += exponent;
    }
    d_perm.reset(n * 4);
    HIPCHK(hipMemcpy(d_perm.ptr, perm.data(), d_perm.bytes, hipMemcpyHostToDevice));
}

// GaloisKey::relinearize_into (fhe/src/bfv/keys/galois_key.rs:90-124).
void Galois::apply(const u64* ct, u64* out, u32 inst, hipStream_t s) {
    check_batch(inst);
    const size_t n = ksk->ct_ctx->N, L = ksk->ct_ctx->L;
    const size_t C = (size_t)inst * L * n;
    u64* c2 = scratch(Scratch::GaloisC2).get(s, C * 8).as<u64>();
    InttRowsArgs a{};
    a.src[0] = ct + L * n;
    a.src_inst_stride = 2 * L * n;
    a.dst = c2;
    a.dst_inst_stride = L * n;
    a.inv_tw = ksk->ct_ctx->inv_tw.as<ulonglong2>();
    a.mods = ksk->ct_ctx->mod.as<ModInfo>();
    a.perm = d_perm.as<u32>();
    intt_rows(*ksk->ct_ctx, a, (u32)L, 1, inst, s);
    ksk->switch_and_finish(c2, L * n, out, inst, nullptr, 0, 0, ct, 2 * L * n, 0, 1, d_perm.as<u32>(), s);
}

// ---------------------------------------------------------------------------
std::shared_ptr<Galois> EvaluationKey::find(u32 exponent) const {
    for (auto& g : keys)
        if (g->exponent == exponent % (2 * degree)) return g;
    return nullptr;
}

u32 EvaluationKey::column_exponent(u32 i) const { return (u32)powmod64(3, i, 2 * (u64)degree); }

void EvaluationKey::rotate(u32 exponent, const u64* ct, u64* out, u32 inst, hipStream_t s) const {
    auto g = find(exponent);
    if (!g) throw Error(-1, "This key does not support the requested rotation");
    g->apply(ct, out, inst, s);
}

// EvaluationKey::computes_inner_sum (evaluation_key.rs:56-82): out starts as ct
// and accumulates the rotations of itself.
void EvaluationKey::rotate_and_sum(const std::vector<u32>& exponents, const u64* ct, u64* out, u32 inst,
                                   hipStream_t s) {
    if (keys.empty()) throw Error(-1, "empty evaluation key");
    for (u32 e : exponents)
        if (!find(e)) throw Error(-1, "This key does not support the inner sum functionality");
    const auto& c = *keys[0]->ksk->ct_ctx;
    const size_t sz = (size_t)inst * 2 * c.L * c.N;
    u64* tmp = scratch(Scratch::EvalTmp).get(s, sz * 8).as<u64>();
    if (out != ct) HIPCHK(hipMemcpyAsync(out, ct, sz * 8, hipMemcpyDeviceToDevice, s));
    for (u32 e : exponents) {
        find(e)->apply(out, tmp, inst, s);
        launch_add_rows(out, tmp, out, c.mod.as<ModInfo>(), c.L, c.logn, (size_t)inst * 2 * c.L, s);
    }
}

// The expansion levels a key supports (EvaluationKey::supports_expansion,
// evaluation_key.rs:134-147) and their monomials -X^(N - 2^l)
// (EvaluationKeyBuilder::build, evaluation_key.rs:355-363), transformed by
// the context's own forward NTT.
void EvaluationKey::prepare_expansion() {
    expansion_levels = 0;
    if (keys.empty()) return;
    const DeviceContext& c = *keys[0]->ksk->ct_ctx;
    while (expansion_levels < (u32)c.logn && find((degree >> expansion_levels) + 1)) expansion_levels++;
    if (!expansion_levels) return;
    const size_t n = c.N, L = c.L, rows = (size_t)expansion_levels * L;
    std::vector<u64> h(rows * n, 0);
    for (u32 l = 0; l < expansion_levels; l++)
        for (size_t m = 0; m < L; m++) h[(l * L + m) * n + n - ((size_t)1 << l)] = c.moduli[m] - 1;
    DevBuf tmp(rows * n * 8);
    HIPCHK(hipMemcpy(tmp.ptr, h.data(), tmp.bytes, hipMemcpyHostToDevice));
    launch_ntt_forward(c, RowMap{tmp.as<u64>(), (u32)L, (u32)L, 0, 0}, (u32)rows, false, nullptr);
    HIPCHK(hipStreamSynchronize(nullptr));
    HIPCHK(hipMemcpy(h.data(), tmp.ptr, tmp.bytes, hipMemcpyDeviceToHost));
    std::vector<ulonglong2> hm(rows * n);
    for (size_t r = 0; r < rows; r++) {
        const u64 q = c.moduli[r % L];
        for (size_t j = 0; j < n; j++) {
            const u64 v = h[r * n + j];
            hm[r * n + j] = ulonglong2{v, shoup64(v, q)};
        }
    }
    monomials.reset(hm.size() * sizeof(ulonglong2));
    HIPCHK(hipMemcpy(monomials.ptr, hm.data(), monomials.bytes, hipMemcpyHostToDevice));
}

// EvaluationKey::expands (evaluation_key.rs:153-193): level l relinearizes
// out[0 .. 2^l) with the key of exponent (N >> l) + 1 as one batch, then one
// elementwise pass forms out[2^l + i] and updates out[i].
// With outputs stored query-minor ([j][query]), the ciphertexts 0 .. 2^l of
// every query are contiguous, and the level step of query q, output i is
// flat index i num + q of a step of 2^l num.
void EvaluationKey::expand(const u64* ct, u32 size, u32 num, u64* out, hipStream_t s) const {
    if (keys.empty()) throw Error(-1, "empty evaluation key");
    u32 level = 0;
    while (((u64)1 << level) < size) level++;
    if (level > expansion_levels) throw Error(-1, "This key does not support expansion at this level");
    if (!num) return;
    if (((u64)num << (level ? level - 1 : 0)) > 65535 || (u64)num * std::max<u32>(size, 1) > 0xffffffffu)
        throw Error(-1, "too many queries for this expansion size");
    const DeviceContext& c = *keys[0]->ksk->ct_ctx;
    const size_t sz = (size_t)2 * c.L * c.N;
    HIPCHK(hipMemcpyAsync(out, ct, (size_t)num * sz * 8, hipMemcpyDeviceToDevice, s));
    if (!level) return;
    u64* sub = scratch(Scratch::EvalTmp).get(s, ((size_t)num << (level - 1)) * sz * 8).as<u64>();
    for (u32 l = 0; l < level; l++) {
        const u32 step = num << l;
        find((degree >> l) + 1)->apply(out, sub, step, s);
        launch_expand_step(out, sub, monomials.as<ulonglong2>() + (size_t)l * c.L * c.N, c.mod.as<ModInfo>(), c.L,
                           c.logn, step, size * num, s);
    }
}

// ---------------------------------------------------------------------------
static CodecConst codec_const(u64 q) {
    CodecConst c{};
    c.q = q;
    barrett128(q, c.bl, c.bh);
    return c;
}

Codec::Codec(std::shared_ptr<DeviceContext> c, std::shared_ptr<DeviceContext> p, u64 t_, const u64* psi_t,
             const int64_t* sk_coeffs)
    : ctx(std::move(c)), plain(std::move(p)), t(t_) {
    if (ctx->N != plain->N) throw Error(-4, "Incompatible degrees");
    if (t < 2 || (t >> 62)) throw Error(-1, "plaintext modulus out of range");
    if (plain->moduli[0] != ctx->moduli[0])
        throw Error(-4, "the plaintext context must start with the first ciphertext modulus");
    const u32 n = ctx->N, L = ctx->L;
    // SIMD needs an NTT modulo t (parameters.rs:497-503); coefficient encoding does not.
    if (supports_ntt64(t, n)) tctx = make_context({t}, n, psi_t);
    // delta_i = (-t)^-1 mod q_i (parameters.rs:512-525); q_mod_t = Q mod t (:536-537)
    std::vector<CodecConst> cc(L);
    q_mod_t = 1 % t;
    for (u32 i = 0; i < L; i++) {
        const u64 q = ctx->moduli[i];
        cc[i] = codec_const(q);
        const u64 tq = t % q;
        if (!tq) throw Error(-1, "Inverse failed");
        cc[i].delta = invmod64(q - tq, q);
        cc[i].delta_shoup = shoup64(cc[i].delta, q);
        q_mod_t = mulmod64(q_mod_t, q % t, t);
    }
    q_mod_t_shoup = shoup64(q_mod_t, t);
    t_const = codec_const(t);
    q0_const = codec_const(plain->moduli[0]);
    consts.reset(cc.size() * sizeof(CodecConst));
    HIPCHK(hipMemcpy(consts.ptr, cc.data(), consts.bytes, hipMemcpyHostToDevice));
    // matrix_reps_index_map (parameters.rs:614-629): slot i of the first row
    // and row_size | i of the second are the NTT positions of 3^i and -3^i.
    std::vector<u32> map(n);
    const u64 two_n = 2 * (u64)n;
    u64 pos = 1;
    for (u32 i = 0; i < n / 2; i++) {
        map[i] = bit_reverse((u32)((pos - 1) >> 1), ctx->logn);
        map[n / 2 + i] = bit_reverse((u32)((two_n - pos - 1) >> 1), ctx->logn);
        pos = (pos * 3) & (two_n - 1);
    }
    index_map.reset(n * sizeof(u32));
    HIPCHK(hipMemcpy(index_map.ptr, map.data(), index_map.bytes, hipMemcpyHostToDevice));
    BigUint Q(1);
    for (u64 q : ctx->moduli) Q = Q * BigUint(q);
    dec.reset(new PolyScaler(ctx, plain, BigUint(t), Q));
    if (sk_coeffs) {
        std::vector<u64> h((size_t)L * n);
        for (u32 i = 0; i < L; i++) {
            const u64 q = ctx->moduli[i];
            for (u32 j = 0; j < n; j++) {
                const int64_t v = sk_coeffs[j];
                const u64 r = (v < 0 ? 0 - (u64)v : (u64)v) % q;
                h[(size_t)i * n + j] = v < 0 && r ? q - r : r;
            }
        }
        sk.reset(h.size() * 8);
        HIPCHK(hipMemcpy(sk.ptr, h.data(), sk.bytes, hipMemcpyHostToDevice));
        launch_ntt_forward(*ctx, RowMap{sk.as<u64>(), L, L, 0, 0}, L, false, nullptr);
        HIPCHK(hipStreamSynchronize(nullptr));
    }
}

void Codec::encode(const u64* values, u64* out, u32 num, bool simd, bool scaled, hipStream_t s) const {
    check_batch(num);
    if (!num) return;
    const u32 n = ctx->N, L = ctx->L;
    const u64* m = values;
    if (simd) {
        if (!tctx) throw Error(-1, "The plaintext modulus does not support SIMD encoding");
        u64* rows = scratch(Scratch::Codec).get(s, (size_t)num * n * 8).as<u64>();
        launch_encode_gather(values, index_map.as<u32>(), rows, ctx->logn, num, s);
        launch_ntt_backward(*tctx, RowMap{rows, 1, 1, 0, 0}, num, s);
        m = rows;
    }
    launch_pt_lift(m, out, consts.as<CodecConst>(), t, q_mod_t, q_mod_t_shoup, L, ctx->logn, num, scaled, s);
    launch_ntt_forward(*ctx, RowMap{out, L, L, 0, 0}, num * L, false, s);
}

void Codec::encrypt(const u64* pt, const u64* a, const int64_t* e, u64* ct, u32 num, hipStream_t s) const {
    check_batch(num);
    if (!sk.ptr) throw Error(-1, "no secret key");
    if (!num) return;
    const u32 n = ctx->N, L = ctx->L;
    u64* en = scratch(Scratch::Codec).get(s, (size_t)num * L * n * 8).as<u64>();
    launch_small_lift(e, en, consts.as<CodecConst>(), L, ctx->logn, num, s);
    launch_ntt_forward(*ctx, RowMap{en, L, L, 0, 0}, num * L, false, s);
    launch_encrypt_sk(pt, a, en, sk.as<u64>(), ct, consts.as<CodecConst>(), L, ctx->logn, num, s);
}

// PublicKey::try_encrypt (public_key.rs:47-87) with caller-drawn u, e1, e2.
void Codec::encrypt_pk(const u64* pk, const u64* pt, const int64_t* u, const int64_t* e1, const int64_t* e2, u64* ct,
                       u32 num, hipStream_t s) const {
    check_batch(num);
    if (!num) return;
    const u32 n = ctx->N, L = ctx->L;
    const size_t stride = (size_t)num * L * n;
    u64* sm = scratch(Scratch::Codec).get(s, 3 * stride * 8).as<u64>();
    const int64_t* src[3] = {u, e1, e2};
    for (int k = 0; k < 3; k++)
        launch_small_lift(src[k], sm + k * stride, consts.as<CodecConst>(), L, ctx->logn, num, s);
    launch_ntt_forward(*ctx, RowMap{sm, L, L, 0, 0}, 3 * num * L, false, s);
    launch_encrypt_pk(pt, pk, sm, stride, ct, consts.as<CodecConst>(), L, ctx->logn, num, s);
}

void Codec::decrypt(const u64* ct, u32 parts, u64* values, u32 num, bool simd, hipStream_t s) const {
    check_batch(num);
    if (!sk.ptr) throw Error(-1, "no secret key");
    if (!parts) throw Error(-1, "ciphertexts without parts");
    if (simd && !tctx) throw Error(-1, "The plaintext modulus does not support SIMD encoding");
    if (!num) return;
    const u32 n = ctx->N, L = ctx->L, k = plain->L;
    const size_t a = (size_t)num * L * n, b = (size_t)num * k * n, c = simd ? (size_t)num * n : 0;
    u64* ph = scratch(Scratch::Codec).get(s, (a + b + c) * 8).as<u64>();
    u64* d = ph + a;
    launch_phase(ct, parts, sk.as<u64>(), ph, consts.as<CodecConst>(), L, ctx->logn, num, s);
    launch_ntt_backward(*ctx, RowMap{ph, L, L, 0, 0}, num * L, s);
    dec->scale(ph, d, num, false, s);
    u64* w = simd ? d + b : values;
    launch_decode_reduce(d, k, w, q0_const, t_const, ctx->logn, num, s);
    if (simd) {
        launch_ntt_forward(*tctx, RowMap{w, 1, 1, 0, 0}, num, false, s);
        launch_decode_gather(w, index_map.as<u32>(), values, ctx->logn, num, s);
    }
}

// ---------------------------------------------------------------------------
void pir_respond(const EvaluationKey& ek, Multiplicator& mul, Ksk& rk, const Switcher& sw, const u64* db, u32 dim1,
                 u32 dim2, const u64* queries, u32 num, u64* responses, hipStream_t s) {
    if (ek.keys.empty()) throw Error(-1, "empty evaluation key");
    const DeviceContext& c = *ek.keys[0]->ksk->ct_ctx;
    if (mul.base->moduli != c.moduli || rk.ct_ctx->moduli != c.moduli || sw.from->moduli != c.moduli)
        throw Error(-4, "the expansion key, multiplicator, relinearization key and switcher must share the query context");
    if (mul.rk || mul.mod_switch) throw Error(-1, "the multiplicator must not relinearize or switch");
    if (!dim1 || !dim2) throw Error(-1, "empty database");
    check_batch(num);
    if ((u64)dim2 * num > 65535) throw Error(-1, "at most 65535 products (dim2 x queries) per call");
    // the final switch_to_level runs on both parts of every response
    if ((u64)2 * num > 65535) throw Error(-1, "at most 32767 queries per call");
    if (!num) return;
    const size_t per = (size_t)2 * c.L * c.N, per3 = (size_t)3 * c.L * c.N, size = (size_t)dim1 + dim2;
    const size_t ex_w = size * num * per, dt_w = (size_t)dim2 * num * per, pr_w = (size_t)dim2 * num * per3;
    u64* ex = scratch(Scratch::Pir).get(s, (ex_w + dt_w + pr_w + (size_t)num * per) * 8).as<u64>();
    u64* dots = ex + ex_w;
    u64* prods = dots + dt_w;
    u64* relin = prods + pr_w;
    ek.expand(queries, (u32)size, num, ex, s);
    dot_product_columns(c, ex, dim1, 2 * num, db, dim2, dots, s);
    mul.multiply(dots, ex + (size_t)dim1 * num * per, prods, dim2 * num, s);
    // out += dot_j * ct_j (mulpir.rs:177-179): pairwise sums over the columns
    for (u32 cnt = dim2; cnt > 1;) {
        const u32 half = cnt / 2;
        launch_add_rows(prods, prods + (size_t)(cnt - half) * num * per3, prods, c.mod.as<ModInfo>(), c.L, c.logn,
                        (size_t)half * num * 3 * c.L, s);
        cnt -= half;
    }
    rk.relinearize(prods, relin, num, s);
    sw.switch_down(relin, responses, 2 * num, s);
}

// ---------------------------------------------------------------------------
// Modulus::serialization_length (zq/mod.rs:773-777) per limb: N bits(q_i - 1) / 8.
static WireTable wire_table(const DeviceContext& c, size_t& per_poly) {
    if (c.L > kWireMaxLimbs) throw Error(-2, "serialization supports at most 64 moduli");
    WireTable t{};
    t.L = c.L;
    per_poly = 0;
    for (u32 i = 0; i < c.L; i++) {
        t.nbits[i] = (u32)(64 - __builtin_clzll(c.moduli[i] - 1));
        t.offset[i] = (u32)per_poly;
        per_poly += (size_t)t.nbits[i] * c.N / 8;
    }
    return t;
}

size_t serialization_length(const DeviceContext& c) {
    size_t per = 0;
    wire_table(c, per);
    return per;
}

void serialize_polys(const DeviceContext& c, const u64* in, u32 num, bool ntt, uint8_t* out, hipStream_t s) {
    check_batch(num);
    size_t per = 0;
    const WireTable t = wire_table(c, per);
    if (!num) return;
    const u64* src = in;
    if (ntt) {
        const size_t rows = (size_t)c.L * c.N;
        u64* pb = scratch(Scratch::Wire).get(s, (size_t)num * rows * 8).as<u64>();
        InttRowsArgs ia{};
        ia.src[0] = in;
        ia.src_inst_stride = rows;
        ia.dst = pb;
        ia.dst_inst_stride = rows;
        ia.dst_part_stride = rows;
        ia.inv_tw = c.inv_tw.as<ulonglong2>();
        ia.mods = c.mod.as<ModInfo>();
        intt_rows(c, ia, c.L, 1, num, s);
        src = pb;
    }
    launch_pack(src, out, t, c.logn, (u32)per, num, s);
}

void deserialize_polys(const DeviceContext& c, const uint8_t* in, u32 num, bool ntt, u64* out, hipStream_t s) {
    check_batch(num);
    size_t per = 0;
    const WireTable t = wire_table(c, per);
    if (!num) return;
    launch_unpack(in, out, t, c.logn, (u32)per, num, s);
    if (ntt) launch_ntt_forward(c, RowMap{out, c.L, c.L, 0, 0}, num * c.L, false, s);
}

// ---------------------------------------------------------------------------
void rgsw_external_product(const Ksk& k0, const Ksk& k1, const u64* ct, u64* out, u32 inst, hipStream_t s) {
    check_batch(inst);
    const DeviceContext& c = *k0.ct_ctx;
    for (const Ksk* k : {&k0, &k1})
        if (k->ct_ctx->moduli != c.moduli || k->key_ctx->moduli != c.moduli || k->ct_ctx->N != c.N)
            throw Error(-4, "Inconsistent key switching levels");
    if (!inst) return;
    const size_t n = c.N, L = c.L, part = L * n;
    u64* pb = scratch(Scratch::Rgsw).get(s, (size_t)inst * 4 * part * 8).as<u64>();
    u64* tmp = pb + (size_t)inst * 2 * part;
    InttRowsArgs ia{};
    ia.src[0] = ct;
    ia.src[1] = ct + part;
    ia.src_inst_stride = 2 * part;
    ia.dst = pb;
    ia.dst_inst_stride = 2 * part;
    ia.dst_part_stride = part;
    ia.inv_tw = c.inv_tw.as<ulonglong2>();
    ia.mods = c.mod.as<ModInfo>();
    intt_rows(c, ia, (u32)L, 2, inst, s);
    k0.key_switch_pb(pb, 2 * part, out, inst, false, s);
    k1.key_switch_pb(pb + part, 2 * part, tmp, inst, false, s);
    launch_add_rows(out, tmp, out, c.mod.as<ModInfo>(), (u32)L, c.logn, (size_t)inst * 2 * L, s);
}

// ---------------------------------------------------------------------------
// dot_product_scalar (fhe/src/bfv/ops/dot_product.rs:56-143).
static void check_dot(u32 count, u32 parts) {
    if (!count) throw Error(-1, "At least one iterator is empty");
    if (!parts) throw Error(-1, "ciphertexts without parts");
    if (count > 0x7fffffffu) throw Error(-1, "at most 2^31 - 1 terms per dot product");
}

void dot_product_scalar(const DeviceContext& c, const u64* ct, size_t ct_stride, const u64* pt, size_t pt_stride,
                        u32 count, u32 parts, u64* out, hipStream_t s) {
    check_dot(count, parts);
    DotArgs a{};
    a.ct = ct;
    a.ct_stride = ct_stride;
    a.pt = pt;
    a.pt_row_stride = pt_stride;
    a.out = out;
    a.mods = c.mod.as<ModInfo>();
    a.count = count;
    a.columns = 1;
    a.words = c.L * c.N;
    a.logn = (u32)c.logn;
    launch_dot_product(a, parts, s);
}

// The column products of a PIR server (examples/mulpir.rs:170-179): column
// j is the plaintexts db[r][j] (database.iter().skip(j).step_by(columns)).
void dot_product_columns(const DeviceContext& c, const u64* ct, u32 rows, u32 parts, const u64* db, u32 columns,
                         u64* out, hipStream_t s) {
    check_dot(rows, parts);
    if (!columns) return;
    const size_t words = (size_t)c.L * c.N;
    DotArgs a{};
    a.ct = ct;
    a.ct_stride = parts * words;
    a.pt = db;
    a.pt_row_stride = columns * words;
    a.pt_col_stride = words;
    a.out = out;
    a.out_col_stride = parts * words;
    a.mods = c.mod.as<ModInfo>();
    a.count = rows;
    a.columns = columns;
    a.words = (u32)words;
    a.logn = (u32)c.logn;
    launch_dot_product(a, parts, s);
}

// ---------------------------------------------------------------------------
StreamScratch& scratch(Scratch k) {
    // Never destroyed: device memory is released when the process ends.
    static StreamScratch* pools = new StreamScratch[(int)Scratch::Count];
    return pools[(int)k];
}

void release_scratch(hipStream_t s) {
    for (int k = 0; k < (int)Scratch::Count; k++) scratch((Scratch)k).release(s);
}

void StreamScratch::release(hipStream_t s) const {
    std::lock_guard<std::mutex> g(m_);
    ws_.erase(s);
}

namespace {
thread_local FixedArena* t_arena = nullptr;
}

ArenaScope::ArenaScope(FixedArena& a) : prev_(t_arena) { t_arena = &a; }
ArenaScope::~ArenaScope() { t_arena = prev_; }

Workspace& FixedArena::carve(size_t bytes) {
    if (views_ == kMaxViews) throw Error(-1, "too many scratch regions for the fixed scratch");
    const size_t need = aligned(bytes);
    if (need > size_ - used_)
        throw Error(-5, "the fixed scratch region holds " + std::to_string(size_) + " bytes; this call needs more");
    Workspace& w = view_[views_++];
    w.view = base_ + used_;
    used_ += need;
    return w;
}

Workspace& StreamScratch::get(hipStream_t s, size_t bytes) const {
    if (t_arena) return t_arena->carve(bytes);
    std::lock_guard<std::mutex> g(m_);
    auto& w = ws_[s];
    if (!w) w.reset(new Workspace);
    if (bytes > w->buf.bytes && w->buf.ptr) {
        // Work already queued on this stream may still use the old buffer.
        HIPCHK(hipStreamSynchronize(s));
    }
    w->ensure(bytes);
    return *w;
}

void Workspace::ensure(size_t bytes) {
    if (bytes <= buf.bytes) return;
    buf.reset(bytes);
}

}  // namespace fhehip
