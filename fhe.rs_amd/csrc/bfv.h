// Host-side objects of the BFV evaluation path.
#pragma once

#include "bfv_launch.h"
#include "rns_device.h"
#include "scale_mfma.h"
#include "tuning.h"

#include <map>
#include <mutex>

namespace fhehip {

// Device scratch owned by a handle; grows to the largest batch seen. A view
// (non-owning) is a region of a FixedArena instead.
struct Workspace {
    DevBuf buf;
    void* view = nullptr;
    void ensure(size_t bytes);
    template <class T>
    T* as() const { return (T*)(view ? view : buf.ptr); }
};

// A fixed device region, owned by the caller, that every scratch request made
// on the calling thread inside an ArenaScope is carved from, in request order
// (each region 256-byte aligned), instead of the per-stream pools: no device
// allocation happens. A request that does not fit throws before its kernels
// are issued.
class FixedArena {
  public:
    static constexpr size_t kAlign = 256;
    static constexpr int kMaxViews = 16;
    FixedArena(void* base, size_t bytes) : base_((uint8_t*)base), size_(bytes) {}
    Workspace& carve(size_t bytes);
    size_t used() const { return used_; }
    static size_t aligned(size_t bytes) { return (bytes + kAlign - 1) / kAlign * kAlign; }

  private:
    uint8_t* base_;
    size_t size_, used_ = 0;
    int views_ = 0;
    Workspace view_[kMaxViews];
};

// Routes the calling thread's scratch requests to `arena` while alive.
struct ArenaScope {
    explicit ArenaScope(FixedArena& a);
    ~ArenaScope();
    ArenaScope(const ArenaScope&) = delete;
    ArenaScope& operator=(const ArenaScope&) = delete;

  private:
    FixedArena* prev_;
};

// Scratch kept per stream: calls issued on different streams never share a
// buffer, so one handle may serve several streams at once, while calls on one
// stream run in stream order and reuse theirs.
class StreamScratch {
  public:
    // Buffer of at least `bytes` for work on stream `s`.
    Workspace& get(hipStream_t s, size_t bytes) const;

    // Frees the buffer of stream `s` (the stream's work must be complete).
    void release(hipStream_t s) const;

  private:
    mutable std::mutex m_;
    mutable std::map<hipStream_t, std::unique_ptr<Workspace>> ws_;
};

// Process-wide scratch pools by purpose. Work on one stream runs in order, so
// every handle shares the pool of each purpose per stream; purposes that are
// live at the same time (a multiplication and the key switch it calls) use
// different pools.
enum class Scratch { Scale, RelinC2, Lift, SwitchDown, Mul, GaloisC2, EvalTmp, ModSwitch, Codec, Rgsw, Wire, Pir, Count };
StreamScratch& scratch(Scratch k);
// Frees every pool's buffer of stream `s`.
void release_scratch(hipStream_t s);

std::vector<u64> scaler_blob(const ScalerConsts& c, const std::vector<ModInfo>& from_mods,
                             const std::vector<ModInfo>& to_mods, u32 start);

// rq::scaler::Scaler (math/src/rq/scaler.rs:18-126).
struct PolyScaler {
    std::shared_ptr<DeviceContext> from, to;
    bool is_one = true;
    u32 common = 0;  // number_common_moduli
    ScalerConsts consts;
    DevBuf d_consts;  // outputs start .. to->L
    u32 words() const { return (u32)(d_consts.bytes / sizeof(u64)); }
    // Digit tables of the matrix-core form (empty when the shape has none).
    DevBuf d_mfma;
    ScaleMfmaLayout mfma;
    // The scaling of `parts` power-basis polynomials per instance (rows of
    // part k of instance b at in + b in_inst + k in_part, likewise out) on
    // the matrix cores when the inputs are reduced residues, the shape has
    // that form and the degree's tuning selects it, else on the VALU.
    void scale_rows(const u64* in, size_t in_inst, size_t in_part, u64* out, size_t out_inst, size_t out_part,
                    u32 parts, u32 inst, bool reduced, hipStream_t s, const char* name) const;
    // A paired launch (pair()) likewise; its inputs are reduced residues, or
    // open halves below 2p with the inverse stage.
    void scale_pair_rows(const ScalePairArgs& a, u32 parts, u32 inst, hipStream_t s, const char* name) const;
    // A paired launch (launch_scale_pair) of these constants; in_open: the
    // inputs are half-row inverse transforms, out_fwd: the outputs get the
    // first forward stage of the destination context.
    ScalePairArgs pair(const u64* in, size_t in_inst, size_t in_part, u64* out, size_t out_inst, size_t out_part,
                       bool in_open, bool out_fwd) const;
    PolyScaler(std::shared_ptr<DeviceContext> f, std::shared_ptr<DeviceContext> t, const BigUint& num,
               const BigUint& den);
    void scale(const u64* in, u64* out, u32 polys, bool ntt_repr, hipStream_t s);
};

// KeySwitchingKey (fhe/src/bfv/keys/key_switching_key.rs:22-45) with resident
// NttShoup rows and the switch-down constants from the key context back to the
// ciphertext context.
struct Ksk {
    std::shared_ptr<DeviceContext> ct_ctx, key_ctx;
    u32 digits = 0, extra = 0, sd_stride = 0;
    // One-modulus key context: power-of-two decomposition into digits of
    // log_base bits (key_switching_key.rs:70-86, 292-329); 0: one digit per
    // ciphertext modulus.
    u32 log_base = 0;
    // The 32-byte seed the c1 rows were drawn from (KeySwitchingKey::seed,
    // key_switching_key.rs:108-123) when the key came from a seeded message;
    // empty otherwise.
    std::vector<uint8_t> seed;
    DevBuf key0, key1;  // [digits][L'][N] (value, shoup)
    // Rows of the key limbs below the dropped one with the switch-down factor
    // q_l^-1 folded in (plain values, [digits][L][N]); used by the fused
    // finish when the key is at most one level above the ciphertext.
    DevBuf fin0, fin1;
    DevBuf extra_mods, sd_consts;
    Ksk(std::shared_ptr<DeviceContext> ct, std::shared_ptr<DeviceContext> key, const u64* c0, const u64* c1,
        u32 digits);
    KeySwitchArgs args(const u64* src, size_t src_inst_stride) const;
    // Scratch bytes of the lifted digits and of the switch-down rows of one
    // switch_and_finish over `inst` instances.
    size_t lift_bytes(u32 inst) const;
    size_t finish_bytes(u32 inst) const;
    // Whether a finish over `inst` instances takes the paired form
    // (launch_ks_z_pair, then the half transforms).
    bool paired_finish(u32 inst) const;
    // Digit lift + lazy forward NTT into scratch [inst][digits][L'][N].
    // `rows` (optional): rows of the ciphertext context transformed alongside
    // (launch_lift_ntt_split).
    u64* lift(const u64* src, size_t src_inst_stride, u32 inst, hipStream_t s,
              const HalfFwdRows* rows = nullptr) const;
    // Full key switch into acc [inst][2][L'][N] (NTT domain, or power basis
    // when to_power_basis).
    void key_switch_pb(const u64* src, size_t src_inst_stride, u64* acc, u32 inst, bool to_power_basis,
                       hipStream_t s) const;
    SdNttArgs sd_args(u64* out, const u64* add_pb, size_t add_inst, size_t add_part, const u64* add_ntt,
                      size_t add_ntt_inst, size_t add_ntt_part, u32 add_ntt_parts, const u32* perm) const;
    // Key switch of power-basis digits followed by switch_down_to the
    // ciphertext context, addends and forward NTT into out [inst][2][L][N]:
    // out = NTT(add_pb) + switch_down(key_switch(src)) + add_ntt[perm].
    // `out` may alias add_ntt only without perm.
    void switch_and_finish(const u64* src, size_t src_inst_stride, u64* out, u32 inst, const u64* add_pb,
                           size_t add_inst, size_t add_part, const u64* add_ntt, size_t add_ntt_inst,
                           size_t add_ntt_part, u32 add_ntt_parts, const u32* perm, hipStream_t s) const;
    void key_switch(const u64* p, u64* out, u32 inst, hipStream_t s) const;
    void relinearize(const u64* ct3, u64* out, u32 inst, hipStream_t s);
};

// Multiplicator (fhe/src/bfv/ops/mul.rs:22-227).
// Ciphertext::switch_down / switch_to_level (fhe/src/bfv/ciphertext.rs:86-119):
// NTT-domain polynomials of `from` rounded into `to`, the first L_to moduli
// of `from`, by one Poly::switch_down (math/src/rq/mod.rs:390-449) per
// dropped modulus. The NTT round trips between fhe.rs's steps are exact, so
// one inverse transform, the steps in power basis and one forward transform
// give the same result.
struct Switcher {
    std::shared_ptr<DeviceContext> from, to;
    DevBuf consts;  // [dropped][L_from - 1] SdConst: step s drops modulus L_from - 1 - s
    Switcher(std::shared_ptr<DeviceContext> from, std::shared_ptr<DeviceContext> to);
    // [polys][L_from][N] -> [polys][L_to][N]; out must not alias in.
    void switch_down(const u64* in, u64* out, u32 polys, hipStream_t s) const;
    // Scratch bytes of one switch_down of `polys` polynomials.
    size_t scratch_bytes(u32 polys) const;
};

struct Multiplicator {
    std::shared_ptr<DeviceContext> base, mul;
    std::unique_ptr<PolyScaler> ext_lhs, ext_rhs, down;
    bool ext_same = false;  // both operands are extended with the same factor (one launch for both)
    std::shared_ptr<Ksk> rk;
    // Multiplicator::enable_mod_switching (mul.rs:146-155): the product is
    // switched down to the next context.
    std::unique_ptr<Switcher> mod_switch;
    Multiplicator(std::shared_ptr<DeviceContext> base, const std::vector<u64>& extended_basis, const BigUint& lhs_num,
                  const BigUint& lhs_den, const BigUint& rhs_num, const BigUint& rhs_den, const BigUint& post_num,
                  const BigUint& post_den);
    static std::unique_ptr<Multiplicator> make_default(std::shared_ptr<Ksk> rk, u64 plaintext);
    void multiply(const u64* lhs, const u64* rhs, u64* out, u32 inst, hipStream_t s);
    // Bytes of the multiplication's own workspace, and of all scratch one
    // multiply over `inst` instances requests (each region aligned as a
    // FixedArena carves it).
    size_t mul_bytes(u32 inst) const;
    size_t scratch_bytes(u32 inst) const;
};

// BfvParameters (fhe/src/bfv/parameters.rs:431-641) on the device: the level
// chain (context_at_level(l) holds the first L - l moduli), the plaintext
// context and each level's MultiplicationParameters (parameters.rs:585-612,
// 694-717): the level's moduli extended by the first ceil((bits + 60) / 62)
// of L + 1 generated 62-bit primes, an extender of factor one and a t / Q_l
// down-scaler, held as a Multiplicator without relinearization.
struct Parameters {
    u32 N = 0;
    u64 t = 0;
    u32 variance = 10;  // the error variance of BfvParametersBuilder (parameters.rs:318-334)
    std::vector<u64> moduli, extension;
    std::vector<std::shared_ptr<DeviceContext>> levels;
    std::shared_ptr<DeviceContext> plain;
    std::vector<std::unique_ptr<Multiplicator>> mul;
    Parameters(u32 degree, u64 plaintext, const std::vector<u64>& moduli);
    const Multiplicator& at(u32 level) const;
    // Mul<&Ciphertext> for &Ciphertext (fhe/src/bfv/ops/mod.rs:259-342) at
    // `level` on num pairs: lhs [num][lp][L][N], rhs [num][rp][L][N] (NTT) ->
    // out [num][lp + rp - 1][L][N].
    void ct_mul(u32 level, const u64* lhs, u32 lp, const u64* rhs, u32 rp, u64* out, u32 num, hipStream_t s);
};

// Times m.multiply on `inst` instances (zero inputs on a private stream)
// under each choice of the fused kernels' launch shapes, one knob at a time
// starting from the current choice, keeps the fastest for the degree and
// returns it. ms: milliseconds per multiply for xcd 0/1, ks_mac instances
// 2/4/8, tensor 4/5, lift 4/5, finish 4/5 log2 values per thread.
FusedTuning autotune_multiply(Multiplicator& m, u32 inst, float (&ms)[19]);

// fhehip_multiply with all scratch in the library's fixed device region
// (no device allocation); refuses batches whose scratch does not fit.
void multiply_noalloc(Multiplicator& m, const u64* lhs, const u64* rhs, u64* out, u32 inst, hipStream_t s);

// GaloisKey (fhe/src/bfv/keys/galois_key.rs:15-124).
struct Galois {
    std::shared_ptr<Ksk> ksk;
    u32 exponent = 1;
    DevBuf d_perm;
    Galois(std::shared_ptr<Ksk> k, u32 exponent);
    void apply(const u64* ct, u64* out, u32 inst, hipStream_t s);
};

// EvaluationKey (fhe/src/bfv/keys/evaluation_key.rs:56-131): Galois keys by
// exponent, column/row rotations and inner sums.
struct EvaluationKey {
    std::vector<std::shared_ptr<Galois>> keys;
    u32 degree = 0;
    // Oblivious expansion (evaluation_key.rs:153-193): levels l < expansion_levels
    // have the key of exponent (N >> l) + 1; monomials [level][L][N] holds
    // -X^(N - 2^l) in the NTT domain as (value, Shoup) pairs.
    u32 expansion_levels = 0;
    DevBuf monomials;
    // Finds the expansion levels and builds their monomials (synchronous).
    void prepare_expansion();
    // EvaluationKey::expands of num queries: ct [num][2][L][N] -> out
    // [max(size, 1)][num][2][L][N] (output j of every query together).
    void expand(const u64* ct, u32 size, u32 num, u64* out, hipStream_t s) const;
    std::shared_ptr<Galois> find(u32 exponent) const;
    // 3^i mod 2N (evaluation_key.rs:195-203)
    u32 column_exponent(u32 i) const;
    void rotate(u32 exponent, const u64* ct, u64* out, u32 inst, hipStream_t s) const;
    // out = ct; for each listed exponent: out += rotate(out)
    void rotate_and_sum(const std::vector<u32>& exponents, const u64* ct, u64* out, u32 inst, hipStream_t s);
};

// BfvParameters at one level for the steps around the evaluation path
// (parameters.rs:508-560, 614-629) and an optional SecretKey: SIMD /
// coefficient encoding (Plaintext::try_encode, plaintext_vec.rs:37-101, and
// Plaintext::to_poly, plaintext.rs:71-98), SecretKey::encrypt_poly
// (secret_key.rs:102-137) with caller-drawn samples, and try_decrypt
// (secret_key.rs:200-282) with decoding.
struct Codec {
    std::shared_ptr<DeviceContext> ctx, plain, tctx;  // level, plaintext context, {t} (null: no SIMD)
    std::unique_ptr<PolyScaler> dec;                   // ctx -> plain by t / Q (cipher_plain_context.scaler)
    u64 t = 0, q_mod_t = 0, q_mod_t_shoup = 0;
    CodecConst t_const{}, q0_const{};
    DevBuf consts;     // [L] CodecConst of the level's moduli
    DevBuf index_map;  // [N] u32 (matrix_reps_index_map)
    DevBuf sk;         // [L][N] NTT-domain secret (empty without a key)
    Codec(std::shared_ptr<DeviceContext> ctx, std::shared_ptr<DeviceContext> plain, u64 t, const u64* psi_t,
          const int64_t* sk_coeffs);
    // values [num][N] (< t) -> [num][L][N]: `scaled` the to_poly polynomial, else poly_ntt.
    void encode(const u64* values, u64* out, u32 num, bool simd, bool scaled, hipStream_t s) const;
    // ct [num][2][L][N] = [e - a s + pt, a]; pt, a [num][L][N] NTT; e [num][N] signed.
    void encrypt(const u64* pt, const u64* a, const int64_t* e, u64* ct, u32 num, hipStream_t s) const;
    // ct [num][2][L][N] = [u pk0 + e1 + pt, u pk1 + e2]; pk [2][L][N] at this
    // level; u, e1, e2 [num][N] signed.
    void encrypt_pk(const u64* pk, const u64* pt, const int64_t* u, const int64_t* e1, const int64_t* e2, u64* ct,
                    u32 num, hipStream_t s) const;
    // ct [num][parts][L][N] -> values [num][N].
    void decrypt(const u64* ct, u32 parts, u64* values, u32 num, bool simd, hipStream_t s) const;
};

// Rq wire format (math/src/rq/convert.rs:15-87): bytes of one polynomial's
// coefficients, and num polynomials [num][L][N] (NTT domain when ntt, which
// are serialized in power basis) to / from their payloads back to back.
size_t serialization_length(const DeviceContext& c);
void serialize_polys(const DeviceContext& c, const u64* in, u32 num, bool ntt, uint8_t* out, hipStream_t s);
void deserialize_polys(const DeviceContext& c, const uint8_t* in, u32 num, bool ntt, u64* out, hipStream_t s);

// The PIR server response of examples/mulpir.rs:151-185 for `num` queries
// [num][2][L][N] over a [dim1][dim2][L][N] database (NTT-domain plaintext
// rows): expansion to dim1 + dim2 (query-minor), the dim2 column products,
// their products with the second-dimension ciphertexts (`mul`: no
// relinearization), the sum, relinearization by `rk` and `sw` (to the
// response level) into responses [num][2][L_out][N].
void pir_respond(const EvaluationKey& ek, Multiplicator& mul, Ksk& rk, const Switcher& sw, const u64* db, u32 dim1,
                 u32 dim2, const u64* queries, u32 num, u64* responses, hipStream_t s);

// Ciphertext x RGSWCiphertext (fhe/src/bfv/rgsw_ciphertext.rs:123-160):
// ct [inst][2][L][N] NTT -> out [inst][2][L][N] = key_switch(c0, k0) +
// key_switch(c1, k1); both keys at the ciphertext's level.
void rgsw_external_product(const Ksk& k0, const Ksk& k1, const u64* ct, u64* out, u32 inst, hipStream_t s);

// dot_product_scalar (fhe/src/bfv/ops/dot_product.rs:56-143) of `count`
// ciphertexts [parts][L][N] (ct_stride words apart) with NTT-domain plaintext
// polynomials [L][N] (pt_stride words apart) of context c, into out [parts][L][N].
void dot_product_scalar(const DeviceContext& c, const u64* ct, size_t ct_stride, const u64* pt, size_t pt_stride,
                        u32 count, u32 parts, u64* out, hipStream_t s);
// Every column of a [rows][columns][L][N] database against rows ciphertexts
// [rows][parts][L][N]: out [columns][parts][L][N].
void dot_product_columns(const DeviceContext& c, const u64* ct, u32 rows, u32 parts, const u64* db, u32 columns,
                         u64* out, hipStream_t s);

}  // namespace fhehip
