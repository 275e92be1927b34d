/*
 * fhehip — MI355X-native RNS-BFV evaluation path (C ABI).
 *
 * A drop-in for the hot path of fhe.rs (crates fhe-math and fhe): negacyclic
 * NTTs, RNS scaling / base extension, key switching, ciphertext multiplication
 * with relinearization, and Galois rotations; around it modulus switching,
 * the PIR server path (oblivious expansion, plaintext dot products), the
 * RGSW external product, ciphertext arithmetic, encoding / encryption /
 * decryption and the wire format's coefficient payloads. Every entry point
 * is extern "C", takes plain pointers and sizes, never aborts, and returns a
 * status code
 * (FHEHIP_OK == 0). All outputs are exact residues, identical to those of the
 * fhe.rs CPU path for the same inputs and NTT roots.
 *
 * Memory: every `d_*` argument is a device pointer (HBM-resident); handles own
 * their device tables (twiddles, Shoup constants, scaler constants, keys).
 * Polynomials are laid out like fhe.rs' Array2<u64>: [limb][degree] row major,
 * and batches of them back to back. Ciphertexts are [part][limb][degree].
 * `stream` is a hipStream_t (NULL = default stream); calls are asynchronous
 * with respect to the host unless stated otherwise. Scratch memory is kept
 * per stream: calls on different streams may run concurrently, also on one
 * handle; calls on one stream run in stream order. Degrees 8 .. 131072:
 * rows of up to 16384 coefficients are transformed in one workgroup, longer
 * rows as halves (one elementwise stage pass per halving); above 32768 the
 * BFV operations run as elementwise passes around those transforms.
 */
#ifndef FHEHIP_H
#define FHEHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (fhe_math::Error / fhe::Error classes, math/src/errors.rs:11-40,
 * fhe/src/errors.rs:15-130). */
#define FHEHIP_OK 0
#define FHEHIP_ERR_INVALID_ARG (-1)   /* bad pointer, size or parameter            */
#define FHEHIP_ERR_UNSUPPORTED (-2)   /* e.g. NttOperator::new returning None      */
#define FHEHIP_ERR_BAD_LEVEL (-3)     /* ciphertext at an unexpected level/size     */
#define FHEHIP_ERR_CTX_MISMATCH (-4)  /* polynomial/key context does not match      */
#define FHEHIP_ERR_OOM (-5)           /* device allocation failed                   */
#define FHEHIP_ERR_DEVICE (-6)        /* HIP runtime error                          */

typedef struct fhehip_context fhehip_context;
typedef struct fhehip_scaler fhehip_scaler;
typedef struct fhehip_ksk fhehip_ksk;
typedef struct fhehip_multiplicator fhehip_multiplicator;
typedef struct fhehip_galois fhehip_galois;
typedef struct fhehip_evaluation_key fhehip_evaluation_key;
typedef struct fhehip_switcher fhehip_switcher;
typedef struct fhehip_codec fhehip_codec;

/* Text of the last error raised on the calling thread. */
const char* fhehip_last_error(void);

/* Library version string. */
const char* fhehip_version(void);

/* ---- Contexts and NTT ---------------------------------------------------
 * fhehip_context mirrors rq::Context::new (math/src/rq/context.rs:42-93) with
 * one NttOperator::new (math/src/ntt/native.rs:35-73) per modulus. `psi` may
 * hold, per modulus, the primitive 2N-th root to use; NULL selects fhe.rs's
 * own root (fhehip_default_psi). Returns FHEHIP_ERR_UNSUPPORTED when a
 * modulus is not an NTT-friendly prime for `degree` (the None case of
 * NttOperator::new, ntt/mod.rs:19-23). */
int fhehip_context_create(const uint64_t* moduli, uint32_t num_moduli, uint32_t degree, const uint64_t* psi,
                          fhehip_context** out);
void fhehip_context_destroy(fhehip_context* ctx);
/* Number of moduli and degree of a context. */
int fhehip_context_info(const fhehip_context* ctx, uint32_t* num_moduli, uint32_t* degree);
/* The roots the context uses, one per modulus. */
int fhehip_context_psi(const fhehip_context* ctx, uint64_t* psi_out);

/* In-place forward NTT of `num_polys` polynomials [num_polys][L][N] in the
 * context (Poly::into_ntt, math/src/rq/mod.rs:294-303; NttOperator::forward,
 * native.rs:77-102). lazy != 0 leaves outputs in [0, 4p) like
 * forward_vt_lazy (native.rs:142-175); otherwise outputs are reduced. */
int fhehip_ntt_forward(const fhehip_context* ctx, uint64_t* d_polys, uint32_t num_polys, int lazy, void* stream);
/* In-place inverse NTT (Poly::into_power_basis, rq/mod.rs:305-313;
 * NttOperator::backward, native.rs:106-132). Inputs must be < 2p. */
int fhehip_ntt_backward(const fhehip_context* ctx, uint64_t* d_polys, uint32_t num_polys, void* stream);
/* No fhe.rs counterpart: times this degree's forward and inverse transforms
 * of `rows` rows with 16 and 32 values per thread and keeps the faster shape
 * per direction for all later transforms of the degree (process-wide).
 * log2 values per thread chosen -> *fwd, *inv; ms_out (may be NULL) receives
 * milliseconds per launch {fwd16, fwd32, inv16, inv32}. Synchronous. */
int fhehip_ntt_autotune(const fhehip_context* ctx, uint32_t rows, uint32_t* fwd, uint32_t* inv, float* ms_out);
/* Sets the shape (4 or 5 log2 values per thread) of the plain NTT kernels for
 * one degree directly, as the autotuner does. */
int fhehip_ntt_set_shape(uint32_t degree, uint32_t fwd, uint32_t inv);
/* No fhe.rs counterpart. Launches of at most `max_rows` polynomial rows at
 * degrees 4096 .. 16384 transform each row as two halves of N/2 on two
 * workgroups (degree 32768 always does), which shortens small launches
 * (single calls, small serving batches); 0 keeps every row on one workgroup.
 * Process-wide; the default is 128. *previous (may be NULL) receives the old
 * value. Results are identical either way. */
int fhehip_split_row_limit(uint32_t max_rows, uint32_t* previous);

/* ---- Scaling (RNS base extension and t/Q rescaling) ----------------------
 * A scaling factor numerator/denominator as little-endian 32-bit words
 * (ScalingFactor, math/src/rns/scaler.rs:20-47). A NULL factor, or NULL
 * words, stand for one. */
typedef struct fhehip_scaling_factor {
    const uint32_t* num;
    uint32_t num_len;
    const uint32_t* den;
    uint32_t den_len;
} fhehip_scaling_factor;

/* rq::scaler::Scaler::new (math/src/rq/scaler.rs:27-49) with the RnsScaler of
 * math/src/rns/scaler.rs:79-175. Contexts must share the degree. */
int fhehip_scaler_create(const fhehip_context* from, const fhehip_context* to, const fhehip_scaling_factor* factor,
                         fhehip_scaler** out);
void fhehip_scaler_destroy(fhehip_scaler* s);
/* Poly::scale (math/src/rq/mod.rs:623-625 -> rq/scaler.rs:52-126) of
 * `num_polys` polynomials [num][L_from][N] into [num][L_to][N]; ntt_repr != 0
 * for NTT-domain polynomials (Ntt representation), 0 for power basis. The
 * scaler's workspace makes concurrent calls on one handle unsafe. */
int fhehip_scaler_scale(fhehip_scaler* s, const uint64_t* d_in, uint64_t* d_out, uint32_t num_polys, int ntt_repr,
                        void* stream);

/* ---- Key switching ---------------------------------------------------------
 * KeySwitchingKey (fhe/src/bfv/keys/key_switching_key.rs:22-45) from its
 * NTT-domain rows c0, c1 given as HOST arrays [digits][L_key][N]; digits must
 * equal the number of ciphertext moduli (RNS-limb decomposition,
 * key_switching_key.rs:89-104), or, for a key context of one modulus q, the
 * ceil(log_modulus / log_base) rows of the power-of-two decomposition with
 * log_modulus = log2(next_power_of_two(q)) and log_base = log_modulus / 2
 * (key_switching_key.rs:70-86, 292-329; Galois keys at the last level). The
 * key context must start with the
 * ciphertext context's moduli and may extend it by any number of moduli (a
 * key at any lower level, RelinearizationKey::new_leveled and GaloisKey::new,
 * relinearization_key.rs:34-67, galois_key.rs:26-60): they are removed by
 * switch_down_to after the switch (math/src/rq/mod.rs:455-462). The rows are
 * copied to the device with their Shoup companions (NttShoup representation). */
int fhehip_ksk_create(const fhehip_context* ct_ctx, const fhehip_context* key_ctx, const uint64_t* c0,
                      const uint64_t* c1, uint32_t digits, fhehip_ksk** out);
void fhehip_ksk_destroy(fhehip_ksk* k);
/* KeySwitchingKey::key_switch (key_switching_key.rs:214-241): power-basis
 * polynomials [num][L_ct][N] -> NTT-domain (c0, c1) pairs [num][2][L_key][N]. */
int fhehip_key_switch(const fhehip_ksk* k, const uint64_t* d_p, uint64_t* d_out, uint32_t num, void* stream);
/* RelinearizationKey::relinearizes (fhe/src/bfv/keys/relinearization_key.rs:71-98)
 * on NTT-domain 3-part ciphertexts [num][3][L][N] -> [num][2][L][N]. */
int fhehip_relinearize(fhehip_ksk* rk, const uint64_t* d_ct3, uint64_t* d_out, uint32_t num, void* stream);

/* ---- Multiplication ---------------------------------------------------------
 * Multiplicator::default (fhe/src/bfv/ops/mul.rs:101-130): extended basis of
 * ceil((bits(Q) + 60) / 62) generated 62-bit primes, extension by one, t/Q
 * down-scaling, relinearization with `rk`. */
int fhehip_multiplicator_default(const fhehip_ksk* rk, uint64_t plaintext_modulus, fhehip_multiplicator** out);
/* Multiplicator::new (mul.rs:37-98): custom factors and extended basis (the
 * full multiplication context moduli). NULL factors stand for one. */
int fhehip_multiplicator_create(const fhehip_context* base_ctx, const uint64_t* extended_basis, uint32_t ext_len,
                                const fhehip_scaling_factor* lhs, const fhehip_scaling_factor* rhs,
                                const fhehip_scaling_factor* post, fhehip_multiplicator** out);
/* Multiplicator::enable_relinearization (mul.rs:133-142). */
int fhehip_multiplicator_enable_relinearization(fhehip_multiplicator* m, const fhehip_ksk* rk);
/* Multiplicator::enable_mod_switching (mul.rs:146-155): every product is then
 * switched down (Ciphertext::switch_down, mul.rs:221-223) into `next`, the
 * base context without its last modulus; fhehip_multiply writes
 * [num][2][L-1][N] (or [num][3][L-1][N] without relinearization). Refused
 * (FHEHIP_ERR_CTX_MISMATCH) when `next` is not that context, and
 * (FHEHIP_ERR_INVALID_ARG) when the base context has one modulus only. Job
 * batches, the request server and the multiply-and-sum graph do not take
 * such a Multiplicator. */
int fhehip_multiplicator_enable_mod_switching(fhehip_multiplicator* m, const fhehip_context* next);
void fhehip_multiplicator_destroy(fhehip_multiplicator* m);
/* Moduli of the multiplication context; *len is in/out (capacity / count). */
int fhehip_multiplicator_extended_basis(const fhehip_multiplicator* m, uint64_t* out, uint32_t* len);
/* Multiplicator::multiply (mul.rs:158-227) on `num` independent pairs:
 * lhs, rhs [num][2][L][N] NTT -> out [num][2][L][N] (or [num][3][L][N]
 * without relinearization). One call runs the whole batch as a fixed
 * sequence of batched kernels. */
int fhehip_multiply(fhehip_multiplicator* m, const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out,
                    uint32_t num, void* stream);

/* ---- Autotuning of the multiplication's fused kernels ------------------------
 * No fhe.rs counterpart. Launch shapes of the fused row kernels, per degree
 * and process-wide: the XCD-aware workgroup order of the 1-D grids (1) or
 * the plain order (0); the instances that share one load of the key rows in
 * the key inner product (2, 4 or 8); the log2 values per thread (4: N/16
 * threads, 5: N/32 threads) of the fused tensor product + inverse NTT, of
 * the digit lift + forward NTT and of the key-switch finish; at N = 4096 ..
 * 16384, which stages run as half rows (two workgroups per CU) in launches
 * of any size, a mask `halves` (without its bit, only launches of at most
 * fhehip_split_row_limit rows do): 1 the key-switch finish, 2 the
 * multiplication's input inverse transforms, base extension and forward
 * transform, 4 the tensor product + inverse transform and the t/Q scaling
 * (2 and 4 pair the halves inside the scalers); whether the base extension
 * and t/Q scaling of unpaired rows form their column sums on the matrix
 * cores (scaler_mfma 1: 8-bit integer MFMA over base-256 digits, for up to 16
 * source moduli and 8 outputs from N = 256; otherwise, or with 0, one thread
 * per column on the VALU). Every choice gives the same, exact results. */
typedef struct fhehip_tuning {
    uint32_t xcd_order;
    uint32_t ks_mac_instances;
    uint32_t tensor_log2_values;
    uint32_t lift_log2_values;
    uint32_t finish_log2_values;
    uint32_t halves;
    uint32_t scaler_mfma;
} fhehip_tuning;
/* Times fhehip_multiply on `num` instances (pseudorandom residues, a
 * private stream) under every choice of each knob in turn, starting from the
 * degree's current choice, keeps the fastest for the multiplicator's degree
 * (8 .. 32768) and writes it to *chosen; ms_out (may be NULL) receives 19
 * milliseconds per multiply: xcd_order 0, 1; ks_mac_instances 2, 4, 8;
 * tensor 4, 5; lift 4, 5; finish 4, 5; scaler_mfma 0, 1; then halves bit 1
 * off, on; bit 2 off, on; bit 4 off, on. Synchronous. Multiplications of the same degree running
 * meanwhile on other threads may use the trial choices. */
int fhehip_multiply_autotune(fhehip_multiplicator* m, uint32_t num, fhehip_tuning* chosen, float* ms_out);
/* Sets or reads the choice for one degree directly. */
int fhehip_tuning_set(uint32_t degree, const fhehip_tuning* t);
int fhehip_tuning_get(uint32_t degree, fhehip_tuning* t);

/* ---- Parameters ---------------------------------------------------------------
 * BfvParameters (fhe/src/bfv/parameters.rs:431-641) built from the degree
 * (a power of two, 8 .. 131072), a plaintext modulus t (2 <= t < 2^62,
 * invertible modulo every ciphertext modulus) and the ciphertext moduli (see
 * fhehip_generate_moduli for BfvParametersBuilder::set_moduli_sizes): the
 * level chain (level l holds the first L - l moduli, max_level = L - 1), the
 * plaintext context (the first moduli whose sizes reach log2(t) + 60 bits)
 * and each level's MultiplicationParameters (parameters.rs:585-612): the
 * level's moduli followed by the first ceil((bits + 60) / 62) of L + 1
 * generated 62-bit primes, an extender of factor one and a t / Q_l
 * down-scaler. */
typedef struct fhehip_parameters fhehip_parameters;
int fhehip_parameters_create(uint32_t degree, uint64_t plaintext, const uint64_t* moduli, uint32_t num_moduli,
                             fhehip_parameters** out);
void fhehip_parameters_destroy(fhehip_parameters* p);
/* BfvParameters::max_level, degree and plaintext (parameters.rs:125-154). */
int fhehip_parameters_info(const fhehip_parameters* p, uint32_t* max_level, uint32_t* degree, uint64_t* plaintext);
/* BfvParameters::context_at_level (parameters.rs:155-178): a new context
 * handle sharing the level's tables (destroy it with fhehip_context_destroy);
 * FHEHIP_ERR_BAD_LEVEL above max_level. */
int fhehip_parameters_context(const fhehip_parameters* p, uint32_t level, fhehip_context** out);
/* The plaintext context (the codec's plain_ctx), a new handle likewise. */
int fhehip_parameters_plaintext_context(const fhehip_parameters* p, fhehip_context** out);
/* Moduli of the level's multiplication context; *len is in/out. */
int fhehip_parameters_mul_basis(const fhehip_parameters* p, uint32_t level, uint64_t* out, uint32_t* len);
/* Mul<&Ciphertext> for &Ciphertext (fhe/src/bfv/ops/mod.rs:259-342) at
 * `level` with that level's MultiplicationParameters, on num pairs: d_lhs
 * [num][lhs_parts][L][N] and d_rhs [num][rhs_parts][L][N] (NTT domain) ->
 * d_out [num][lhs_parts + rhs_parts - 1][L][N], no relinearization; d_lhs
 * may equal d_rhs (squaring). At most 16 output parts. */
int fhehip_ciphertext_mul(fhehip_parameters* p, uint32_t level, const uint64_t* d_lhs, uint32_t lhs_parts,
                          const uint64_t* d_rhs, uint32_t rhs_parts, uint64_t* d_out, uint32_t num, void* stream);

/* No fhe.rs counterpart. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): fhehip_multiply with every scratch buffer carved from one fixed
 * device region of FHEHIP_NOALLOC_SCRATCH_BYTES, which each GPU the library
 * runs on holds from the library's load until the process ends; the call
 * allocates nothing. Its first call on a GPU looks the region up (make that
 * call outside graph capture). A batch whose scratch does not fit is refused
 * with FHEHIP_ERR_OOM before any work: at N = 16384 with 6 moduli and a key
 * one modulus above them an instance needs 19.1 MiB (13 per call); at
 * N = 8192 with 3 moduli and a key at their level 4.2 MiB (61 per call). Every call on one GPU shares
 * the region, and the region keeps the last call's contents (no part of it
 * has to start at zero): calls that overlap in time on one GPU race and must
 * be serialized by the caller. Results equal fhehip_multiply's. */
#define FHEHIP_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int fhehip_multiply_noalloc(fhehip_multiplicator* m, const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out,
                            uint32_t num, void* stream);

/* ---- Galois automorphisms -------------------------------------------------
 * GaloisKey (fhe/src/bfv/keys/galois_key.rs:15-60) from a key-switching key
 * built for s(x^exponent) and the exponent (odd modulo 2N). */
int fhehip_galois_create(const fhehip_ksk* ksk, uint32_t exponent, fhehip_galois** out);
void fhehip_galois_destroy(fhehip_galois* g);
/* GaloisKey::relinearize (galois_key.rs:63-124) on [num][2][L][N] NTT
 * ciphertexts; out must not alias ct. */
int fhehip_galois_relinearize(fhehip_galois* g, const uint64_t* d_ct, uint64_t* d_out, uint32_t num, void* stream);

/* ---- Evaluation keys: rotations and inner sums -------------------------------
 * EvaluationKey (fhe/src/bfv/keys/evaluation_key.rs:19-131) over a set of
 * Galois keys for one ciphertext context; it shares ownership of them. */
int fhehip_evaluation_key_create(fhehip_galois* const* keys, uint32_t count, fhehip_evaluation_key** out);
void fhehip_evaluation_key_destroy(fhehip_evaluation_key* ek);
/* EvaluationKey::rotates_columns_by (evaluation_key.rs:117-131): Galois
 * element 3^step mod 2N (:195-203); 0 < step < N/2. out must not alias ct. */
int fhehip_rotate_columns(fhehip_evaluation_key* ek, const uint64_t* d_ct, uint64_t* d_out, uint32_t step,
                          uint32_t num, void* stream);
/* EvaluationKey::rotates_rows (evaluation_key.rs:92-103): element 2N - 1. */
int fhehip_rotate_rows(fhehip_evaluation_key* ek, const uint64_t* d_ct, uint64_t* d_out, uint32_t num, void* stream);
/* EvaluationKey::computes_inner_sum (evaluation_key.rs:56-82). */
int fhehip_inner_sum(fhehip_evaluation_key* ek, const uint64_t* d_ct, uint64_t* d_out, uint32_t num, void* stream);
/* The first `log2_slots` steps of the inner sum (rotations by 1, 2, ..., with
 * an add after each), plus the row rotation when rows != 0: the sum over
 * 2^log2_slots consecutive slots of each row. No single fhe.rs call; it is
 * the computes_inner_sum loop truncated. */
int fhehip_rotate_and_sum(fhehip_evaluation_key* ek, const uint64_t* d_ct, uint64_t* d_out, uint32_t log2_slots,
                          int rows, uint32_t num, void* stream);
/* Ciphertext x RGSWCiphertext (fhe/src/bfv/rgsw_ciphertext.rs:123-160): an
 * RGSW ciphertext is the pair of KeySwitchingKeys (ksk0 from m, ksk1 from
 * m s) created with fhehip_ksk_create at the ciphertext's level (key context
 * = ciphertext context). d_out [num][2][L][N] = key_switch(c0, ksk0) +
 * key_switch(c1, ksk1) for d_ct [num][2][L][N]; FHEHIP_ERR_CTX_MISMATCH for
 * keys at other levels ("Inconsistent key switching levels"). */
int fhehip_rgsw_external_product(const fhehip_ksk* ksk0, const fhehip_ksk* ksk1, const uint64_t* d_ct,
                                 uint64_t* d_out, uint32_t num, void* stream);

/* ---- Wire format ------------------------------------------------------------------
 * The coefficients field of the Rq protobuf message (math/src/rq/convert.rs:
 * 15-87, rq.proto), i.e. the bulk of a serialized Ciphertext, KeySwitchingKey
 * or Plaintext: each limb's power-basis coefficients packed least significant
 * bit first with the bit width of q_i - 1 (transcode_to_bytes,
 * fhe-util/src/lib.rs:60-97), limbs back to back. The message framing
 * (representation, degree, seeds) stays with the caller's protobuf codec. */
/* Bytes of one polynomial of the context (Modulus::serialization_length summed). */
int fhehip_serialization_length(const fhehip_context* ctx, uint64_t* bytes_per_poly);
/* num polynomials [num][L][N] (NTT domain when ntt != 0; they are serialized
 * in power basis, as fhe.rs does) -> num payloads back to back in d_bytes. */
int fhehip_serialize(const fhehip_context* ctx, const uint64_t* d_polys, uint32_t num, int ntt, uint8_t* d_bytes,
                     void* stream);
/* The inverse (Poly::try_convert_from(&Rq), convert.rs:42-120): payloads ->
 * [num][L][N], transformed to the NTT domain when ntt != 0. */
int fhehip_deserialize(const fhehip_context* ctx, const uint8_t* d_bytes, uint32_t num, int ntt, uint64_t* d_polys,
                       void* stream);

/* ---- Protobuf messages (host bytes) ---------------------------------------
 * fhe.rs's serialized forms (prost encodings of fhe/src/proto/bfv.proto and
 * math/src/proto/rq.proto) for the objects of a parameters handle. Each Rq
 * polynomial carries its coefficients in power basis, bit-packed as
 * fhehip_serialize does. Writers fill `out` (capacity `cap`) and set *len;
 * out == NULL with cap == 0 only reports the size. Messages are read with
 * fields in any order and unknown fields skipped. These calls wait for their
 * device work; `stream` (may be NULL) orders it after the caller's.
 *
 * Ciphertext::to_bytes (fhe/src/bfv/ciphertext.rs:167-196): d_ct [parts][L][N]
 * (NTT) at `level`. With a 32-byte `seed` the last polynomial is the one drawn
 * from it (random_from_seed) and is replaced by the seed. Bit i of
 * variable_time_mask is polynomial i's allow_variable_time flag. */
int fhehip_ciphertext_to_bytes(const fhehip_parameters* p, uint32_t level, const uint64_t* d_ct, uint32_t parts,
                               const uint8_t* seed, uint32_t variable_time_mask, uint8_t* out, size_t cap,
                               size_t* len, void* stream);
/* Ciphertext::from_bytes (ciphertext.rs:198-242): *parts (polynomials,
 * including the one a seed stands for), *level and, for a seeded message,
 * *seeded = 1 and the seed (seed_out, 32 bytes, may be NULL); d_out receives
 * [parts][L][N] in the NTT domain (at most max_parts). d_out == NULL with
 * max_parts == 0 only parses. Errors as fhe.rs: fewer than two polynomials,
 * a level above max_level, a seed not of 32 bytes, a representation other
 * than NTT, a degree or payload length not of the level's context. */
int fhehip_ciphertext_from_bytes(const fhehip_parameters* p, const uint8_t* bytes, size_t len, uint64_t* d_out,
                                 uint32_t max_parts, uint32_t* parts, uint32_t* level, uint8_t* seed_out,
                                 int* seeded, void* stream);
/* KeySwitchingKey::try_convert_from (fhe/src/bfv/keys/key_switching_key.rs
 * :354-426): the key of a KeySwitchingKey message; c1 is drawn from the seed
 * (generate_c1) when the message has one, and the key keeps the seed. */
int fhehip_ksk_from_bytes(const fhehip_parameters* p, const uint8_t* bytes, size_t len, fhehip_ksk** out);
/* From<&KeySwitchingKey> (key_switching_key.rs:332-352): c0 (and c1, or the
 * seed the key came with) as NttShoup Rq messages whose allow_variable_time
 * flag is `variable_time`; levels from the parameters. */
int fhehip_ksk_to_bytes(const fhehip_parameters* p, const fhehip_ksk* k, int variable_time, uint8_t* out, size_t cap,
                        size_t* len);
/* RelinearizationKey { ksk } (fhe/src/bfv/keys/relinearization_key.rs:109-127). */
int fhehip_relinearization_key_from_bytes(const fhehip_parameters* p, const uint8_t* bytes, size_t len,
                                          fhehip_ksk** out);
int fhehip_relinearization_key_to_bytes(const fhehip_parameters* p, const fhehip_ksk* k, int variable_time,
                                        uint8_t* out, size_t cap, size_t* len);
/* GaloisKey { ksk, exponent } (fhe/src/bfv/keys/galois_key.rs:127-150). */
int fhehip_galois_key_from_bytes(const fhehip_parameters* p, const uint8_t* bytes, size_t len, fhehip_galois** out);
int fhehip_galois_key_to_bytes(const fhehip_parameters* p, const fhehip_galois* g, int variable_time, uint8_t* out,
                               size_t cap, size_t* len);
/* EvaluationKey { gk, ciphertext_level, evaluation_key_level }
 * (fhe/src/bfv/keys/evaluation_key.rs:382-431): every Galois key must have the
 * message's two levels; one key per exponent (the last one). The keys are
 * written in the handle's order (fhe.rs writes its map's order). */
int fhehip_evaluation_key_from_bytes(const fhehip_parameters* p, const uint8_t* bytes, size_t len,
                                     fhehip_evaluation_key** out, uint32_t* ciphertext_level);
int fhehip_evaluation_key_to_bytes(const fhehip_parameters* p, const fhehip_evaluation_key* ek, int variable_time,
                                   uint8_t* out, size_t cap, size_t* len);
/* BfvParameters::try_deserialize / to_bytes (fhe/src/bfv/parameters.rs
 * :644-691): degree, moduli, variance (1 .. 16; 10 for handles made by
 * fhehip_parameters_create) and the plaintext modulus (a plaintext_big of
 * more than 62 bits is FHEHIP_ERR_UNSUPPORTED). */
int fhehip_parameters_from_bytes(const uint8_t* bytes, size_t len, fhehip_parameters** out);
int fhehip_parameters_to_bytes(const fhehip_parameters* p, uint8_t* out, size_t cap, size_t* len);

/* ---- Encoding, encryption, decryption ------------------------------------------
 * The steps on either side of the evaluation path, for the BfvParameters
 * level whose ciphertext context is `ctx`: `plain_ctx` is the parameters'
 * plaintext context (their first moduli whose sizes sum to at least
 * log2(t) + 60, parameters.rs:479-494), t the plaintext modulus, psi_t the
 * primitive 2N-th root of unity modulo t of the parameters' NttOperator
 * (NULL: the library's root; it fixes the slot order of SIMD encoding) and
 * sk_coeffs the SecretKey's N signed coefficients (NULL: encoding only).
 * SIMD encoding needs t = 1 mod 2N. */
int fhehip_codec_create(const fhehip_context* ctx, const fhehip_context* plain_ctx, uint64_t t, const uint64_t* psi_t,
                        const int64_t* sk_coeffs, fhehip_codec** out);
void fhehip_codec_destroy(fhehip_codec* c);
/* Plaintext::try_encode (plaintext_vec.rs:37-101) of num x N values (< t),
 * as SIMD slots (simd != 0) or coefficients, into d_out [num][L][N] (NTT
 * domain): with scaled != 0 the polynomial encryption adds (Plaintext::to_poly,
 * plaintext.rs:71-98), else Plaintext::poly_ntt (the operand of
 * ciphertext-plaintext products and dot_product_scalar). */
int fhehip_encode(const fhehip_codec* c, const uint64_t* d_values, uint32_t num, int simd, int scaled,
                  uint64_t* d_out, void* stream);
/* SecretKey::encrypt_poly (secret_key.rs:102-137) with caller-drawn samples:
 * d_ct [num][2][L][N] = [e - a s + pt, a] for the scaled plaintexts d_pt
 * [num][L][N], the uniform NTT-domain polynomials d_a [num][L][N] (fhe.rs draws
 * them with Poly::random_from_seed) and the error coefficients d_e [num][N]
 * (Poly::small). */
int fhehip_encrypt_sk(const fhehip_codec* c, const uint64_t* d_pt, const uint64_t* d_a, const int64_t* d_e,
                      uint64_t* d_ct, uint32_t num, void* stream);
/* PublicKey::try_encrypt (public_key.rs:47-87) with caller-drawn samples:
 * d_ct [num][2][L][N] = [u pk0 + e1 + pt, u pk1 + e2] for the public key d_pk
 * [2][L][N] at the codec's level (switch it down first if it is above),
 * the scaled plaintexts d_pt [num][L][N] and the small coefficients d_u,
 * d_e1, d_e2 [num][N] (Poly::small). No secret key needed. */
int fhehip_encrypt_pk(const fhehip_codec* c, const uint64_t* d_pk, const uint64_t* d_pt, const int64_t* d_u,
                      const int64_t* d_e1, const int64_t* d_e2, uint64_t* d_ct, uint32_t num, void* stream);
/* FheDecrypter::try_decrypt (secret_key.rs:200-282) of num ciphertexts of
 * `parts` polynomials, then decoding as SIMD slots (simd != 0) or
 * coefficients: d_values [num][N]. */
int fhehip_decrypt(const fhehip_codec* c, const uint64_t* d_ct, uint32_t parts, uint32_t num, int simd,
                   uint64_t* d_values, void* stream);

/* ---- PIR server path ---------------------------------------------------------
 * EvaluationKey::expands (fhe/src/bfv/keys/evaluation_key.rs:153-193): the
 * oblivious expansion (eprint 2019/1483) of the ciphertext d_ct [2][L][N]
 * into d_out [size][2][L][N], with the key's Galois keys of exponents
 * (N >> l) + 1 for l < ceil(log2(size)). As in fhe.rs, size 0 gives one
 * ciphertext (the input). FHEHIP_ERR_INVALID_ARG when a needed key is missing
 * ("This key does not support expansion at this level") or out aliases ct. */
int fhehip_expand(fhehip_evaluation_key* ek, const uint64_t* d_ct, uint32_t size, uint64_t* d_out, void* stream);
/* fhehip_expand of `num` queries d_cts [num][2][L][N] at once, into d_out
 * [max(size, 1)][num][2][L][N] (output j of every query together), so that
 * each level relinearizes 2^l num ciphertexts as one batch and row j of the
 * expanded queries is one [num x 2]-part operand of fhehip_dot_product_columns
 * (a database pass serves all num queries). No fhe.rs counterpart: fhe.rs
 * expands one query per call. */
int fhehip_expand_batch(fhehip_evaluation_key* ek, const uint64_t* d_cts, uint32_t num, uint32_t size,
                        uint64_t* d_out, void* stream);
/* The PIR server response of fhe.rs's MulPIR example (examples/mulpir.rs:
 * 151-185) for num queries d_queries [num][2][L][N] (the query level's
 * context) over the database d_db [dim1][dim2][L][N] (Plaintext::poly_ntt
 * rows at that level): expansion to dim1 + dim2 ciphertexts, the dim2 column
 * dot products (one pass for all queries), their products with the last dim2
 * expanded ciphertexts by `mul` (a Multiplicator::new without
 * relinearization), the sum, relinearization by `rk` and Ciphertext::
 * switch_to_level by `sw` into d_responses [num][2][L_out][N]. The handles
 * must share the query context. No single fhe.rs call. */
int fhehip_pir_respond(fhehip_evaluation_key* ek, fhehip_multiplicator* mul, fhehip_ksk* rk, const fhehip_switcher* sw,
                       const uint64_t* d_db, uint32_t dim1, uint32_t dim2, const uint64_t* d_queries, uint32_t num,
                       uint64_t* d_responses, void* stream);
/* EvaluationKey::supports_expansion (evaluation_key.rs:134-147): the largest
 * level (ceil(log2(size))) fhehip_expand accepts with this key. */
int fhehip_expansion_levels(const fhehip_evaluation_key* ek, uint32_t* levels);
/* dot_product_scalar (fhe/src/bfv/ops/dot_product.rs:56-143): d_out
 * [parts][L][N] = sum over i < count of ciphertext i ([parts][L][N], NTT
 * domain, at d_ct + i * ct_stride) times plaintext i (its NTT-domain
 * polynomial Plaintext::poly_ntt [L][N] at the ciphertexts' level, at d_pt + i
 * * pt_stride); strides in 64-bit words. Exact, fully reduced, for any count.
 * count 0 -> FHEHIP_ERR_INVALID_ARG ("At least one iterator is empty"). */
int fhehip_dot_product_scalar(const fhehip_context* ctx, const uint64_t* d_ct, uint64_t ct_stride,
                              const uint64_t* d_pt, uint64_t pt_stride, uint32_t count, uint32_t parts,
                              uint64_t* d_out, void* stream);
/* The column products of a PIR server (examples/mulpir.rs:170-179,
 * sealpir.rs:165-173) in one pass over the database: for each column j <
 * columns, d_out[j] ([parts][L][N]) = dot_product_scalar of the d_ct rows
 * ([rows][parts][L][N]) with plaintexts db[0][j], db[1][j], ... of d_db
 * ([rows][columns][L][N], i.e. database.iter().skip(j).step_by(columns)). */
int fhehip_dot_product_columns(const fhehip_context* ctx, const uint64_t* d_ct, uint32_t rows, uint32_t parts,
                               const uint64_t* d_db, uint32_t columns, uint64_t* d_out, void* stream);
/* ---- Leveled operations ------------------------------------------------------
 * Ciphertext::switch_down (fhe/src/bfv/ciphertext.rs:86-97) and
 * Ciphertext::switch_to_level (:99-119): the NTT-domain polynomials of `from`
 * are taken to power basis, rounded by each dropped modulus in turn, last
 * first (Poly::switch_down, math/src/rq/mod.rs:390-449), and transformed back
 * in `to`, which must hold the first moduli of `from`
 * (FHEHIP_ERR_CTX_MISMATCH otherwise): `from` without its last modulus is
 * switch_down, a shorter prefix switch_to_level, `from` itself a copy. */
int fhehip_switcher_create(const fhehip_context* from, const fhehip_context* to, fhehip_switcher** out);
void fhehip_switcher_destroy(fhehip_switcher* s);
/* [num_polys][L_from][N] -> [num_polys][L_to][N]; out must not alias in. */
int fhehip_switch_down(const fhehip_switcher* s, const uint64_t* d_in, uint64_t* d_out, uint32_t num_polys,
                       void* stream);
/* Ciphertext / polynomial addition (fhe/src/bfv/ops/mod.rs:15-69, Poly
 * AddAssign math/src/rq/ops.rs:92-118) of `num_polys` polynomials of the
 * context: out = a + b (out may alias a or b). */
int fhehip_ciphertext_add(const fhehip_context* ctx, const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_out,
                          uint32_t num_polys, void* stream);
/* Sub for &Ciphertext (ops/mod.rs:110-164): out = a - b (out may alias a or
 * b). Adding or subtracting a plaintext (:71-107, :166-203) is this or
 * fhehip_ciphertext_add on part 0 with its fhehip_encode(scaled) polynomial. */
int fhehip_ciphertext_sub(const fhehip_context* ctx, const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_out,
                          uint32_t num_polys, void* stream);
/* Neg for &Ciphertext (ops/mod.rs:205-227): out = -a (out may alias a). */
int fhehip_ciphertext_neg(const fhehip_context* ctx, const uint64_t* d_a, uint64_t* d_out, uint32_t num_polys,
                          void* stream);
/* Mul<&Plaintext> for &Ciphertext (ops/mod.rs:229-258): each of the `parts`
 * polynomials of ciphertext i (d_ct [num][parts][L][N], NTT domain) times the
 * plaintext's poly_ntt at d_pt + i * pt_stride (words; 0: one plaintext for
 * all) into d_out [num][parts][L][N] (may alias d_ct). */
int fhehip_ciphertext_mul_plain(const fhehip_context* ctx, const uint64_t* d_ct, uint32_t parts, uint32_t num,
                                const uint64_t* d_pt, uint64_t pt_stride, uint64_t* d_out, void* stream);

/* ---- HIP graphs ---------------------------------------------------------------
 * No fhe.rs counterpart: a fixed sequence of launches captured once and
 * replayed with one call. Capture runs the sequence once eagerly on a
 * private stream (sizing its scratch), then records it. Buffers and handles
 * are fixed at capture and must outlive the graph; replays of one graph must
 * not overlap each other. */
typedef struct fhehip_graph fhehip_graph;
/* fhehip_multiply(m, lhs, rhs, out, num) as a graph. */
int fhehip_multiply_graph(fhehip_multiplicator* m, const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out,
                          uint32_t num, fhehip_graph** out);
/* fhehip_multiply followed by fhehip_rotate_and_sum(ek, .., log2_slots, rows)
 * (the request chain of the server) as a graph. */
int fhehip_multiply_sum_graph(fhehip_multiplicator* m, fhehip_evaluation_key* ek, uint32_t log2_slots, int rows,
                              const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out, uint32_t num,
                              fhehip_graph** out);
int fhehip_graph_launch(fhehip_graph* g, void* stream);
/* Waits for the device to be idle, then frees the graph and its scratch. */
void fhehip_graph_destroy(fhehip_graph* g);

/* ---- Batches of mixed parameters ----------------------------------------------
 * No single fhe.rs call: independent evaluations, each Multiplicator::multiply
 * (mul.rs:158-227) with relinearization followed by up to four
 * EvaluationKey::rotates_columns_by (evaluation_key.rs:117-131) in order.
 * Jobs may use different parameter sets (degree, moduli) and keys. */
#define FHEHIP_JOB_MAX_ROTATIONS 4
typedef struct fhehip_job {
    fhehip_multiplicator* mult; /* relinearization enabled; defines the parameter set */
    fhehip_evaluation_key* ek;  /* Galois keys of the steps; may be NULL without rotations */
    const uint64_t* d_lhs;      /* [2][L][N] NTT ciphertext */
    const uint64_t* d_rhs;      /* [2][L][N] NTT ciphertext */
    uint64_t* d_out;            /* [2][L][N]; must not overlap any job's inputs */
    uint32_t num_rotations;     /* 0 .. FHEHIP_JOB_MAX_ROTATIONS */
    uint32_t steps[FHEHIP_JOB_MAX_ROTATIONS]; /* column rotation steps */
} fhehip_job;
/* Runs all jobs. Jobs that share a Multiplicator form one batch (staged
 * through device pointer tables when their buffers are not contiguous);
 * batches run concurrently on library-owned streams ordered after `stream`,
 * and `stream` is ordered after all of them. Every job's result equals
 * fhehip_multiply then fhehip_rotate_columns per step on that job alone.
 * Every job is validated before any work is issued. */
int fhehip_submit(const fhehip_job* jobs, uint32_t n, void* stream);

/* ---- Serving ------------------------------------------------------------------
 * No fhe.rs counterpart: a request server for encrypted slot sums. A request
 * is a ct x ct multiply + relinearize (Multiplicator::multiply) followed by
 * the sum over 2^log2_slots consecutive slots (the first log2_slots steps of
 * EvaluationKey::computes_inner_sum). A server thread admits queued requests
 * in arrival order into batches of one parameter set while their device
 * footprint fits the budget, and launches each batch on a free lane (its own
 * stream) as soon as one is free, so requests arriving while others run are
 * batched and overlap on the device. Keys and contexts stay resident; request
 * buffers are device memory owned by the caller and must stay valid until
 * the request is done. */
typedef struct fhehip_server fhehip_server;
typedef struct fhehip_server_config {
    uint64_t hbm_budget_bytes; /* device bytes admitted requests may occupy at once; 0: no limit */
    uint32_t max_batch;        /* requests per batch */
    uint32_t lanes;            /* batches in flight at once (1..16) */
} fhehip_server_config;
typedef struct fhehip_server_statistics {
    uint64_t submitted, completed, failed, batches;
    uint64_t peak_bytes;       /* largest admitted footprint at once */
    uint32_t max_batch_seen;
} fhehip_server_statistics;
int fhehip_server_create(const fhehip_server_config* cfg, fhehip_server** out);
/* Stops admitting, fails what is still queued, waits for running batches. */
void fhehip_server_destroy(fhehip_server* s);
/* Registers a request kind; the server shares ownership of neither handle:
 * both must outlive it. footprint_bytes (may be NULL) receives the device
 * bytes one admitted request of this kind is charged. */
int fhehip_server_add_params(fhehip_server* s, fhehip_multiplicator* mult, fhehip_evaluation_key* ek,
                             uint32_t log2_slots, uint32_t* id, uint64_t* footprint_bytes);
/* Queues one request: lhs, rhs, out are [2][L][N] NTT ciphertexts. */
int fhehip_server_submit(fhehip_server* s, uint32_t params, const uint64_t* d_lhs, const uint64_t* d_rhs,
                         uint64_t* d_out, uint64_t* ticket);
/* state: 0 queued, 1 running, 2 done, -1 failed. latency_us: from submit to
 * the server observing completion (done requests). */
int fhehip_server_query(fhehip_server* s, uint64_t ticket, int* state, double* latency_us);
/* Waits until every request submitted so far is done or failed;
 * FHEHIP_ERR_DEVICE on timeout or when a request failed. */
int fhehip_server_drain(fhehip_server* s, double timeout_s);
int fhehip_server_stats(fhehip_server* s, fhehip_server_statistics* out);

/* BfvParametersBuilder's generate_moduli (fhe/src/bfv/parameters.rs:383-423)
 * with generate_prime (math/src/zq/primes.rs:30-59): for each bit size in
 * order, the largest prime below 2^size congruent to 1 mod 2 * degree not
 * already chosen. out receives `count` moduli. */
int fhehip_generate_moduli(const uint32_t* sizes, uint32_t count, uint32_t degree, uint64_t* out);

/* ---- Seeded values (host) -------------------------------------------------
 * The pseudo-random draws fhe.rs makes from seeds, on the host (no device
 * needed): ChaCha8Rng (rand_chacha 0.9), SHA-256 (sha2 0.10) and rand 0.9's
 * uniform integers.
 * NttOperator::new's root (math/src/ntt/native.rs:320-336): the first of 100
 * draws random_range(0..p) from ChaCha8Rng::seed_from_u64(0), raised to
 * (p - 1) / 2N, of order 2N. Contexts created without `psi` use it. */
int fhehip_default_psi(uint32_t degree, uint64_t modulus, uint64_t* psi);
/* Poly::random_from_seed (math/src/rq/mod.rs:241-257): out[l][j], l <
 * num_moduli, j < degree, uniform below moduli[l] from ChaCha8Rng seeded with
 * SHA-256(seed); NTT-domain values. seed: 32 bytes. */
int fhehip_random_from_seed(const uint64_t* moduli, uint32_t num_moduli, uint32_t degree, const uint8_t* seed,
                            uint64_t* out);
/* KeySwitchingKey::generate_c1 (fhe/src/bfv/keys/key_switching_key.rs:108-123):
 * `count` polynomials [count][num_moduli][degree], each random_from_seed of a
 * 32-byte seed filled from ChaCha8Rng(seed). */
int fhehip_generate_c1(const uint64_t* moduli, uint32_t num_moduli, uint32_t degree, const uint8_t* seed,
                       uint32_t count, uint64_t* out);

/* ---- Measurement ------------------------------------------------------------
 * No counterpart in fhe.rs. When enabled, every kernel launch is bracketed by a
 * pair of HIP events on its stream; enabling (or disabling) clears the totals.
 * The report is text, one line per kernel and degree:
 * "<kernel>/N=<degree> <launches> <total milliseconds>". `needed` receives the
 * buffer size the full report requires. */
int fhehip_profiler_enable(int on);
int fhehip_profiler_report(char* buf, size_t cap, size_t* needed);

#ifdef __cplusplus
}
#endif

#endif /* FHEHIP_H */
