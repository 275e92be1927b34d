/*
 * fhehip — MI355X-native RNS-BFV evaluation path (C ABI).
 *
 * A drop-in for the hot path of fhe.rs (crates fhe-math and fhe): negacyclic
 * NTTs, RNS scaling / base extension, key switching, ciphertext multiplication
 * with relinearization, and Galois rotations. Every entry point is extern "C",
 * takes plain pointers and sizes, never aborts, and returns a status code
 * (FHEHIP_OK == 0). All outputs are exact residues, identical to those of the
 * fhe.rs CPU path for the same inputs and NTT roots.
 *
 * Memory: every `d_*` argument is a device pointer (HBM-resident); handles own
 * their device tables (twiddles, Shoup constants, scaler constants, keys).
 * Polynomials are laid out like fhe.rs' Array2<u64>: [limb][degree] row major,
 * and batches of them back to back. Ciphertexts are [part][limb][degree].
 * `stream` is a hipStream_t (NULL = default stream); calls are asynchronous
 * with respect to the host unless stated otherwise.
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

/* Text of the last error raised on the calling thread. */
const char* fhehip_last_error(void);

/* Library version string. */
const char* fhehip_version(void);

/* ---- Contexts and NTT ---------------------------------------------------
 * fhehip_context mirrors rq::Context::new (math/src/rq/context.rs:42-93) with
 * one NttOperator::new (math/src/ntt/native.rs:35-73) per modulus. `psi` may
 * hold, per modulus, the primitive 2N-th root to use; NULL selects the
 * library's deterministic root (first g = 2,3,... with g^((p-1)/2N) of order
 * 2N). Returns FHEHIP_ERR_UNSUPPORTED when a modulus is not an NTT-friendly
 * prime for `degree` (the None case of NttOperator::new, ntt/mod.rs:19-23). */
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

#ifdef __cplusplus
}
#endif

#endif /* FHEHIP_H */
