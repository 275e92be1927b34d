// extern "C" boundary: scaler, key switching, Multiplicator, Galois keys.
#include "../../include/fhehip.h"

#include <algorithm>
#include <cstring>

#include "bfv.h"
#include "capi_internal.h"
#include "graph.h"
#include "jobs.h"
#include "prng.h"
#include "proto.h"
#include "server.h"

using namespace fhehip;

struct fhehip_scaler {
    std::unique_ptr<PolyScaler> s;
};
struct fhehip_ksk {
    std::shared_ptr<Ksk> k;
};
struct fhehip_multiplicator {
    std::unique_ptr<Multiplicator> m;
};
struct fhehip_parameters {
    std::unique_ptr<Parameters> p;
};
struct fhehip_galois {
    std::shared_ptr<Galois> g;
};
struct fhehip_evaluation_key {
    std::unique_ptr<EvaluationKey> ek;
};
struct fhehip_server {
    std::unique_ptr<Server> s;
};
struct fhehip_graph {
    std::unique_ptr<Graph> g;
};
struct fhehip_switcher {
    std::unique_ptr<Switcher> s;
};
struct fhehip_codec {
    std::unique_ptr<Codec> c;
};

static BigUint factor_num(const fhehip_scaling_factor* f) {
    if (!f || !f->num) return BigUint(1);
    return BigUint::from_words32(f->num, f->num_len);
}
static BigUint factor_den(const fhehip_scaling_factor* f) {
    if (!f || !f->den) return BigUint(1);
    return BigUint::from_words32(f->den, f->den_len);
}

static hipStream_t S(void* s) { return (hipStream_t)s; }

extern "C" {

int fhehip_scaler_create(const fhehip_context* from, const fhehip_context* to, const fhehip_scaling_factor* factor,
                         fhehip_scaler** out) {
    FHEHIP_API_BEGIN
    if (!from || !to || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    BigUint den = factor_den(factor);
    if (den.is_zero()) return fail(FHEHIP_ERR_INVALID_ARG, "zero denominator");
    std::unique_ptr<PolyScaler> s(new PolyScaler(from->c, to->c, factor_num(factor), den));
    *out = new fhehip_scaler{std::move(s)};
    FHEHIP_API_END
}

void fhehip_scaler_destroy(fhehip_scaler* s) { delete s; }

int fhehip_scaler_scale(fhehip_scaler* s, const uint64_t* d_in, uint64_t* d_out, uint32_t num_polys, int ntt_repr,
                        void* stream) {
    FHEHIP_API_BEGIN
    if (!s || ((!d_in || !d_out) && num_polys)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    s->s->scale(d_in, d_out, num_polys, ntt_repr != 0, S(stream));
    FHEHIP_API_END
}

int fhehip_ksk_create(const fhehip_context* ct_ctx, const fhehip_context* key_ctx, const uint64_t* c0,
                      const uint64_t* c1, uint32_t digits, fhehip_ksk** out) {
    FHEHIP_API_BEGIN
    if (!ct_ctx || !key_ctx || !c0 || !c1 || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    auto k = std::make_shared<Ksk>(ct_ctx->c, key_ctx->c, c0, c1, digits);
    *out = new fhehip_ksk{k};
    FHEHIP_API_END
}

void fhehip_ksk_destroy(fhehip_ksk* k) { delete k; }

int fhehip_key_switch(const fhehip_ksk* k, const uint64_t* d_p, uint64_t* d_out, uint32_t num, void* stream) {
    FHEHIP_API_BEGIN
    if (!k || ((!d_p || !d_out) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    k->k->key_switch(d_p, d_out, num, S(stream));
    FHEHIP_API_END
}

int fhehip_serialization_length(const fhehip_context* ctx, uint64_t* bytes_per_poly) {
    FHEHIP_API_BEGIN
    if (!ctx || !bytes_per_poly) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *bytes_per_poly = serialization_length(*ctx->c);
    FHEHIP_API_END
}

int fhehip_serialize(const fhehip_context* ctx, const uint64_t* d_polys, uint32_t num, int ntt, uint8_t* d_bytes,
                     void* stream) {
    FHEHIP_API_BEGIN
    if (!ctx || ((!d_polys || !d_bytes) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    serialize_polys(*ctx->c, d_polys, num, ntt != 0, d_bytes, S(stream));
    FHEHIP_API_END
}

int fhehip_deserialize(const fhehip_context* ctx, const uint8_t* d_bytes, uint32_t num, int ntt, uint64_t* d_polys,
                       void* stream) {
    FHEHIP_API_BEGIN
    if (!ctx || ((!d_polys || !d_bytes) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    deserialize_polys(*ctx->c, d_bytes, num, ntt != 0, d_polys, S(stream));
    FHEHIP_API_END
}

int fhehip_rgsw_external_product(const fhehip_ksk* ksk0, const fhehip_ksk* ksk1, const uint64_t* d_ct,
                                 uint64_t* d_out, uint32_t num, void* stream) {
    FHEHIP_API_BEGIN
    if (!ksk0 || !ksk1 || ((!d_ct || !d_out) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (d_ct == d_out && num) return fail(FHEHIP_ERR_INVALID_ARG, "out must not alias ct");
    rgsw_external_product(*ksk0->k, *ksk1->k, d_ct, d_out, num, S(stream));
    FHEHIP_API_END
}

int fhehip_relinearize(fhehip_ksk* rk, const uint64_t* d_ct3, uint64_t* d_out, uint32_t num, void* stream) {
    FHEHIP_API_BEGIN
    if (!rk || ((!d_ct3 || !d_out) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    rk->k->relinearize(d_ct3, d_out, num, S(stream));
    FHEHIP_API_END
}

int fhehip_multiplicator_default(const fhehip_ksk* rk, uint64_t plaintext_modulus, fhehip_multiplicator** out) {
    FHEHIP_API_BEGIN
    if (!rk || !out || plaintext_modulus < 2) return fail(FHEHIP_ERR_INVALID_ARG, "bad argument");
    *out = nullptr;
    auto m = Multiplicator::make_default(rk->k, plaintext_modulus);
    *out = new fhehip_multiplicator{std::move(m)};
    FHEHIP_API_END
}

int fhehip_multiplicator_create(const fhehip_context* base_ctx, const uint64_t* extended_basis, uint32_t ext_len,
                                const fhehip_scaling_factor* lhs, const fhehip_scaling_factor* rhs,
                                const fhehip_scaling_factor* post, fhehip_multiplicator** out) {
    FHEHIP_API_BEGIN
    if (!base_ctx || !extended_basis || !ext_len || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    std::vector<u64> ext(extended_basis, extended_basis + ext_len);
    std::unique_ptr<Multiplicator> m(new Multiplicator(base_ctx->c, ext, factor_num(lhs), factor_den(lhs),
                                                       factor_num(rhs), factor_den(rhs), factor_num(post),
                                                       factor_den(post)));
    *out = new fhehip_multiplicator{std::move(m)};
    FHEHIP_API_END
}

int fhehip_multiplicator_enable_relinearization(fhehip_multiplicator* m, const fhehip_ksk* rk) {
    FHEHIP_API_BEGIN
    if (!m || !rk) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (rk->k->ct_ctx->moduli != m->m->base->moduli || rk->k->ct_ctx->N != m->m->base->N)
        return fail(FHEHIP_ERR_CTX_MISMATCH, "Invalid relinearization key context");
    m->m->rk = rk->k;
    FHEHIP_API_END
}

int fhehip_multiplicator_enable_mod_switching(fhehip_multiplicator* m, const fhehip_context* next) {
    FHEHIP_API_BEGIN
    if (!m || !next) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (m->m->base->L < 2)
        return fail(FHEHIP_ERR_INVALID_ARG, "Cannot modulo switch as this is already the last level");
    if (next->c->L + 1 != m->m->base->L)
        return fail(FHEHIP_ERR_CTX_MISMATCH, "the next context must drop exactly the last modulus");
    m->m->mod_switch.reset(new Switcher(m->m->base, next->c));
    FHEHIP_API_END
}

void fhehip_multiplicator_destroy(fhehip_multiplicator* m) { delete m; }

int fhehip_codec_create(const fhehip_context* ctx, const fhehip_context* plain_ctx, uint64_t t, const uint64_t* psi_t,
                        const int64_t* sk_coeffs, fhehip_codec** out) {
    FHEHIP_API_BEGIN
    if (!ctx || !plain_ctx || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    std::unique_ptr<Codec> c(new Codec(ctx->c, plain_ctx->c, t, psi_t, sk_coeffs));
    *out = new fhehip_codec{std::move(c)};
    FHEHIP_API_END
}

void fhehip_codec_destroy(fhehip_codec* c) { delete c; }

int fhehip_encode(const fhehip_codec* c, const uint64_t* d_values, uint32_t num, int simd, int scaled,
                  uint64_t* d_out, void* stream) {
    FHEHIP_API_BEGIN
    if (!c || ((!d_values || !d_out) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    c->c->encode(d_values, d_out, num, simd != 0, scaled != 0, S(stream));
    FHEHIP_API_END
}

int fhehip_encrypt_sk(const fhehip_codec* c, const uint64_t* d_pt, const uint64_t* d_a, const int64_t* d_e,
                      uint64_t* d_ct, uint32_t num, void* stream) {
    FHEHIP_API_BEGIN
    if (!c || ((!d_pt || !d_a || !d_e || !d_ct) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    c->c->encrypt(d_pt, d_a, d_e, d_ct, num, S(stream));
    FHEHIP_API_END
}

int fhehip_encrypt_pk(const fhehip_codec* c, const uint64_t* d_pk, const uint64_t* d_pt, const int64_t* d_u,
                      const int64_t* d_e1, const int64_t* d_e2, uint64_t* d_ct, uint32_t num, void* stream) {
    FHEHIP_API_BEGIN
    if (!c || ((!d_pk || !d_pt || !d_u || !d_e1 || !d_e2 || !d_ct) && num))
        return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    c->c->encrypt_pk(d_pk, d_pt, d_u, d_e1, d_e2, d_ct, num, S(stream));
    FHEHIP_API_END
}

int fhehip_decrypt(const fhehip_codec* c, const uint64_t* d_ct, uint32_t parts, uint32_t num, int simd,
                   uint64_t* d_values, void* stream) {
    FHEHIP_API_BEGIN
    if (!c || ((!d_ct || !d_values) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    c->c->decrypt(d_ct, parts, d_values, num, simd != 0, S(stream));
    FHEHIP_API_END
}

int fhehip_switcher_create(const fhehip_context* from, const fhehip_context* to, fhehip_switcher** out) {
    FHEHIP_API_BEGIN
    if (!from || !to || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    std::unique_ptr<Switcher> sw(new Switcher(from->c, to->c));
    *out = new fhehip_switcher{std::move(sw)};
    FHEHIP_API_END
}

void fhehip_switcher_destroy(fhehip_switcher* s) { delete s; }

int fhehip_switch_down(const fhehip_switcher* s, const uint64_t* d_in, uint64_t* d_out, uint32_t num_polys,
                       void* stream) {
    FHEHIP_API_BEGIN
    if (!s || ((!d_in || !d_out) && num_polys)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (num_polys && d_in == d_out) return fail(FHEHIP_ERR_INVALID_ARG, "the output must not alias the input");
    s->s->switch_down(d_in, d_out, num_polys, S(stream));
    FHEHIP_API_END
}

int fhehip_multiplicator_extended_basis(const fhehip_multiplicator* m, uint64_t* out, uint32_t* len) {
    FHEHIP_API_BEGIN
    if (!m || !len) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    const auto& mod = m->m->mul->moduli;
    if (out) {
        if (*len < mod.size()) return fail(FHEHIP_ERR_INVALID_ARG, "output too small");
        std::memcpy(out, mod.data(), mod.size() * 8);
    }
    *len = (uint32_t)mod.size();
    FHEHIP_API_END
}

int fhehip_multiply(fhehip_multiplicator* m, const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out,
                    uint32_t num, void* stream) {
    FHEHIP_API_BEGIN
    if (!m || ((!d_lhs || !d_rhs || !d_out) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    m->m->multiply(d_lhs, d_rhs, d_out, num, S(stream));
    FHEHIP_API_END
}

static FusedTuning from_c(const fhehip_tuning& t) {
    FusedTuning f;
    f.xcd = (int)t.xcd_order;
    f.ks_mac_inst = (int)t.ks_mac_instances;
    f.re_tensor = (int)t.tensor_log2_values;
    f.re_lift = (int)t.lift_log2_values;
    f.re_finish = (int)t.finish_log2_values;
    f.halves = (int)t.halves;
    f.scaler_mfma = (int)t.scaler_mfma;
    return f;
}

static fhehip_tuning to_c(const FusedTuning& f) {
    return fhehip_tuning{(uint32_t)f.xcd,     (uint32_t)f.ks_mac_inst, (uint32_t)f.re_tensor, (uint32_t)f.re_lift,
                         (uint32_t)f.re_finish, (uint32_t)f.halves,     (uint32_t)f.scaler_mfma};
}

static int degree_log2(uint32_t degree) {
    if (degree < 8 || degree > 32768 || (degree & (degree - 1))) throw Error(-1, "degree must be 8 .. 32768, a power of two");
    return 31 - __builtin_clz(degree);
}

int fhehip_multiply_autotune(fhehip_multiplicator* m, uint32_t num, fhehip_tuning* chosen, float* ms_out) {
    FHEHIP_API_BEGIN
    if (!m) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    float ms[19] = {};
    const FusedTuning t = autotune_multiply(*m->m, num, ms);
    if (chosen) *chosen = to_c(t);
    if (ms_out) std::memcpy(ms_out, ms, sizeof ms);
    FHEHIP_API_END
}

int fhehip_tuning_set(uint32_t degree, const fhehip_tuning* t) {
    FHEHIP_API_BEGIN
    if (!t) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    set_fused_tuning(degree_log2(degree), from_c(*t));
    FHEHIP_API_END
}

int fhehip_tuning_get(uint32_t degree, fhehip_tuning* t) {
    FHEHIP_API_BEGIN
    if (!t) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *t = to_c(fused_tuning(degree_log2(degree)));
    FHEHIP_API_END
}

int fhehip_parameters_create(uint32_t degree, uint64_t plaintext, const uint64_t* moduli, uint32_t num_moduli,
                             fhehip_parameters** out) {
    FHEHIP_API_BEGIN
    if (!moduli || !out || !num_moduli) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    std::unique_ptr<Parameters> p(new Parameters(degree, plaintext, std::vector<u64>(moduli, moduli + num_moduli)));
    *out = new fhehip_parameters{std::move(p)};
    FHEHIP_API_END
}

void fhehip_parameters_destroy(fhehip_parameters* p) { delete p; }

int fhehip_parameters_info(const fhehip_parameters* p, uint32_t* max_level, uint32_t* degree, uint64_t* plaintext) {
    FHEHIP_API_BEGIN
    if (!p) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (max_level) *max_level = (uint32_t)p->p->levels.size() - 1;
    if (degree) *degree = p->p->N;
    if (plaintext) *plaintext = p->p->t;
    FHEHIP_API_END
}

int fhehip_parameters_context(const fhehip_parameters* p, uint32_t level, fhehip_context** out) {
    FHEHIP_API_BEGIN
    if (!p || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    p->p->at(level);
    *out = new fhehip_context{p->p->levels[level]};
    FHEHIP_API_END
}

int fhehip_parameters_plaintext_context(const fhehip_parameters* p, fhehip_context** out) {
    FHEHIP_API_BEGIN
    if (!p || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = new fhehip_context{p->p->plain};
    FHEHIP_API_END
}

int fhehip_parameters_mul_basis(const fhehip_parameters* p, uint32_t level, uint64_t* out, uint32_t* len) {
    FHEHIP_API_BEGIN
    if (!p || !len) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    const auto& mod = p->p->at(level).mul->moduli;
    if (out) {
        if (*len < mod.size()) return fail(FHEHIP_ERR_INVALID_ARG, "output too small");
        std::memcpy(out, mod.data(), mod.size() * 8);
    }
    *len = (uint32_t)mod.size();
    FHEHIP_API_END
}

int fhehip_ciphertext_mul(fhehip_parameters* p, uint32_t level, const uint64_t* d_lhs, uint32_t lhs_parts,
                          const uint64_t* d_rhs, uint32_t rhs_parts, uint64_t* d_out, uint32_t num, void* stream) {
    FHEHIP_API_BEGIN
    if (!p || ((!d_lhs || !d_rhs || !d_out) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    p->p->ct_mul(level, d_lhs, lhs_parts, d_rhs, rhs_parts, d_out, num, S(stream));
    FHEHIP_API_END
}

int fhehip_multiply_noalloc(fhehip_multiplicator* m, const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out,
                            uint32_t num, void* stream) {
    FHEHIP_API_BEGIN
    if (!m || ((!d_lhs || !d_rhs || !d_out) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    multiply_noalloc(*m->m, d_lhs, d_rhs, d_out, num, S(stream));
    FHEHIP_API_END
}

int fhehip_galois_create(const fhehip_ksk* ksk, uint32_t exponent, fhehip_galois** out) {
    FHEHIP_API_BEGIN
    if (!ksk || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    auto g = std::make_shared<Galois>(ksk->k, exponent);
    *out = new fhehip_galois{g};
    FHEHIP_API_END
}

void fhehip_galois_destroy(fhehip_galois* g) { delete g; }

int fhehip_galois_relinearize(fhehip_galois* g, const uint64_t* d_ct, uint64_t* d_out, uint32_t num, void* stream) {
    FHEHIP_API_BEGIN
    if (!g || ((!d_ct || !d_out) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (d_ct == d_out && num) return fail(FHEHIP_ERR_INVALID_ARG, "in-place Galois relinearization is not supported");
    g->g->apply(d_ct, d_out, num, S(stream));
    FHEHIP_API_END
}

int fhehip_evaluation_key_create(fhehip_galois* const* keys, uint32_t count, fhehip_evaluation_key** out) {
    FHEHIP_API_BEGIN
    if (!keys || !count || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    std::unique_ptr<EvaluationKey> ek(new EvaluationKey);
    for (uint32_t i = 0; i < count; i++) {
        if (!keys[i]) return fail(FHEHIP_ERR_INVALID_ARG, "null Galois key");
        const auto& g = keys[i]->g;
        if (i && (g->ksk->ct_ctx->moduli != ek->keys[0]->ksk->ct_ctx->moduli ||
                  g->ksk->ct_ctx->N != ek->keys[0]->ksk->ct_ctx->N))
            return fail(FHEHIP_ERR_CTX_MISMATCH, "Galois keys for different ciphertext contexts");
        ek->keys.push_back(g);
    }
    ek->degree = ek->keys[0]->ksk->ct_ctx->N;
    ek->prepare_expansion();
    *out = new fhehip_evaluation_key{std::move(ek)};
    FHEHIP_API_END
}

void fhehip_evaluation_key_destroy(fhehip_evaluation_key* ek) { delete ek; }

int fhehip_rotate_columns(fhehip_evaluation_key* ek, const uint64_t* d_ct, uint64_t* d_out, uint32_t step,
                          uint32_t num, void* stream) {
    FHEHIP_API_BEGIN
    if (!ek || ((!d_ct || !d_out) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (d_ct == d_out && num) return fail(FHEHIP_ERR_INVALID_ARG, "out must not alias ct");
    if (step == 0 || step >= ek->ek->degree / 2) return fail(FHEHIP_ERR_INVALID_ARG, "rotation step out of range");
    ek->ek->rotate(ek->ek->column_exponent(step), d_ct, d_out, num, S(stream));
    FHEHIP_API_END
}

int fhehip_expand(fhehip_evaluation_key* ek, const uint64_t* d_ct, uint32_t size, uint64_t* d_out, void* stream) {
    FHEHIP_API_BEGIN
    if (!ek || !d_ct || !d_out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (d_ct == d_out) return fail(FHEHIP_ERR_INVALID_ARG, "out must not alias ct");
    ek->ek->expand(d_ct, size, 1, d_out, S(stream));
    FHEHIP_API_END
}

int fhehip_expand_batch(fhehip_evaluation_key* ek, const uint64_t* d_cts, uint32_t num, uint32_t size,
                        uint64_t* d_out, void* stream) {
    FHEHIP_API_BEGIN
    if (!ek || ((!d_cts || !d_out) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (d_cts == d_out && num) return fail(FHEHIP_ERR_INVALID_ARG, "out must not alias ct");
    ek->ek->expand(d_cts, size, num, d_out, S(stream));
    FHEHIP_API_END
}

int fhehip_pir_respond(fhehip_evaluation_key* ek, fhehip_multiplicator* mul, fhehip_ksk* rk, const fhehip_switcher* sw,
                       const uint64_t* d_db, uint32_t dim1, uint32_t dim2, const uint64_t* d_queries, uint32_t num,
                       uint64_t* d_responses, void* stream) {
    FHEHIP_API_BEGIN
    if (!ek || !mul || !rk || !sw || ((!d_db || !d_queries || !d_responses) && num))
        return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    pir_respond(*ek->ek, *mul->m, *rk->k, *sw->s, d_db, dim1, dim2, d_queries, num, d_responses, S(stream));
    FHEHIP_API_END
}

int fhehip_expansion_levels(const fhehip_evaluation_key* ek, uint32_t* levels) {
    FHEHIP_API_BEGIN
    if (!ek || !levels) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *levels = ek->ek->expansion_levels;
    FHEHIP_API_END
}

int fhehip_dot_product_scalar(const fhehip_context* ctx, const uint64_t* d_ct, uint64_t ct_stride,
                              const uint64_t* d_pt, uint64_t pt_stride, uint32_t count, uint32_t parts,
                              uint64_t* d_out, void* stream) {
    FHEHIP_API_BEGIN
    if (!ctx || ((!d_ct || !d_pt || !d_out) && count)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    dot_product_scalar(*ctx->c, d_ct, ct_stride, d_pt, pt_stride, count, parts, d_out, S(stream));
    FHEHIP_API_END
}

int fhehip_dot_product_columns(const fhehip_context* ctx, const uint64_t* d_ct, uint32_t rows, uint32_t parts,
                               const uint64_t* d_db, uint32_t columns, uint64_t* d_out, void* stream) {
    FHEHIP_API_BEGIN
    if (!ctx || ((!d_ct || !d_db || !d_out) && rows && columns)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    dot_product_columns(*ctx->c, d_ct, rows, parts, d_db, columns, d_out, S(stream));
    FHEHIP_API_END
}

int fhehip_rotate_rows(fhehip_evaluation_key* ek, const uint64_t* d_ct, uint64_t* d_out, uint32_t num, void* stream) {
    FHEHIP_API_BEGIN
    if (!ek || ((!d_ct || !d_out) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (d_ct == d_out && num) return fail(FHEHIP_ERR_INVALID_ARG, "out must not alias ct");
    ek->ek->rotate(2 * ek->ek->degree - 1, d_ct, d_out, num, S(stream));
    FHEHIP_API_END
}

int fhehip_inner_sum(fhehip_evaluation_key* ek, const uint64_t* d_ct, uint64_t* d_out, uint32_t num, void* stream) {
    FHEHIP_API_BEGIN
    if (!ek || ((!d_ct || !d_out) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    std::vector<u32> ex;
    for (u32 i = 1; i < ek->ek->degree / 2; i *= 2) ex.push_back(ek->ek->column_exponent(i));
    ex.push_back(2 * ek->ek->degree - 1);
    ek->ek->rotate_and_sum(ex, d_ct, d_out, num, S(stream));
    FHEHIP_API_END
}

int fhehip_rotate_and_sum(fhehip_evaluation_key* ek, const uint64_t* d_ct, uint64_t* d_out, uint32_t log2_slots,
                          int rows, uint32_t num, void* stream) {
    FHEHIP_API_BEGIN
    if (!ek || ((!d_ct || !d_out) && num)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if ((1ull << log2_slots) > ek->ek->degree / 2) return fail(FHEHIP_ERR_INVALID_ARG, "log2_slots too large");
    std::vector<u32> ex;
    for (u32 k = 0; k < log2_slots; k++) ex.push_back(ek->ek->column_exponent(1u << k));
    if (rows) ex.push_back(2 * ek->ek->degree - 1);
    ek->ek->rotate_and_sum(ex, d_ct, d_out, num, S(stream));
    FHEHIP_API_END
}

int fhehip_ciphertext_add(const fhehip_context* ctx, const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_out,
                          uint32_t num_polys, void* stream) {
    FHEHIP_API_BEGIN
    if (!ctx || ((!d_a || !d_b || !d_out) && num_polys)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    const auto& c = *ctx->c;
    launch_add_rows(d_a, d_b, d_out, c.mod.as<ModInfo>(), c.L, c.logn, (size_t)num_polys * c.L, S(stream));
    FHEHIP_API_END
}

int fhehip_ciphertext_sub(const fhehip_context* ctx, const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_out,
                          uint32_t num_polys, void* stream) {
    FHEHIP_API_BEGIN
    if (!ctx || ((!d_a || !d_b || !d_out) && num_polys)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    const auto& c = *ctx->c;
    launch_rows_sub_neg(d_a, d_b, d_out, c.mod.as<ModInfo>(), c.L, c.logn, (size_t)num_polys * c.L, false,
                        S(stream));
    FHEHIP_API_END
}

int fhehip_ciphertext_neg(const fhehip_context* ctx, const uint64_t* d_a, uint64_t* d_out, uint32_t num_polys,
                          void* stream) {
    FHEHIP_API_BEGIN
    if (!ctx || ((!d_a || !d_out) && num_polys)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    const auto& c = *ctx->c;
    launch_rows_sub_neg(d_a, d_a, d_out, c.mod.as<ModInfo>(), c.L, c.logn, (size_t)num_polys * c.L, true, S(stream));
    FHEHIP_API_END
}

int fhehip_ciphertext_mul_plain(const fhehip_context* ctx, const uint64_t* d_ct, uint32_t parts, uint32_t num,
                                const uint64_t* d_pt, uint64_t pt_stride, uint64_t* d_out, void* stream) {
    FHEHIP_API_BEGIN
    if (!ctx || ((!d_ct || !d_pt || !d_out) && num && parts)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (num > 65535) return fail(FHEHIP_ERR_INVALID_ARG, "at most 65535 ciphertexts per call");
    const auto& c = *ctx->c;
    if ((size_t)parts * c.L > 65535) return fail(FHEHIP_ERR_INVALID_ARG, "too many parts");
    launch_mul_plain(d_ct, parts, d_pt, pt_stride, d_out, c.mod.as<ModInfo>(), c.L, c.logn, num, S(stream));
    FHEHIP_API_END
}

int fhehip_multiply_graph(fhehip_multiplicator* m, const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out,
                          uint32_t num, fhehip_graph** out) {
    FHEHIP_API_BEGIN
    if (!m || !out || !d_lhs || !d_rhs || !d_out || !num) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    Multiplicator* mm = m->m.get();
    auto g = capture_graph(std::unique_ptr<Graph>(new Graph),
                           [&](hipStream_t s) { mm->multiply(d_lhs, d_rhs, d_out, num, s); });
    *out = new fhehip_graph{std::move(g)};
    FHEHIP_API_END
}

int fhehip_multiply_sum_graph(fhehip_multiplicator* m, fhehip_evaluation_key* ek, uint32_t log2_slots, int rows,
                              const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out, uint32_t num,
                              fhehip_graph** out) {
    FHEHIP_API_BEGIN
    if (!m || !ek || !out || !d_lhs || !d_rhs || !d_out || !num) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (!m->m->rk) return fail(FHEHIP_ERR_INVALID_ARG, "the multiplicator has no relinearization key");
    if (m->m->mod_switch) return fail(FHEHIP_ERR_UNSUPPORTED, "modulus switching in a multiply-and-sum graph");
    EvaluationKey* e = ek->ek.get();
    const auto& b = *m->m->base;
    if (e->degree != b.N || e->keys[0]->ksk->ct_ctx->moduli != b.moduli)
        return fail(FHEHIP_ERR_CTX_MISMATCH, "evaluation key and multiplicator use different ciphertext contexts");
    if ((1ull << log2_slots) > b.N / 2) return fail(FHEHIP_ERR_INVALID_ARG, "log2_slots too large");
    std::vector<u32> ex;
    for (u32 k = 0; k < log2_slots; k++) ex.push_back(e->column_exponent(1u << k));
    if (rows) ex.push_back(2 * e->degree - 1);
    for (u32 x : ex)
        if (!e->find(x)) return fail(FHEHIP_ERR_INVALID_ARG, "the evaluation key lacks a rotation of the sum");
    std::unique_ptr<Graph> g(new Graph);
    g->tmp.reset((size_t)num * 2 * b.L * b.N * 8);
    u64* prod = g->tmp.as<u64>();
    Multiplicator* mm = m->m.get();
    g = capture_graph(std::move(g), [&](hipStream_t s) {
        mm->multiply(d_lhs, d_rhs, prod, num, s);
        e->rotate_and_sum(ex, prod, d_out, num, s);
    });
    *out = new fhehip_graph{std::move(g)};
    FHEHIP_API_END
}

int fhehip_graph_launch(fhehip_graph* g, void* stream) {
    FHEHIP_API_BEGIN
    if (!g) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    g->g->launch(S(stream));
    FHEHIP_API_END
}

void fhehip_graph_destroy(fhehip_graph* g) { delete g; }

int fhehip_submit(const fhehip_job* jobs, uint32_t n, void* stream) {
    FHEHIP_API_BEGIN
    if (!jobs && n) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    std::vector<Job> js(n);
    for (uint32_t k = 0; k < n; k++) {
        const fhehip_job& j = jobs[k];
        if (!j.mult) return fail(FHEHIP_ERR_INVALID_ARG, "null multiplicator in a job");
        if (j.mult->m->mod_switch) return fail(FHEHIP_ERR_UNSUPPORTED, "modulus switching in a job");
        js[k].mult = j.mult->m.get();
        js[k].ek = j.ek ? j.ek->ek.get() : nullptr;
        js[k].lhs = j.d_lhs;
        js[k].rhs = j.d_rhs;
        js[k].out = j.d_out;
        js[k].num_rotations = j.num_rotations;
        std::memcpy(js[k].steps, j.steps, sizeof(js[k].steps));
    }
    JobRunner::instance().submit(js.data(), n, S(stream));
    FHEHIP_API_END
}

int fhehip_server_create(const fhehip_server_config* cfg, fhehip_server** out) {
    FHEHIP_API_BEGIN
    if (!cfg || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    ServerConfig c;
    c.hbm_budget = cfg->hbm_budget_bytes;
    c.max_batch = cfg->max_batch;
    c.lanes = cfg->lanes;
    *out = new fhehip_server{std::unique_ptr<Server>(new Server(c))};
    FHEHIP_API_END
}

void fhehip_server_destroy(fhehip_server* s) { delete s; }

int fhehip_server_add_params(fhehip_server* s, fhehip_multiplicator* mult, fhehip_evaluation_key* ek,
                             uint32_t log2_slots, uint32_t* id, uint64_t* footprint_bytes) {
    FHEHIP_API_BEGIN
    if (!s || !mult || !ek || !id) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (mult->m->mod_switch) return fail(FHEHIP_ERR_UNSUPPORTED, "modulus switching in a served parameter set");
    *id = s->s->add_params(mult->m.get(), ek->ek.get(), log2_slots);
    if (footprint_bytes) *footprint_bytes = s->s->params(*id).footprint;
    FHEHIP_API_END
}

int fhehip_server_submit(fhehip_server* s, uint32_t params, const uint64_t* d_lhs, const uint64_t* d_rhs,
                         uint64_t* d_out, uint64_t* ticket) {
    FHEHIP_API_BEGIN
    if (!s || !ticket) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *ticket = s->s->submit(params, d_lhs, d_rhs, d_out);
    FHEHIP_API_END
}

int fhehip_server_query(fhehip_server* s, uint64_t ticket, int* state, double* latency_us) {
    FHEHIP_API_BEGIN
    if (!s || !state || !latency_us) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    s->s->query(ticket, *state, *latency_us);
    FHEHIP_API_END
}

int fhehip_server_drain(fhehip_server* s, double timeout_s) {
    FHEHIP_API_BEGIN
    if (!s) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (!s->s->drain(timeout_s)) return fail(FHEHIP_ERR_DEVICE, "server drain timed out");
    const ServerStats st = s->s->stats();
    if (st.failed) return fail(FHEHIP_ERR_DEVICE, "requests failed: " + s->s->error());
    FHEHIP_API_END
}

int fhehip_server_stats(fhehip_server* s, fhehip_server_statistics* out) {
    FHEHIP_API_BEGIN
    if (!s || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    const ServerStats st = s->s->stats();
    out->submitted = st.submitted;
    out->completed = st.completed;
    out->failed = st.failed;
    out->batches = st.batches;
    out->peak_bytes = st.peak_bytes;
    out->max_batch_seen = st.max_batch_seen;
    FHEHIP_API_END
}

int fhehip_generate_moduli(const uint32_t* sizes, uint32_t count, uint32_t degree, uint64_t* out) {
    FHEHIP_API_BEGIN
    if ((!sizes || !out) && count) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (degree < 8 || (degree & (degree - 1))) return fail(FHEHIP_ERR_INVALID_ARG, "degree is not a power of two >= 8");
    std::vector<u64> chosen;
    for (uint32_t k = 0; k < count; k++) {
        const u32 sz = sizes[k];
        if (sz < 10 || sz > 62) return fail(FHEHIP_ERR_INVALID_ARG, "modulus size outside [10, 62] bits");
        u64 ub = (u64)1 << sz;
        for (;;) {
            const u64 p = generate_prime64(sz, 2 * (u64)degree, ub);
            if (!p) return fail(FHEHIP_ERR_UNSUPPORTED, "not enough primes of the requested size");
            if (std::find(chosen.begin(), chosen.end(), p) == chosen.end()) {
                chosen.push_back(p);
                break;
            }
            ub = p;
        }
    }
    std::copy(chosen.begin(), chosen.end(), out);
    FHEHIP_API_END
}

}  // extern "C"

// ---- Protobuf messages -------------------------------------------------------
namespace {

// Level of context c in the parameters' chain (its moduli are a prefix).
u32 level_of(const Parameters& par, const DeviceContext& c) {
    const u32 L0 = (u32)par.moduli.size();
    if (c.N != par.N || c.L > L0 || !std::equal(c.moduli.begin(), c.moduli.end(), par.moduli.begin()))
        throw Error(FHEHIP_ERR_CTX_MISMATCH, "the context is not a level of these parameters");
    return L0 - c.L;
}

const std::shared_ptr<DeviceContext>& level_ctx(const Parameters& par, u32 level) {
    if (level >= par.levels.size()) throw Error(FHEHIP_ERR_BAD_LEVEL, "level above the parameters' max_level");
    return par.levels[level];
}

int copy_message(const std::vector<uint8_t>& msg, uint8_t* out, size_t cap, size_t* len) {
    if (!len) return fail(FHEHIP_ERR_INVALID_ARG, "null length");
    *len = msg.size();
    if (!out && !cap) return FHEHIP_OK;  // size query
    if (!out || cap < msg.size()) return fail(FHEHIP_ERR_INVALID_ARG, "output buffer too small (*len holds the size)");
    std::memcpy(out, msg.data(), msg.size());
    return FHEHIP_OK;
}

// Payloads of `num` NTT-domain polynomials of c (power basis, bit-packed),
// on the host; waits for the stream.
std::vector<uint8_t> payloads(const DeviceContext& c, const u64* d_polys, u32 num, hipStream_t s) {
    const size_t per = serialization_length(c);
    std::vector<uint8_t> h(per * num);
    if (!num) return h;
    DevBuf d(h.size());
    serialize_polys(c, d_polys, num, true, d.as<uint8_t>(), s);
    HIPCHK(hipMemcpyAsync(h.data(), d.ptr, h.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return h;
}

// Poly::<R>::from_bytes (math/src/rq/serialize.rs:23-26, convert.rs:42-120)
// of each message: representation `repr`, the context's degree and payload
// length; into d_out [num][L][N] in the NTT domain. Waits for the stream.
void rq_to_device(const DeviceContext& c, const std::vector<pb::Span>& msgs, uint32_t repr, u64* d_out, hipStream_t s) {
    if (msgs.empty()) return;
    const size_t per = serialization_length(c);
    std::vector<uint8_t> h(per * msgs.size());
    for (size_t i = 0; i < msgs.size(); i++) {
        const pb::Rq r = pb::parse_rq(msgs[i]);
        if (r.representation == pb::REPR_UNKNOWN || r.representation > pb::REPR_NTTSHOUP)
            throw Error(-1, "Unknown representation");
        if (r.representation != repr)
            throw Error(-1, "The representation asked for does not match the representation in the serialization");
        if (r.degree % 8 || r.degree < 8) throw Error(-1, "Invalid degree");
        if (r.degree != c.N || r.coefficients.len != per) throw Error(-1, "Invalid coefficients");
        std::memcpy(h.data() + i * per, r.coefficients.data, per);
    }
    DevBuf d(h.size());
    HIPCHK(hipMemcpyAsync(d.ptr, h.data(), h.size(), hipMemcpyHostToDevice, s));
    deserialize_polys(c, d.as<uint8_t>(), (u32)msgs.size(), true, d_out, s);
    HIPCHK(hipStreamSynchronize(s));
}

std::vector<uint8_t> seed_of(pb::Span s) {
    if (s.len != 32) throw Error(-1, "Invalid seed (32 bytes expected)");
    return std::vector<uint8_t>(s.data, s.data + 32);
}

// KeySwitchingKey::try_convert_from (fhe/src/bfv/keys/key_switching_key.rs:354-426).
std::shared_ptr<Ksk> ksk_from_message(const Parameters& par, pb::Span bytes) {
    const pb::KeySwitchingKey m = pb::parse_ksk(bytes);
    const auto& key_ctx = level_ctx(par, m.ksk_level);
    const auto& ct_ctx = level_ctx(par, m.ciphertext_level);
    size_t c0_size;
    if (m.log_base) {
        const u32 max_level = (u32)par.levels.size() - 1;
        if (m.ksk_level != max_level || m.ciphertext_level != max_level)
            throw Error(-1, "A decomposition size is specified but the levels are not maximal");
        u32 log_modulus = 0;
        while (log_modulus < 64 && ((u64)1 << log_modulus) < par.moduli[0]) log_modulus++;
        if (m.log_base != log_modulus / 2)
            throw Error(FHEHIP_ERR_UNSUPPORTED, "only the decomposition of KeySwitchingKey::new (log2(q) / 2 bits)");
        c0_size = (log_modulus + m.log_base - 1) / m.log_base;
    } else {
        c0_size = ct_ctx->L;
    }
    if (m.c0.size() != c0_size) throw Error(-1, "Incorrect number of values in c0");
    std::vector<uint8_t> seed;
    if (m.seed.len == 0) {
        if (m.c1.size() != c0_size) throw Error(-1, "Incorrect number of values in c1");
    } else {
        seed = seed_of(m.seed);
    }
    const size_t row = (size_t)key_ctx->L * key_ctx->N;
    std::vector<u64> c0(c0_size * row), c1(c0_size * row);
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    try {
        DevBuf d(c0.size() * 8);
        rq_to_device(*key_ctx, m.c0, pb::REPR_NTTSHOUP, d.as<u64>(), s);
        HIPCHK(hipMemcpy(c0.data(), d.ptr, d.bytes, hipMemcpyDeviceToHost));
        if (seed.empty()) {
            rq_to_device(*key_ctx, m.c1, pb::REPR_NTTSHOUP, d.as<u64>(), s);
            HIPCHK(hipMemcpy(c1.data(), d.ptr, d.bytes, hipMemcpyDeviceToHost));
        } else {
            Seed32 sd;
            std::memcpy(sd.data(), seed.data(), 32);
            generate_c1(key_ctx->moduli, key_ctx->N, sd, (u32)c0_size, c1.data());
        }
    } catch (...) {
        (void)hipStreamDestroy(s);
        throw;
    }
    HIPCHK(hipStreamDestroy(s));
    auto k = std::make_shared<Ksk>(ct_ctx, key_ctx, c0.data(), c1.data(), (u32)c0_size);
    k->seed = std::move(seed);
    return k;
}

// From<&KeySwitchingKey> for KeySwitchingKeyProto (key_switching_key.rs:332-352).
std::vector<uint8_t> ksk_message(const Parameters& par, const Ksk& k, bool variable_time) {
    const u32 ct_level = level_of(par, *k.ct_ctx), key_level = level_of(par, *k.key_ctx);
    const DeviceContext& kc = *k.key_ctx;
    const size_t row = (size_t)kc.L * kc.N, rows = (size_t)k.digits * row;
    std::vector<ulonglong2> h(rows);
    std::vector<u64> vals(rows);
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::vector<std::vector<uint8_t>> c0, c1;
    try {
        DevBuf d(rows * 8);
        for (int part = 0; part < 2; part++) {
            if (part == 1 && !k.seed.empty()) break;  // the seed stands for c1
            HIPCHK(hipMemcpy(h.data(), (part ? k.key1 : k.key0).ptr, rows * sizeof(ulonglong2), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < rows; i++) vals[i] = h[i].x;
            HIPCHK(hipMemcpy(d.ptr, vals.data(), rows * 8, hipMemcpyHostToDevice));
            const std::vector<uint8_t> pl = payloads(kc, d.as<u64>(), k.digits, s);
            const size_t per = pl.size() / k.digits;
            for (u32 r = 0; r < k.digits; r++)
                (part ? c1 : c0).push_back(pb::write_rq(pb::REPR_NTTSHOUP, kc.N, pl.data() + r * per, per, variable_time));
        }
    } catch (...) {
        (void)hipStreamDestroy(s);
        throw;
    }
    HIPCHK(hipStreamDestroy(s));
    return pb::write_ksk(c0, c1, k.seed.empty() ? nullptr : k.seed.data(), ct_level, key_level, k.log_base);
}

std::vector<uint8_t> wrap_ksk(const std::vector<uint8_t>& ksk, uint32_t exponent) {
    pb::Writer w;
    w.bytes_field(1, ksk.data(), ksk.size());
    w.uint_field(2, exponent);
    return std::move(w.out);
}

}  // namespace

extern "C" {

int fhehip_ciphertext_to_bytes(const fhehip_parameters* p, uint32_t level, const uint64_t* d_ct, uint32_t parts,
                               const uint8_t* seed, uint32_t variable_time_mask, uint8_t* out, size_t cap,
                               size_t* len, void* stream) {
    FHEHIP_API_BEGIN
    if (!p || !d_ct || !len) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (parts < (seed ? 2u : 1u)) return fail(FHEHIP_ERR_INVALID_ARG, "too few polynomials");
    const auto& ctx = level_ctx(*p->p, level);
    const u32 written = seed ? parts - 1 : parts;
    const std::vector<uint8_t> pl = payloads(*ctx, d_ct, written, S(stream));
    const size_t per = serialization_length(*ctx);
    pb::Writer w;
    for (u32 i = 0; i < written; i++) {
        const bool vt = i < 32 && ((variable_time_mask >> i) & 1);
        const std::vector<uint8_t> rq = pb::write_rq(pb::REPR_NTT, ctx->N, pl.data() + i * per, per, vt);
        w.bytes_field(1, rq.data(), rq.size());
    }
    if (seed) w.bytes_field(2, seed, 32);
    w.uint_field(3, level);
    return copy_message(w.out, out, cap, len);
    FHEHIP_API_END
}

int fhehip_ciphertext_from_bytes(const fhehip_parameters* p, const uint8_t* bytes, size_t len, uint64_t* d_out,
                                 uint32_t max_parts, uint32_t* parts, uint32_t* level, uint8_t* seed_out,
                                 int* seeded, void* stream) {
    FHEHIP_API_BEGIN
    if (!p || !bytes || !parts) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    const pb::Ciphertext m = pb::parse_ciphertext(pb::Span{bytes, len});
    if (m.c.empty() || (m.c.size() == 1 && m.seed.len == 0))
        return fail(FHEHIP_ERR_INVALID_ARG, "Not enough polynomials");
    const auto& ctx = level_ctx(*p->p, m.level);
    const u32 total = (u32)m.c.size() + (m.seed.len ? 1 : 0);
    std::vector<uint8_t> seed;
    if (m.seed.len) seed = seed_of(m.seed);
    *parts = total;
    if (level) *level = m.level;
    if (seeded) *seeded = seed.empty() ? 0 : 1;
    if (seed_out && !seed.empty()) std::memcpy(seed_out, seed.data(), 32);
    if (!d_out && !max_parts) return FHEHIP_OK;  // size query
    if (!d_out || total > max_parts) return fail(FHEHIP_ERR_INVALID_ARG, "more polynomials than max_parts");
    rq_to_device(*ctx, m.c, pb::REPR_NTT, d_out, S(stream));
    if (!seed.empty()) {
        // Poly::<Ntt>::random_from_seed (ciphertext.rs:224-235)
        const size_t row = (size_t)ctx->L * ctx->N;
        std::vector<u64> c1(row);
        Seed32 sd;
        std::memcpy(sd.data(), seed.data(), 32);
        random_from_seed(ctx->moduli, ctx->N, sd, c1.data());
        HIPCHK(hipMemcpyAsync(d_out + m.c.size() * row, c1.data(), row * 8, hipMemcpyHostToDevice, S(stream)));
        HIPCHK(hipStreamSynchronize(S(stream)));
    }
    FHEHIP_API_END
}

int fhehip_ksk_from_bytes(const fhehip_parameters* p, const uint8_t* bytes, size_t len, fhehip_ksk** out) {
    FHEHIP_API_BEGIN
    if (!p || !bytes || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    auto k = ksk_from_message(*p->p, pb::Span{bytes, len});
    *out = new fhehip_ksk{k};
    FHEHIP_API_END
}

int fhehip_ksk_to_bytes(const fhehip_parameters* p, const fhehip_ksk* k, int variable_time, uint8_t* out, size_t cap,
                        size_t* len) {
    FHEHIP_API_BEGIN
    if (!p || !k || !len) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    return copy_message(ksk_message(*p->p, *k->k, variable_time != 0), out, cap, len);
    FHEHIP_API_END
}

int fhehip_relinearization_key_from_bytes(const fhehip_parameters* p, const uint8_t* bytes, size_t len,
                                          fhehip_ksk** out) {
    FHEHIP_API_BEGIN
    if (!p || !bytes || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    const pb::WrappedKsk w = pb::parse_wrapped_ksk(pb::Span{bytes, len});
    if (!w.has_ksk) return fail(FHEHIP_ERR_INVALID_ARG, "Invalid serialization");
    auto k = ksk_from_message(*p->p, w.ksk);
    *out = new fhehip_ksk{k};
    FHEHIP_API_END
}

int fhehip_relinearization_key_to_bytes(const fhehip_parameters* p, const fhehip_ksk* k, int variable_time,
                                        uint8_t* out, size_t cap, size_t* len) {
    FHEHIP_API_BEGIN
    if (!p || !k || !len) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    return copy_message(wrap_ksk(ksk_message(*p->p, *k->k, variable_time != 0), 0), out, cap, len);
    FHEHIP_API_END
}

int fhehip_galois_key_from_bytes(const fhehip_parameters* p, const uint8_t* bytes, size_t len, fhehip_galois** out) {
    FHEHIP_API_BEGIN
    if (!p || !bytes || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    const pb::WrappedKsk w = pb::parse_wrapped_ksk(pb::Span{bytes, len});
    if (!w.has_ksk) return fail(FHEHIP_ERR_INVALID_ARG, "Invalid serialization");
    auto g = std::make_shared<Galois>(ksk_from_message(*p->p, w.ksk), w.exponent);
    *out = new fhehip_galois{g};
    FHEHIP_API_END
}

int fhehip_galois_key_to_bytes(const fhehip_parameters* p, const fhehip_galois* g, int variable_time, uint8_t* out,
                               size_t cap, size_t* len) {
    FHEHIP_API_BEGIN
    if (!p || !g || !len) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    return copy_message(wrap_ksk(ksk_message(*p->p, *g->g->ksk, variable_time != 0), g->g->exponent), out, cap, len);
    FHEHIP_API_END
}

int fhehip_evaluation_key_from_bytes(const fhehip_parameters* p, const uint8_t* bytes, size_t len,
                                     fhehip_evaluation_key** out, uint32_t* ciphertext_level) {
    FHEHIP_API_BEGIN
    if (!p || !bytes || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    const pb::EvaluationKey m = pb::parse_evaluation_key(pb::Span{bytes, len});
    const auto& ct_ctx = level_ctx(*p->p, m.ciphertext_level);
    level_ctx(*p->p, m.evaluation_key_level);
    std::unique_ptr<EvaluationKey> ek(new EvaluationKey);
    for (const pb::Span& gs : m.gk) {
        const pb::WrappedKsk w = pb::parse_wrapped_ksk(gs);
        if (!w.has_ksk) return fail(FHEHIP_ERR_INVALID_ARG, "Invalid serialization");
        auto k = ksk_from_message(*p->p, w.ksk);
        if (level_of(*p->p, *k->ct_ctx) != m.ciphertext_level)
            return fail(FHEHIP_ERR_INVALID_ARG, "Galois key has incorrect ciphertext level");
        if (level_of(*p->p, *k->key_ctx) != m.evaluation_key_level)
            return fail(FHEHIP_ERR_INVALID_ARG, "Galois key has incorrect evaluation key level");
        auto g = std::make_shared<Galois>(k, w.exponent);
        // one key per exponent, the last one kept (a map in fhe.rs)
        auto same = std::find_if(ek->keys.begin(), ek->keys.end(),
                                 [&](const std::shared_ptr<Galois>& x) { return x->exponent == g->exponent; });
        if (same != ek->keys.end()) *same = g;
        else ek->keys.push_back(g);
    }
    ek->degree = ct_ctx->N;
    if (!ek->keys.empty()) ek->prepare_expansion();
    if (ciphertext_level) *ciphertext_level = m.ciphertext_level;
    *out = new fhehip_evaluation_key{std::move(ek)};
    FHEHIP_API_END
}

int fhehip_evaluation_key_to_bytes(const fhehip_parameters* p, const fhehip_evaluation_key* ek, int variable_time,
                                   uint8_t* out, size_t cap, size_t* len) {
    FHEHIP_API_BEGIN
    if (!p || !ek || !len) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (ek->ek->keys.empty()) return fail(FHEHIP_ERR_INVALID_ARG, "an evaluation key without Galois keys");
    pb::Writer w;
    for (const auto& g : ek->ek->keys) {
        const std::vector<uint8_t> gk = wrap_ksk(ksk_message(*p->p, *g->ksk, variable_time != 0), g->exponent);
        w.bytes_field(2, gk.data(), gk.size());
    }
    const Ksk& k0 = *ek->ek->keys[0]->ksk;
    w.uint_field(3, level_of(*p->p, *k0.ct_ctx));
    w.uint_field(4, level_of(*p->p, *k0.key_ctx));
    return copy_message(w.out, out, cap, len);
    FHEHIP_API_END
}

int fhehip_parameters_from_bytes(const uint8_t* bytes, size_t len, fhehip_parameters** out) {
    FHEHIP_API_BEGIN
    if (!bytes || !out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    const pb::Parameters m = pb::parse_parameters(pb::Span{bytes, len});
    if (!m.plaintext_kind) return fail(FHEHIP_ERR_INVALID_ARG, "Missing required field: Parameters.plaintext_modulus");
    u64 t = m.plaintext;
    if (m.plaintext_kind == 2) {
        // BigUint::from_bytes_le; only values of at most 62 bits are supported
        t = 0;
        for (size_t i = 0; i < m.plaintext_big.len; i++) {
            const uint8_t b = m.plaintext_big.data[i];
            if (b && i >= 8) return fail(FHEHIP_ERR_UNSUPPORTED, "plaintext modulus above 2^62");
            if (i < 8) t |= (u64)b << (8 * i);
        }
    }
    if (m.variance < 1 || m.variance > 16)
        return fail(FHEHIP_ERR_INVALID_ARG, "The variance should be an integer between 1 and 16");
    std::unique_ptr<Parameters> par(new Parameters(m.degree, t, m.moduli));
    par->variance = m.variance;
    *out = new fhehip_parameters{std::move(par)};
    FHEHIP_API_END
}

int fhehip_parameters_to_bytes(const fhehip_parameters* p, uint8_t* out, size_t cap, size_t* len) {
    FHEHIP_API_BEGIN
    if (!p || !len) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    return copy_message(pb::write_parameters(p->p->N, p->p->moduli, p->p->t, p->p->variance), out, cap, len);
    FHEHIP_API_END
}

static bool ntt_friendly(u64 p, u32 degree) {
    return degree >= 8 && !(degree & (degree - 1)) && p >= 2 && !(p >> 62) && supports_ntt64(p, degree);
}

int fhehip_default_psi(uint32_t degree, uint64_t modulus, uint64_t* psi) {
    FHEHIP_API_BEGIN
    if (!psi) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (!ntt_friendly(modulus, degree)) return fail(FHEHIP_ERR_UNSUPPORTED, "not an NTT-friendly prime for degree");
    u64 r = fhers_primitive_root(degree, modulus);
    *psi = r ? r : primitive_root64(degree, modulus);
    FHEHIP_API_END
}

static int check_seeded(const uint64_t* moduli, uint32_t num_moduli, uint32_t degree, const uint8_t* seed,
                        const uint64_t* out) {
    if (!moduli || !seed || !out || !num_moduli) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    if (degree < 8 || (degree & (degree - 1))) return fail(FHEHIP_ERR_INVALID_ARG, "degree is not a power of two >= 8");
    for (uint32_t i = 0; i < num_moduli; i++)
        if (moduli[i] < 2 || (moduli[i] >> 62)) return fail(FHEHIP_ERR_INVALID_ARG, "InvalidModulus");
    return FHEHIP_OK;
}

int fhehip_random_from_seed(const uint64_t* moduli, uint32_t num_moduli, uint32_t degree, const uint8_t* seed,
                            uint64_t* out) {
    FHEHIP_API_BEGIN
    if (int rc = check_seeded(moduli, num_moduli, degree, seed, out)) return rc;
    Seed32 s;
    std::memcpy(s.data(), seed, 32);
    random_from_seed(std::vector<u64>(moduli, moduli + num_moduli), degree, s, out);
    FHEHIP_API_END
}

int fhehip_generate_c1(const uint64_t* moduli, uint32_t num_moduli, uint32_t degree, const uint8_t* seed,
                       uint32_t count, uint64_t* out) {
    FHEHIP_API_BEGIN
    if (int rc = check_seeded(moduli, num_moduli, degree, seed, out)) return rc;
    Seed32 s;
    std::memcpy(s.data(), seed, 32);
    generate_c1(std::vector<u64>(moduli, moduli + num_moduli), degree, s, count, out);
    FHEHIP_API_END
}

}  // extern "C"
