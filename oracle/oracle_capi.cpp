// ORACLE — test infrastructure only. C entry points over the restatement so that
// Python tests (ctypes) can drive it. Every function returns 0 on success and a
// negative value on error; the error text is available from or_last_error().
#include <chrono>
#include <cstdio>
#include <thread>

#include "oracle_bfv.hpp"

using namespace oracle;

static thread_local std::string g_err;

#define OR_TRY(...)                       \
    try {                                 \
        __VA_ARGS__;                      \
        return 0;                         \
    } catch (const std::exception& e) {   \
        g_err = e.what();                 \
        return -1;                        \
    }

extern "C" {

const char* or_last_error() { return g_err.c_str(); }

// ---- Zq / primes ----------------------------------------------------------
uint64_t or_generate_prime(uint32_t bits, uint64_t modulo, uint64_t upper) {
    return generate_prime(bits, modulo, upper);
}
int or_supports_opt(uint64_t p) { return supports_opt(p) ? 1 : 0; }
int or_is_prime(uint64_t p) { return is_prime(p) ? 1 : 0; }
int or_generate_moduli(const uint32_t* sizes, uint32_t n, uint32_t degree, uint64_t* out) {
    OR_TRY({
        std::vector<size_t> s(sizes, sizes + n);
        std::vector<u64> m = generate_moduli(s, degree);
        std::copy(m.begin(), m.end(), out);
    })
}
uint64_t or_primitive_root(uint32_t n, uint64_t p) {
    try { return primitive_root(n, Modulus(p)); } catch (...) { return 0; }
}

// ---- Pseudo-random streams (oracle_rng.hpp) ------------------------------------
static std::array<uint8_t, 32> seed32(const uint8_t* s) {
    std::array<uint8_t, 32> a;
    std::memcpy(a.data(), s, 32);
    return a;
}
void or_sha256(const uint8_t* data, size_t len, uint8_t* out) {
    const auto h = oracle_rng::sha256(data, len);
    std::memcpy(out, h.data(), 32);
}
void or_chacha_block(const uint32_t* in, int rounds, uint32_t* out) { oracle_rng::chacha_block(in, rounds, out); }
// kind 0: n x next_u32 (u32 out), 1: n x next_u64, 2: n bytes of fill_bytes;
// kind 3: n x next_u32 then next_u64 alternating (u64 out) for the buffer edges.
void or_chacha8_stream(const uint8_t* seed, int kind, size_t n, void* out) {
    oracle_rng::ChaCha8Rng r(seed32(seed));
    if (kind == 0) for (size_t i = 0; i < n; i++) ((uint32_t*)out)[i] = r.next_u32();
    else if (kind == 1) for (size_t i = 0; i < n; i++) ((uint64_t*)out)[i] = r.next_u64();
    else if (kind == 2) r.fill_bytes((uint8_t*)out, n);
    else for (size_t i = 0; i < n; i++) ((uint64_t*)out)[i] = (i & 1) ? r.next_u64() : r.next_u32();
}
void or_pcg32_seed(uint64_t state, uint8_t* out) {
    const auto s = oracle_rng::pcg32_seed(state);
    std::memcpy(out, s.data(), 32);
}
uint32_t or_pcg32_output(uint64_t state) { return oracle_rng::pcg32_output(state); }
// n draws of Uniform::new(0, range).sample (single = 0) or random_range(low..high) (single = 1)
void or_uniform(const uint8_t* seed, int single, uint64_t low, uint64_t high, size_t n, uint64_t* out) {
    oracle_rng::ChaCha8Rng r(seed32(seed));
    for (size_t i = 0; i < n; i++)
        out[i] = single ? oracle_rng::uniform_single(r, low, high) : oracle_rng::uniform_sample(r, high);
}
void or_random_from_seed(const uint64_t* moduli, uint32_t L, uint32_t n, const uint8_t* seed, uint64_t* out) {
    const auto v = oracle_rng::random_from_seed(std::vector<u64>(moduli, moduli + L), n, seed32(seed));
    std::copy(v.begin(), v.end(), out);
}
void or_generate_c1(const uint64_t* moduli, uint32_t L, uint32_t n, const uint8_t* seed, uint32_t size,
                    uint64_t* out) {
    const auto v = oracle_rng::generate_c1(std::vector<u64>(moduli, moduli + L), n, seed32(seed), size);
    std::copy(v.begin(), v.end(), out);
}

// Elementwise Modulus operations for the property tests.
// op: 0 add, 1 sub, 2 mul, 3 mul_opt, 4 shoup(a), 5 mul_shoup(a,b), 6 lazy_mul_shoup,
//     7 reduce(a), 8 lazy_reduce(a), 9 reduce_u128(a_lo | b_hi<<64), 10 lazy_reduce_u128,
//     11 neg, 12 lazy_reduce_opt(a), 13 reduce_opt_u128(a|b<<64), 14 pow(a,b), 15 inv(a),
//     16 center(a) (as i64 bits), 17 reduce_i64(a as i64)
int or_modulus_op(uint64_t p, uint32_t op, const uint64_t* a, const uint64_t* b, uint32_t n, uint64_t* out) {
    OR_TRY({
        Modulus m(p);
        for (uint32_t i = 0; i < n; i++) {
            u64 x = a[i], y = b ? b[i] : 0;
            u64 r = 0;
            switch (op) {
                case 0: r = m.add(x, y); break;
                case 1: r = m.sub(x, y); break;
                case 2: r = m.mul(x, y); break;
                case 3: r = m.mul_opt(x, y); break;
                case 4: r = m.shoup(x); break;
                case 5: r = m.mul_shoup(x, y, m.shoup(y)); break;
                case 6: r = m.lazy_mul_shoup(x, y, m.shoup(y)); break;
                case 7: r = m.reduce(x); break;
                case 8: r = m.lazy_reduce(x); break;
                case 9: r = m.reduce_u128(((u128)y << 64) | x); break;
                case 10: r = m.lazy_reduce_u128(((u128)y << 64) | x); break;
                case 11: r = m.neg(x); break;
                case 12: r = m.lazy_reduce_opt(x); break;
                case 13: r = m.reduce_opt_u128(((u128)y << 64) | x); break;
                case 14: r = m.pow(x, y); break;
                case 15: r = m.inv(x); break;
                case 16: r = (u64)m.center(x); break;
                case 17: r = m.reduce_i64((i64)x); break;
                default: throw std::runtime_error("bad op");
            }
            out[i] = r;
        }
    })
}

// ---- NTT -----------------------------------------------------------------
// variant: 0 forward, 1 forward_vt, 2 forward_vt_lazy, 3 backward
int or_ntt(uint64_t p, uint32_t n, uint64_t psi, uint32_t variant, uint64_t* a, uint32_t count) {
    OR_TRY({
        NttOperator op(Modulus(p), n, psi);
        for (uint32_t c = 0; c < count; c++) {
            u64* x = a + (size_t)c * n;
            switch (variant) {
                case 0: op.forward(x); break;
                case 1: op.forward_vt(x); break;
                case 2: op.forward_lazy(x); break;
                case 3: op.backward(x); break;
                default: throw std::runtime_error("bad variant");
            }
        }
    })
}
// Twiddle tables as NttOperator::new builds them: [omegas, omegas_shoup, zetas_inv, zetas_inv_shoup]
int or_ntt_tables(uint64_t p, uint32_t n, uint64_t psi, uint64_t* out4n, uint64_t* size_inv) {
    OR_TRY({
        NttOperator op(Modulus(p), n, psi);
        std::copy(op.omegas.begin(), op.omegas.end(), out4n);
        std::copy(op.omegas_shoup.begin(), op.omegas_shoup.end(), out4n + n);
        std::copy(op.zetas_inv.begin(), op.zetas_inv.end(), out4n + 2 * n);
        std::copy(op.zetas_inv_shoup.begin(), op.zetas_inv_shoup.end(), out4n + 3 * n);
        *size_inv = op.size_inv;
    })
}

// ---- RNS -------------------------------------------------------------------
struct OrBig {
    static Big from_words(const uint32_t* w, uint32_t n) {
        Big b;
        b.w.assign(w, w + n);
        b.trim();
        return b;
    }
};
// garner coefficients as 32-bit words, garner_i at out + i*stride
int or_rns_garner(const uint64_t* moduli, uint32_t n, uint32_t* out, uint32_t stride) {
    OR_TRY({
        RnsContext r(std::vector<u64>(moduli, moduli + n));
        for (uint32_t i = 0; i < n; i++) {
            if (r.garner[i].w.size() > stride) throw std::runtime_error("stride too small");
            std::fill(out + i * stride, out + (i + 1) * stride, 0u);
            std::copy(r.garner[i].w.begin(), r.garner[i].w.end(), out + i * stride);
        }
    })
}
int or_rns_lift(const uint64_t* moduli, uint32_t n, const uint64_t* rests, uint32_t* out, uint32_t stride) {
    OR_TRY({
        RnsContext r(std::vector<u64>(moduli, moduli + n));
        Big b = r.lift(rests);
        if (b.w.size() > stride) throw std::runtime_error("stride too small");
        std::fill(out, out + stride, 0u);
        std::copy(b.w.begin(), b.w.end(), out);
    })
}

struct OrScaler {
    std::unique_ptr<RnsContext> from, to;
    std::unique_ptr<RnsScaler> s;
};
void* or_rns_scaler_new(const uint64_t* from, uint32_t nf, const uint64_t* to, uint32_t nt,
                        const uint32_t* num, uint32_t nnum, const uint32_t* den, uint32_t nden) {
    try {
        OrScaler* h = new OrScaler;
        h->from.reset(new RnsContext(std::vector<u64>(from, from + nf)));
        h->to.reset(new RnsContext(std::vector<u64>(to, to + nt)));
        h->s.reset(new RnsScaler(h->from.get(), h->to.get(),
                                 ScalingFactor(OrBig::from_words(num, nnum), OrBig::from_words(den, nden))));
        return h;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void or_rns_scaler_free(void* h) { delete (OrScaler*)h; }
// rests: [count][nf] ; out: [count][len]
int or_rns_scale(void* h, const uint64_t* rests, uint32_t count, uint64_t* out, uint32_t start, uint32_t len) {
    OR_TRY({
        OrScaler* s = (OrScaler*)h;
        size_t nf = s->from->moduli.size();
        for (uint32_t c = 0; c < count; c++) s->s->scale(rests + c * nf, out + (size_t)c * len, len, start);
    })
}
// Scaler constants in the layout the tests compare against.
int or_rns_scaler_consts(void* h, uint64_t* gamma, uint64_t* theta_gamma /*lo,hi,sign*/,
                         uint64_t* omega /*[nt][nf]*/, uint64_t* theta_omega /*[nf][3]*/,
                         uint64_t* theta_garner /*[nf][2]*/, uint64_t* shift) {
    OR_TRY({
        RnsScaler& s = *((OrScaler*)h)->s;
        size_t nf = s.from->moduli.size(), nt = s.to->moduli.size();
        std::copy(s.gamma.begin(), s.gamma.end(), gamma);
        theta_gamma[0] = s.theta_gamma_lo; theta_gamma[1] = s.theta_gamma_hi; theta_gamma[2] = s.theta_gamma_sign;
        for (size_t j = 0; j < nt; j++)
            for (size_t i = 0; i < nf; i++) omega[j * nf + i] = s.omega[j][i];
        for (size_t i = 0; i < nf; i++) {
            theta_omega[3 * i] = s.theta_omega_lo[i];
            theta_omega[3 * i + 1] = s.theta_omega_hi[i];
            theta_omega[3 * i + 2] = s.theta_omega_sign[i];
            theta_garner[2 * i] = s.theta_garner_lo[i];
            theta_garner[2 * i + 1] = s.theta_garner_hi[i];
        }
        *shift = s.theta_garner_shift;
    })
}

// ---- rq ----------------------------------------------------------------------
void* or_ctx_new(const uint64_t* moduli, uint32_t L, uint32_t n) {
    try {
        return new CtxPtr(std::make_shared<Context>(std::vector<u64>(moduli, moduli + L), n));
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void or_ctx_free(void* h) { delete (CtxPtr*)h; }
int or_ctx_psi(void* h, uint64_t* psi) {
    OR_TRY({
        const Context& c = **(CtxPtr*)h;
        for (size_t i = 0; i < c.L(); i++) psi[i] = c.ops[i].psi;
    })
}
// poly ops on [count][L][N]; op 0 forward, 1 backward
int or_ctx_ntt(void* h, uint32_t op, uint64_t* a, uint32_t count) {
    OR_TRY({
        const Context& c = **(CtxPtr*)h;
        for (uint32_t k = 0; k < count; k++) {
            u64* x = a + (size_t)k * c.L() * c.degree;
            if (op == 0) ntt_forward(c, x); else ntt_backward(c, x);
        }
    })
}
int or_switch_down(void* h, const uint64_t* in, uint64_t* out, uint32_t count) {
    OR_TRY({
        const Context& c = **(CtxPtr*)h;
        size_t in_sz = c.L() * c.degree, out_sz = (c.L() - 1) * c.degree;
        for (uint32_t k = 0; k < count; k++) {
            PolyData r = switch_down(c, in + k * in_sz);
            std::copy(r.begin(), r.end(), out + k * out_sz);
        }
    })
}
int or_substitute(void* h, uint32_t exponent, uint32_t ntt_repr, const uint64_t* in, uint64_t* out,
                  uint32_t count) {
    OR_TRY({
        const Context& c = **(CtxPtr*)h;
        size_t sz = c.L() * c.degree;
        for (uint32_t k = 0; k < count; k++) {
            PolyData r = substitute(c, exponent, in + k * sz, ntt_repr != 0);
            std::copy(r.begin(), r.end(), out + k * sz);
        }
    })
}
void* or_poly_scaler_new(void* from, void* to, const uint32_t* num, uint32_t nnum, const uint32_t* den,
                         uint32_t nden) {
    try {
        ScalingFactor f = (nnum == 0) ? ScalingFactor()
                                      : ScalingFactor(OrBig::from_words(num, nnum), OrBig::from_words(den, nden));
        return new PolyScaler(*(CtxPtr*)from, *(CtxPtr*)to, f);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void or_poly_scaler_free(void* h) { delete (PolyScaler*)h; }
int or_poly_scale(void* h, uint32_t ntt_repr, const uint64_t* in, uint64_t* out, uint32_t count) {
    OR_TRY({
        PolyScaler& s = *(PolyScaler*)h;
        size_t isz = s.from->L() * s.from->degree, osz = s.to->L() * s.to->degree;
        for (uint32_t k = 0; k < count; k++) {
            PolyData r = s.scale(in + k * isz, ntt_repr != 0);
            std::copy(r.begin(), r.end(), out + k * osz);
        }
    })
}

// ---- BFV ---------------------------------------------------------------------
struct OrBfv {
    std::unique_ptr<BfvParams> par;
    SecretKey sk;
};
void* or_bfv_new(uint32_t degree, uint64_t t, const uint64_t* moduli, uint32_t L, uint32_t variance,
                 uint64_t sk_seed) {
    try {
        OrBfv* h = new OrBfv;
        h->par.reset(new BfvParams(degree, t, std::vector<u64>(moduli, moduli + L), variance));
        Rng rng(sk_seed);
        h->sk.coeffs = sample_vec_cbd(degree, variance, rng);
        return h;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void or_bfv_free(void* h) { delete (OrBfv*)h; }
int or_bfv_level_moduli(void* h, uint32_t level, uint64_t* out, uint32_t* L) {
    OR_TRY({
        const Context& c = ((OrBfv*)h)->par->ctx(level);
        std::copy(c.moduli.begin(), c.moduli.end(), out);
        *L = (uint32_t)c.L();
    })
}
int or_bfv_secret(void* h, int64_t* out) {
    OR_TRY({
        auto& s = ((OrBfv*)h)->sk.coeffs;
        std::copy(s.begin(), s.end(), out);
    })
}
// poly != 0: Encoding::poly_at_level instead of Encoding::simd.
int or_bfv_encrypt(void* h, uint32_t level, const uint64_t* slots, uint32_t nslots, uint64_t seed, uint32_t poly,
                   uint64_t* out) {
    OR_TRY({
        OrBfv& b = *(OrBfv*)h;
        Rng rng(seed);
        PolyData ct = encrypt(*b.par, b.sk, level, std::vector<u64>(slots, slots + nslots), rng, poly != 0);
        std::copy(ct.begin(), ct.end(), out);
    })
}
// Plaintext::try_encode (plaintext_vec.rs:37-101): `count` values (< t) as SIMD
// slots (poly == 0) or coefficients -> [L][N] NTT: with `scaled` the
// polynomial Plaintext::to_poly adds in encryption (plaintext.rs:71-98), else
// Plaintext::poly_ntt (the values reduced per modulus, plaintext_vec.rs:75).
int or_bfv_encode(void* h, uint32_t level, const uint64_t* values, uint32_t count, uint32_t poly, uint32_t scaled,
                  uint64_t* out) {
    OR_TRY({
        OrBfv& b = *(OrBfv*)h;
        std::vector<u64> v(values, values + count);
        v.resize(b.par->degree, 0);
        if (!poly) v = encode_simd(*b.par, std::vector<u64>(values, values + count));
        PolyData p;
        if (scaled) {
            p = plaintext_to_poly(*b.par, level, v);
        } else {
            p = poly_from_u64(b.par->ctx(level), v);
            ntt_forward(b.par->ctx(level), p.data());
        }
        std::copy(p.begin(), p.end(), out);
    })
}
// SecretKey::encrypt_poly with the given a ([L][N], NTT) and e (N signed coefficients).
int or_bfv_encrypt_with(void* h, uint32_t level, const uint64_t* pt, const uint64_t* a, const int64_t* e,
                        uint64_t* out) {
    OR_TRY({
        OrBfv& b = *(OrBfv*)h;
        PolyData ct = encrypt_with(*b.par, b.sk, level, pt, a, std::vector<i64>(e, e + b.par->degree));
        std::copy(ct.begin(), ct.end(), out);
    })
}
// PublicKey::try_encrypt with the given u, e1, e2 (N signed coefficients each).
int or_bfv_encrypt_pk_with(void* h, uint32_t level, const uint64_t* pk, const uint64_t* pt, const int64_t* u,
                           const int64_t* e1, const int64_t* e2, uint64_t* out) {
    OR_TRY({
        OrBfv& b = *(OrBfv*)h;
        size_t n = b.par->degree;
        PolyData ct = encrypt_pk_with(*b.par, level, pk, pt, std::vector<i64>(u, u + n), std::vector<i64>(e1, e1 + n),
                                      std::vector<i64>(e2, e2 + n));
        std::copy(ct.begin(), ct.end(), out);
    })
}
int or_bfv_decrypt(void* h, uint32_t level, const uint64_t* ct, uint32_t nparts, uint32_t poly, uint64_t* slots) {
    OR_TRY({
        OrBfv& b = *(OrBfv*)h;
        std::vector<u64> s = decrypt(*b.par, b.sk, level, ct, nparts, poly != 0);
        std::copy(s.begin(), s.end(), slots);
    })
}
// kind 0: relinearization key ; kind 1: Galois key with `exponent`
void* or_ksk_new(void* h, uint32_t kind, uint32_t exponent, uint32_t ct_level, uint32_t key_level, uint64_t seed) {
    try {
        OrBfv& b = *(OrBfv*)h;
        Rng rng(seed);
        if (kind == 0) return relin_key_new(*b.par, b.sk, ct_level, key_level, rng);
        return galois_key_new(*b.par, b.sk, exponent, ct_level, key_level, rng);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
// One key of an RGSW encryption of the plaintext whose poly_ntt is pt_ntt.
void* or_rgsw_key_new(void* h, uint32_t level, const uint64_t* pt_ntt, uint32_t which, uint64_t seed) {
    try {
        OrBfv& b = *(OrBfv*)h;
        Rng rng(seed);
        return rgsw_key_new(*b.par, b.sk, level, pt_ntt, which, rng);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
int or_rgsw_external_product(void* k0, void* k1, const uint64_t* ct, uint64_t* out, uint32_t count) {
    OR_TRY({
        const KeySwitchingKey& a = *(KeySwitchingKey*)k0;
        const KeySwitchingKey& b = *(KeySwitchingKey*)k1;
        size_t sz = 2 * a.ctx_ct->L() * a.ctx_ct->degree;
        for (uint32_t i = 0; i < count; i++) {
            PolyData r = rgsw_external_product(a, b, ct + i * sz);
            std::copy(r.begin(), r.end(), out + i * sz);
        }
    })
}
void or_ksk_free(void* k) { delete (KeySwitchingKey*)k; }
void* or_ksk_from_rows(void* h, uint32_t ct_level, uint32_t key_level, const uint64_t* c0, const uint64_t* c1) {
    try {
        OrBfv& b = *(OrBfv*)h;
        return new KeySwitchingKey(*b.par, ct_level, key_level, c0, c1);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
// c0, c1: [digits][L'][N] NTT-domain values
int or_ksk_export(void* k, uint64_t* c0, uint64_t* c1, uint32_t* digits, uint32_t* key_L) {
    OR_TRY({
        KeySwitchingKey& ks = *(KeySwitchingKey*)k;
        size_t sz = ks.ctx_ksk->L() * ks.ctx_ksk->degree;
        for (size_t i = 0; i < ks.c0.size(); i++) {
            if (c0) std::copy(ks.c0[i].begin(), ks.c0[i].end(), c0 + i * sz);
            if (c1) std::copy(ks.c1[i].begin(), ks.c1[i].end(), c1 + i * sz);
        }
        *digits = (uint32_t)ks.c0.size();
        *key_L = (uint32_t)ks.ctx_ksk->L();
    })
}
int or_key_switch(void* k, const uint64_t* p, uint64_t* c0, uint64_t* c1, uint32_t count) {
    OR_TRY({
        KeySwitchingKey& ks = *(KeySwitchingKey*)k;
        size_t isz = ks.ctx_ct->L() * ks.ctx_ct->degree, osz = ks.ctx_ksk->L() * ks.ctx_ksk->degree;
        for (uint32_t i = 0; i < count; i++) {
            PolyData a, b;
            ks.key_switch(p + i * isz, a, b);
            std::copy(a.begin(), a.end(), c0 + i * osz);
            std::copy(b.begin(), b.end(), c1 + i * osz);
        }
    })
}
void* or_mult_new(void* h, void* rk, uint32_t level) {
    try {
        OrBfv& b = *(OrBfv*)h;
        return new Multiplicator(*b.par, (KeySwitchingKey*)rk, level);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
// Multiplicator::new_leveled with factors given as 32-bit words (n = 0: one).
void* or_mult_new_general(void* h, void* rk, uint32_t level, const uint64_t* ext, uint32_t n_ext,
                          const uint32_t* const* words, const uint32_t* counts) {
    try {
        OrBfv& b = *(OrBfv*)h;
        ScalingFactor f[3];
        for (int i = 0; i < 3; i++)
            if (counts[2 * i])
                f[i] = ScalingFactor(OrBig::from_words(words[2 * i], counts[2 * i]),
                                     OrBig::from_words(words[2 * i + 1], counts[2 * i + 1]));
        return new Multiplicator(*b.par, (KeySwitchingKey*)rk, level, std::vector<u64>(ext, ext + n_ext), f[0], f[1],
                                 f[2]);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void or_mult_free(void* m) { delete (Multiplicator*)m; }
int or_mult_ext_moduli(void* m, uint64_t* out, uint32_t* n) {
    OR_TRY({
        Multiplicator& M = *(Multiplicator*)m;
        std::copy(M.mul_ctx->moduli.begin(), M.mul_ctx->moduli.end(), out);
        *n = (uint32_t)M.mul_ctx->L();
    })
}
// lhs/rhs: [count][2][L][N] ; out: [count][2 or 3][L][N]
int or_mult_multiply(void* m, const uint64_t* lhs, const uint64_t* rhs, uint64_t* out, uint32_t count) {
    OR_TRY({
        Multiplicator& M = *(Multiplicator*)m;
        size_t sz = M.base_ctx->L() * M.base_ctx->degree;
        size_t osz = (M.rk ? 2 : 3) * sz;
        for (uint32_t i = 0; i < count; i++) {
            PolyData r = M.multiply(lhs + i * 2 * sz, rhs + i * 2 * sz);
            std::copy(r.begin(), r.end(), out + i * osz);
        }
    })
}
// ct x ct without relinearization through the per-level MultiplicationParameters
// (fhe/src/bfv/ops/mod.rs:259-342): [2] x [2] -> [3]
int or_ct_mul(void* h, uint32_t level, const uint64_t* lhs, const uint64_t* rhs, uint64_t* out) {
    OR_TRY({
        OrBfv& b = *(OrBfv*)h;
        const Level& L = b.par->levels[level];
        const Context& c = *L.ctx;
        const Context& m = *L.mp.to;
        size_t sz = c.L() * c.degree;
        PolyData a0 = L.mp.extender->scale(lhs, true), a1 = L.mp.extender->scale(lhs + sz, true);
        PolyData b0 = L.mp.extender->scale(rhs, true), b1 = L.mp.extender->scale(rhs + sz, true);
        PolyData c0 = a0, c1 = a0, t = a1, c2 = a1;
        poly_mul(m, c0.data(), b0.data());
        poly_mul(m, c1.data(), b1.data());
        poly_mul(m, t.data(), b0.data());
        poly_add(m, c1.data(), t.data());
        poly_mul(m, c2.data(), b1.data());
        PolyData d0 = L.mp.down_scaler->scale(c0.data(), true);
        PolyData d1 = L.mp.down_scaler->scale(c1.data(), true);
        PolyData d2 = L.mp.down_scaler->scale(c2.data(), true);
        std::copy(d0.begin(), d0.end(), out);
        std::copy(d1.begin(), d1.end(), out + sz);
        std::copy(d2.begin(), d2.end(), out + 2 * sz);
    })
}
// Mul<&Ciphertext> for &Ciphertext (fhe/src/bfv/ops/mod.rs:259-342) for any
// part counts: every part scaled by the level's extender (NTT representation),
// c[i + j] += a_i * b_j in the multiplication context, every product part
// scaled back by the down-scaler. out: [lp + rp - 1][L][N].
int or_ct_mul_parts(void* h, uint32_t level, const uint64_t* lhs, uint32_t lp, const uint64_t* rhs, uint32_t rp,
                    uint64_t* out) {
    OR_TRY({
        OrBfv& b = *(OrBfv*)h;
        if (!lp || !rp) throw std::runtime_error("empty ciphertext");
        const Level& L = b.par->levels.at(level);
        const Context& c = *L.ctx;
        const Context& m = *L.mp.to;
        const size_t sz = c.L() * c.degree;
        std::vector<PolyData> a, bb;
        for (uint32_t i = 0; i < lp; i++) a.push_back(L.mp.extender->scale(lhs + i * sz, true));
        for (uint32_t j = 0; j < rp; j++) bb.push_back(L.mp.extender->scale(rhs + j * sz, true));
        std::vector<PolyData> cc(lp + rp - 1, PolyData(m.L() * m.degree, 0));
        for (uint32_t i = 0; i < lp; i++)
            for (uint32_t j = 0; j < rp; j++) {
                PolyData t = a[i];
                poly_mul(m, t.data(), bb[j].data());
                poly_add(m, cc[i + j].data(), t.data());
            }
        for (size_t k = 0; k < cc.size(); k++) {
            PolyData d = L.mp.down_scaler->scale(cc[k].data(), true);
            std::copy(d.begin(), d.end(), out + k * sz);
        }
    })
}
// The multiplication context moduli of a level (MultiplicationParameters,
// fhe/src/bfv/parameters.rs:585-612).
int or_bfv_mul_moduli(void* h, uint32_t level, uint64_t* out, uint32_t* len) {
    OR_TRY({
        const Context& m = *((OrBfv*)h)->par->levels.at(level).mp.to;
        if (*len < m.L()) throw std::runtime_error("buffer too small");
        std::copy(m.moduli.begin(), m.moduli.end(), out);
        *len = (uint32_t)m.L();
    })
}
int or_relinearize(void* h, void* rk, const uint64_t* ct3, uint64_t* out) {
    OR_TRY({
        OrBfv& b = *(OrBfv*)h;
        KeySwitchingKey& k = *(KeySwitchingKey*)rk;
        PolyData r = Multiplicator::relinearize(*b.par, k, ct3);
        std::copy(r.begin(), r.end(), out);
    })
}
int or_galois(void* h, void* gk, uint32_t exponent, const uint64_t* ct, uint64_t* out, uint32_t count) {
    OR_TRY({
        OrBfv& b = *(OrBfv*)h;
        KeySwitchingKey& k = *(KeySwitchingKey*)gk;
        size_t sz = 2 * k.ctx_ct->L() * k.ctx_ct->degree;
        for (uint32_t i = 0; i < count; i++) {
            PolyData r = galois_apply(*b.par, k, exponent, ct + i * sz);
            std::copy(r.begin(), r.end(), out + i * sz);
        }
    })
}
// dot_product_scalar over `count` (ciphertext, plaintext) pairs; out [parts][L][N].
int or_dot_product_scalar(void* ctx, const uint64_t* cts, uint64_t ct_stride, const uint64_t* pts,
                          uint64_t pt_stride, uint32_t count, uint32_t parts, uint64_t* out) {
    OR_TRY({
        const Context& c = **(CtxPtr*)ctx;
        PolyData r = dot_product_scalar(c, cts, ct_stride, pts, pt_stride, count, parts);
        std::copy(r.begin(), r.end(), out);
    })
}
// EvaluationKey::expands with the Galois keys `gks` of exponents `exps`;
// out [size][2][L][N].
int or_expands(void* h, void* const* gks, const uint32_t* exps, uint32_t ngk, uint32_t ct_level, const uint64_t* ct,
               uint32_t size, uint64_t* out) {
    OR_TRY({
        OrBfv& b = *(OrBfv*)h;
        const Context& c = b.par->ctx(ct_level);
        auto key = [&](size_t l) -> const KeySwitchingKey* {
            for (uint32_t k = 0; k < ngk; k++)
                if (exps[k] == (c.degree >> l) + 1) return (const KeySwitchingKey*)gks[k];
            return nullptr;
        };
        std::vector<PolyData> r = expands(*b.par, c, key, ct, size);
        for (auto& x : r) out = std::copy(x.begin(), x.end(), out);
    })
}
// Rq coefficient payloads of `count` polynomials [L][N] of ctx, back to back.
int or_poly_serialize(void* ctx, const uint64_t* polys, uint32_t count, uint32_t ntt, uint8_t* out) {
    OR_TRY({
        const Context& c = **(CtxPtr*)ctx;
        for (uint32_t k = 0; k < count; k++) {
            std::vector<uint8_t> v = serialize_poly(c, polys + (size_t)k * c.L() * c.degree, ntt != 0);
            out = std::copy(v.begin(), v.end(), out);
        }
    })
}
int or_poly_deserialize(void* ctx, const uint8_t* bytes, uint32_t count, uint32_t ntt, uint64_t* out) {
    OR_TRY({
        const Context& c = **(CtxPtr*)ctx;
        size_t per = 0;
        for (u64 q : c.moduli) per += coeff_bits(q) * c.degree / 8;
        for (uint32_t k = 0; k < count; k++) {
            PolyData p = deserialize_poly(c, bytes + k * per, ntt != 0);
            out = std::copy(p.begin(), p.end(), out);
        }
    })
}
// Seeded uniform residues for [count][L][N] in the given moduli.
int or_random_polys(const uint64_t* moduli, uint32_t L, uint32_t n, uint32_t count, uint64_t seed, uint64_t* out) {
    OR_TRY({
        Rng rng(seed);
        for (uint32_t c = 0; c < count; c++)
            for (uint32_t l = 0; l < L; l++)
                for (uint32_t j = 0; j < n; j++) *out++ = rng.uniform(moduli[l]);
    })
}

// ---- CPU baseline ------------------------------------------------------------
// Times Multiplicator::multiply on `threads` worker threads, each running `reps`
// independent operations on its own copy of the inputs. Returns wall seconds and
// fills per-op latencies (seconds) of thread 0.
int or_bench_mul_relin(void* m, const uint64_t* lhs, const uint64_t* rhs, uint32_t threads, uint32_t reps,
                       double* wall_s, double* lat0) {
    OR_TRY({
        Multiplicator& M = *(Multiplicator*)m;
        auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> th;
        for (uint32_t k = 0; k < threads; k++) {
            th.emplace_back([&, k]() {
                for (uint32_t r = 0; r < reps; r++) {
                    auto a = std::chrono::steady_clock::now();
                    PolyData out = M.multiply(lhs, rhs);
                    auto b = std::chrono::steady_clock::now();
                    if (k == 0 && lat0) lat0[r] = std::chrono::duration<double>(b - a).count();
                }
            });
        }
        for (auto& t : th) t.join();
        *wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    })
}
int or_bench_ntt(uint64_t p, uint32_t n, uint64_t* a, uint32_t count, uint32_t threads, double* wall_s) {
    OR_TRY({
        NttOperator op(Modulus(p), n);
        auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> th;
        for (uint32_t k = 0; k < threads; k++) {
            th.emplace_back([&, k]() {
                for (uint32_t c = k; c < count; c += threads) {
                    op.forward(a + (size_t)c * n);
                    op.backward(a + (size_t)c * n);
                }
            });
        }
        for (auto& t : th) t.join();
        *wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    })
}

}  // extern "C"
