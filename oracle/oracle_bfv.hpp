// ORACLE — test infrastructure only (see oracle_core.hpp for the scope note).
//
// Ring layer (math/src/rq) and BFV layer (fhe/src/bfv) of the restatement.
// Polynomials are flat std::vector<u64> of shape [num_moduli][degree], row
// major, like the ndarray::Array2 in math/src/rq/mod.rs:125-132.
#pragma once

#include "oracle_core.hpp"

namespace oracle {

typedef std::vector<u64> PolyData;

// ---------------------------------------------------------------------------
// rq::Context — math/src/rq/context.rs:42-93
// ---------------------------------------------------------------------------
struct Context {
    std::vector<u64> moduli;
    std::vector<Modulus> q;
    std::shared_ptr<RnsContext> rns;
    std::vector<NttOperator> ops;
    size_t degree = 0;
    int logn = 0;
    std::vector<size_t> bitrev_;
    std::vector<u64> inv_last_qi_mod_qj, inv_last_qi_mod_qj_shoup;

    Context(const std::vector<u64>& m, size_t n) : moduli(m), degree(n) {
        if (n < 8 || (n & (n - 1))) throw std::runtime_error("bad degree");
        rns = std::make_shared<RnsContext>(m);
        for (u64 p : m) {
            q.emplace_back(p);
            ops.emplace_back(q.back(), n);
        }
        logn = ilog2(n);
        for (size_t j = 0; j < n; j++) bitrev_.push_back(bitrev(j, logn));
        u64 q_last = m.back();
        for (size_t i = 0; i + 1 < m.size(); i++) {
            u64 inv = q[i].inv(q[i].reduce(q_last));
            inv_last_qi_mod_qj.push_back(inv);
            inv_last_qi_mod_qj_shoup.push_back(q[i].shoup(inv));
        }
    }
    size_t L() const { return moduli.size(); }
    bool same(const Context& o) const { return moduli == o.moduli && degree == o.degree; }
};
typedef std::shared_ptr<Context> CtxPtr;

// rq/mod.rs:294-313
static inline void ntt_forward(const Context& c, u64* a) {
    for (size_t i = 0; i < c.L(); i++) c.ops[i].forward(a + i * c.degree);
}
static inline void ntt_backward(const Context& c, u64* a) {
    for (size_t i = 0; i < c.L(); i++) c.ops[i].backward(a + i * c.degree);
}

// rq/ops.rs:92-118, 136-160, 174-206 (exact modular pointwise ops)
static inline void poly_add(const Context& c, u64* a, const u64* b) {
    for (size_t i = 0; i < c.L(); i++)
        for (size_t j = 0; j < c.degree; j++) a[i * c.degree + j] = c.q[i].add(a[i * c.degree + j], b[i * c.degree + j]);
}
static inline void poly_sub(const Context& c, u64* a, const u64* b) {
    for (size_t i = 0; i < c.L(); i++)
        for (size_t j = 0; j < c.degree; j++) a[i * c.degree + j] = c.q[i].sub(a[i * c.degree + j], b[i * c.degree + j]);
}
static inline void poly_mul(const Context& c, u64* a, const u64* b) {
    for (size_t i = 0; i < c.L(); i++)
        for (size_t j = 0; j < c.degree; j++)
            a[i * c.degree + j] = c.q[i].mul_auto(a[i * c.degree + j], b[i * c.degree + j]);
}
// rq/ops.rs:297-350 : multiply by a big scalar projected on each modulus
static inline void poly_mul_big(const Context& c, u64* a, const Big& s) {
    std::vector<u64> crt = c.rns->project(s);
    for (size_t i = 0; i < c.L(); i++) {
        u64 b = crt[i], bs = c.q[i].shoup(b);
        for (size_t j = 0; j < c.degree; j++) a[i * c.degree + j] = c.q[i].mul_shoup(a[i * c.degree + j], b, bs);
    }
}

// rq/convert.rs:296-318 : signed coefficients -> power basis
static inline PolyData poly_from_i64(const Context& c, const std::vector<i64>& v) {
    PolyData p(c.L() * c.degree, 0);
    for (size_t i = 0; i < c.L(); i++)
        for (size_t j = 0; j < v.size(); j++) p[i * c.degree + j] = c.q[i].reduce_i64(v[j]);
    return p;
}
// rq/convert.rs:137-177 (short vector branch): unsigned values, reduced per modulus
static inline PolyData poly_from_u64(const Context& c, const std::vector<u64>& v) {
    PolyData p(c.L() * c.degree, 0);
    for (size_t i = 0; i < c.L(); i++)
        for (size_t j = 0; j < v.size(); j++) p[i * c.degree + j] = c.q[i].reduce(v[j]);
    return p;
}

// ---------------------------------------------------------------------------
// rq::Scaler — math/src/rq/scaler.rs:18-126
// ---------------------------------------------------------------------------
struct PolyScaler {
    CtxPtr from, to;
    size_t number_common_moduli = 0;
    std::unique_ptr<RnsScaler> scaler;

    PolyScaler(CtxPtr f, CtxPtr t, const ScalingFactor& factor) : from(f), to(t) {
        if (f->degree != t->degree) throw std::runtime_error("Incompatible degrees");
        if (factor.is_one) {
            size_t k = 0;
            while (k < f->L() && k < t->L() && f->moduli[k] == t->moduli[k]) k++;
            number_common_moduli = k;
        }
        scaler.reset(new RnsScaler(f->rns.get(), t->rns.get(), factor));
    }
    // in: from-context poly ; ntt_repr selects the Ntt branch (INTT in, NTT out)
    PolyData scale(const u64* in, bool ntt_repr) const {
        size_t n = from->degree, lf = from->L(), lt = to->L();
        PolyData out(lt * n, 0);
        size_t c = number_common_moduli;
        if (c > 0) std::copy(in, in + c * n, out.begin());
        if (c < lt) {
            PolyData pb(in, in + lf * n);
            if (ntt_repr) ntt_backward(*from, pb.data());
            std::vector<u64> col(lf);
            for (size_t j = 0; j < n; j++) {
                for (size_t i = 0; i < lf; i++) col[i] = pb[i * n + j];
                scaler->scale(col.data(), out.data() + c * n + j, lt - c, c, n);
            }
            if (ntt_repr)
                for (size_t i = c; i < lt; i++) to->ops[i].forward(out.data() + i * n);
        }
        return out;
    }
};

// ---------------------------------------------------------------------------
// Poly::switch_down — math/src/rq/mod.rs:390-449 (power basis; drops last limb)
// ---------------------------------------------------------------------------
static inline PolyData switch_down(const Context& c, const u64* in) {
    size_t n = c.degree, L = c.L();
    if (L < 2) throw std::runtime_error("NoMoreContext");
    const Modulus& ql = c.q[L - 1];
    u64 half = ql.p / 2;
    std::vector<u64> last(in + (L - 1) * n, in + L * n);
    for (auto& x : last) x = ql.add(x, half);
    PolyData out(in, in + (L - 1) * n);
    for (size_t i = 0; i + 1 < L; i++) {
        const Modulus& qi = c.q[i];
        u64 half_mod = qi.p - qi.reduce(half);
        for (size_t j = 0; j < n; j++) {
            u64 tmp = qi.lazy_reduce(last[j]) + half_mod;
            u64 v = out[i * n + j] + 3 * qi.p - tmp;
            out[i * n + j] = qi.mul_shoup(v, c.inv_last_qi_mod_qj[i], c.inv_last_qi_mod_qj_shoup[i]);
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
// SubstitutionExponent / Poly::substitute — math/src/rq/mod.rs:88-121, 319-369
// ---------------------------------------------------------------------------
static inline std::vector<size_t> power_bitrev(const Context& c, size_t exponent) {
    exponent %= 2 * c.degree;
    if ((exponent & 1) == 0) throw std::runtime_error("The exponent should be odd modulo 2 * degree");
    size_t power = (exponent - 1) / 2, mask = c.degree - 1;
    std::vector<size_t> pb(c.degree);
    for (size_t k = 0; k < c.degree; k++) {
        pb[k] = bitrev(power & mask, c.logn);
        power += exponent;
    }
    return pb;
}
static inline PolyData substitute(const Context& c, size_t exponent, const u64* in, bool ntt_repr) {
    size_t n = c.degree;
    PolyData out(c.L() * n, 0);
    if (ntt_repr) {
        std::vector<size_t> pb = power_bitrev(c, exponent);
        for (size_t i = 0; i < c.L(); i++)
            for (size_t j = 0; j < n; j++) out[i * n + c.bitrev_[j]] = in[i * n + pb[j]];
    } else {
        exponent %= 2 * n;
        size_t power = 0, mask = n - 1;
        for (size_t j = 0; j < n; j++) {
            for (size_t i = 0; i < c.L(); i++) {
                u64& d = out[i * n + (power & mask)];
                d = (power & n) ? c.q[i].sub(d, in[i * n + j]) : c.q[i].add(d, in[i * n + j]);
            }
            power += exponent;
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
// Seeded generator for the harness (fhe.rs draws from rand::rng(); its streams
// are not reproducible, so every test input here comes from this generator).
// ---------------------------------------------------------------------------
struct Rng {
    u64 s;
    explicit Rng(u64 seed) : s(seed) {}
    u64 next() {  // splitmix64
        u64 z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    u64 uniform(u64 p) {
        u64 lim = (~0ULL) - ((~0ULL) % p);
        u64 x;
        do { x = next(); } while (x >= lim);
        return x % p;
    }
};

// fhe-util/src/lib.rs:22-55 (same pooling of 64-bit draws)
static inline std::vector<i64> sample_vec_cbd(size_t n, size_t variance, Rng& rng) {
    if (variance < 1 || variance > 16) throw std::runtime_error("bad variance");
    std::vector<i64> out;
    size_t number_bits = 4 * variance;
    u128 mask_add = (u128)((~0ULL >> (64 - number_bits)) >> (2 * variance));
    u128 mask_sub = mask_add << (2 * variance);
    u128 pool = 0;
    size_t nbits = 0;
    for (size_t i = 0; i < n; i++) {
        if (nbits < number_bits) {
            pool |= (u128)rng.next() << nbits;
            nbits += 64;
        }
        u128 a = pool & mask_add, s = pool & mask_sub;
        int ca = __builtin_popcountll((u64)a) + __builtin_popcountll((u64)(a >> 64));
        int cs = __builtin_popcountll((u64)s) + __builtin_popcountll((u64)(s >> 64));
        out.push_back((i64)ca - (i64)cs);
        pool >>= number_bits;
        nbits -= number_bits;
    }
    return out;
}

static inline PolyData random_poly(const Context& c, Rng& rng) {
    PolyData p(c.L() * c.degree);
    for (size_t i = 0; i < c.L(); i++)
        for (size_t j = 0; j < c.degree; j++) p[i * c.degree + j] = rng.uniform(c.moduli[i]);
    return p;
}

// ---------------------------------------------------------------------------
// BfvParameters — fhe/src/bfv/parameters.rs:383-641
// ---------------------------------------------------------------------------
static inline std::vector<u64> generate_moduli(const std::vector<size_t>& sizes, size_t degree) {
    std::vector<u64> moduli;
    for (size_t s : sizes) {
        if (s > 62 || s < 10) throw std::runtime_error("InvalidModulusSize");
        u64 ub = 1ULL << s;
        for (;;) {
            u64 p = generate_prime(s, 2 * (u64)degree, ub);
            if (!p) throw std::runtime_error("NotEnoughPrimes");
            if (std::find(moduli.begin(), moduli.end(), p) == moduli.end()) {
                moduli.push_back(p);
                break;
            }
            ub = p;
        }
    }
    return moduli;
}

struct MulParams {
    CtxPtr from, to;
    std::unique_ptr<PolyScaler> extender, down_scaler;
};

struct Level {
    CtxPtr ctx;
    std::vector<u64> delta_rests;  // (-t)^-1 mod q_i
    u64 q_mod_t = 0;
    std::unique_ptr<PolyScaler> decrypt_scaler;  // ctx -> plaintext ctx, t/Q
    MulParams mp;
};

struct BfvParams {
    size_t degree = 0;
    u64 t = 0;
    size_t variance = 10;
    std::vector<u64> moduli;
    std::vector<size_t> moduli_sizes;
    CtxPtr plaintext_ctx;
    std::unique_ptr<NttOperator> t_op;
    Modulus t_mod;
    std::vector<size_t> index_map;
    std::vector<Level> levels;

    BfvParams(size_t n, u64 t_, const std::vector<u64>& m, size_t var) : degree(n), t(t_), variance(var), moduli(m) {
        for (u64 q : m) moduli_sizes.push_back(64 - lzcnt64(q));
        // plaintext context: enough moduli to exceed t_bits + 60 (parameters.rs:479-494)
        size_t t_bits = 64 - lzcnt64(t), acc = 0, count = 0;
        for (size_t s : moduli_sizes) {
            acc += s;
            count++;
            if (acc >= t_bits + 60) break;
        }
        count = std::max<size_t>(count, 1);
        count = std::min(count, m.size());
        plaintext_ctx = std::make_shared<Context>(std::vector<u64>(m.begin(), m.begin() + count), n);
        t_mod = Modulus(t);
        if (supports_ntt(t, n)) t_op.reset(new NttOperator(t_mod, n));
        // extended basis (parameters.rs:585-593)
        std::vector<u64> ext;
        u64 ub = 1ULL << 62;
        while (ext.size() != m.size() + 1) {
            ub = generate_prime(62, 2 * (u64)n, ub);
            if (std::find(ext.begin(), ext.end(), ub) == ext.end() &&
                std::find(m.begin(), m.end(), ub) == m.end())
                ext.push_back(ub);
        }
        levels.resize(m.size());
        for (size_t lvl = 0; lvl < m.size(); lvl++) {
            Level& L = levels[lvl];
            std::vector<u64> lm(m.begin(), m.end() - lvl);
            L.ctx = std::make_shared<Context>(lm, n);
            for (u64 q : lm) {
                Modulus qm(q);
                L.delta_rests.push_back(qm.inv(qm.neg(t % q)));
            }
            L.q_mod_t = L.ctx->rns->product.mod_u64(t);
            L.decrypt_scaler.reset(new PolyScaler(L.ctx, plaintext_ctx, ScalingFactor(Big(t), L.ctx->rns->product)));
            size_t msize = 0;
            for (size_t i = 0; i < lm.size(); i++) msize += moduli_sizes[i];
            size_t n_moduli = (msize + 60 + 61) / 62;
            std::vector<u64> mm = lm;
            mm.insert(mm.end(), ext.begin(), ext.begin() + n_moduli);
            L.mp.from = L.ctx;
            L.mp.to = std::make_shared<Context>(mm, n);
            L.mp.extender.reset(new PolyScaler(L.ctx, L.mp.to, ScalingFactor()));
            L.mp.down_scaler.reset(new PolyScaler(L.mp.to, L.ctx, ScalingFactor(Big(t), L.ctx->rns->product)));
        }
        // SIMD index map (parameters.rs:614-629)
        index_map.assign(n, 0);
        size_t row = n >> 1, mm2 = n << 1, pos = 1;
        int logn = ilog2(n);
        for (size_t i = 0; i < row; i++) {
            size_t i1 = (pos - 1) >> 1, i2 = (mm2 - pos - 1) >> 1;
            index_map[i] = bitrev(i1, logn);
            index_map[row | i] = bitrev(i2, logn);
            pos *= 3;
            pos &= mm2 - 1;
        }
    }
    const Context& ctx(size_t level) const { return *levels.at(level).ctx; }
};

// ---------------------------------------------------------------------------
// Plaintext encode/decode (harness) — plaintext_vec.rs:37-101, plaintext.rs:71-98,
// secret_key.rs:200-282 and plaintext.rs:310-330
// ---------------------------------------------------------------------------
static inline std::vector<u64> encode_simd(const BfvParams& par, const std::vector<u64>& slots) {
    if (!par.t_op) throw std::invalid_argument("the plaintext modulus does not support SIMD encoding at this degree");
    std::vector<u64> v(par.degree, 0);
    for (size_t i = 0; i < slots.size(); i++) v[par.index_map[i]] = slots[i];
    par.t_op->backward(v.data());
    return v;
}
static inline std::vector<u64> decode_simd(const BfvParams& par, std::vector<u64> w) {
    if (!par.t_op) throw std::invalid_argument("the plaintext modulus does not support SIMD encoding at this degree");
    par.t_op->forward(w.data());
    std::vector<u64> out(par.degree);
    for (size_t i = 0; i < par.degree; i++) out[i] = w[par.index_map[i]];
    return out;
}
// Plaintext::to_poly : NTT-domain Delta * m at the given level
static inline PolyData plaintext_to_poly(const BfvParams& par, size_t level, const std::vector<u64>& v) {
    const Level& L = par.levels[level];
    std::vector<u64> m = v;
    u64 b = L.q_mod_t, bs = par.t_mod.shoup(b);
    for (auto& x : m) x = par.t_mod.mul_shoup(x, b, bs);
    PolyData p = poly_from_u64(*L.ctx, m);
    ntt_forward(*L.ctx, p.data());
    size_t n = par.degree;
    for (size_t i = 0; i < L.ctx->L(); i++) {
        u64 d = L.delta_rests[i], ds = L.ctx->q[i].shoup(d);
        for (size_t j = 0; j < n; j++) p[i * n + j] = L.ctx->q[i].mul_shoup(p[i * n + j], d, ds);
    }
    return p;
}

struct SecretKey {
    std::vector<i64> coeffs;
};

// SecretKey::encrypt_poly (secret_key.rs:102-137) of the NTT-domain plaintext
// polynomial pt with the given uniform NTT-domain `a` and error coefficients
// `e`: [e - a s + pt, a].
static inline PolyData encrypt_with(const BfvParams& par, const SecretKey& sk, size_t level, const u64* pt,
                                    const u64* a, const std::vector<i64>& e) {
    const Context& c = par.ctx(level);
    size_t sz = c.L() * c.degree;
    PolyData s = poly_from_i64(c, sk.coeffs);
    ntt_forward(c, s.data());
    PolyData as(a, a + sz);
    poly_mul(c, as.data(), s.data());
    PolyData b = poly_from_i64(c, e);
    ntt_forward(c, b.data());
    poly_sub(c, b.data(), as.data());
    poly_add(c, b.data(), pt);
    PolyData ct(2 * sz);
    std::copy(b.begin(), b.end(), ct.begin());
    std::copy(a, a + sz, ct.begin() + sz);
    return ct;
}

// PublicKey::try_encrypt (public_key.rs:47-87) of the NTT-domain plaintext
// polynomial pt with the public key pk [2][L][N] at the same level and the
// given small coefficients u, e1, e2: [u pk0 + e1 + pt, u pk1 + e2].
static inline PolyData encrypt_pk_with(const BfvParams& par, size_t level, const u64* pk, const u64* pt,
                                       const std::vector<i64>& u, const std::vector<i64>& e1,
                                       const std::vector<i64>& e2) {
    const Context& c = par.ctx(level);
    size_t sz = c.L() * c.degree;
    PolyData un = poly_from_i64(c, u), e1n = poly_from_i64(c, e1), e2n = poly_from_i64(c, e2);
    ntt_forward(c, un.data());
    ntt_forward(c, e1n.data());
    ntt_forward(c, e2n.data());
    PolyData c0(pk, pk + sz), c1(pk + sz, pk + 2 * sz);
    poly_mul(c, c0.data(), un.data());
    poly_add(c, c0.data(), e1n.data());
    poly_add(c, c0.data(), pt);
    poly_mul(c, c1.data(), un.data());
    poly_add(c, c1.data(), e2n.data());
    PolyData ct(2 * sz);
    std::copy(c0.begin(), c0.end(), ct.begin());
    std::copy(c1.begin(), c1.end(), ct.begin() + sz);
    return ct;
}

// SecretKey::encrypt_poly (secret_key.rs:102-137) with `a` and `e` drawn from the
// seeded generator. Returns [b, a] in NTT domain.
// `poly`: Encoding::poly_at_level (the values are the plaintext's
// coefficients, plaintext.rs:71-98) instead of Encoding::simd.
static inline PolyData encrypt(const BfvParams& par, const SecretKey& sk, size_t level,
                               const std::vector<u64>& slots, Rng& rng, bool poly = false) {
    const Context& c = par.ctx(level);
    size_t sz = c.L() * c.degree;
    std::vector<u64> v = slots;
    if (poly) v.resize(par.degree, 0);
    PolyData pt = plaintext_to_poly(par, level, poly ? v : encode_simd(par, slots));
    PolyData s = poly_from_i64(c, sk.coeffs);
    ntt_forward(c, s.data());
    PolyData a = random_poly(c, rng);
    PolyData as = a;
    poly_mul(c, as.data(), s.data());
    PolyData b = poly_from_i64(c, sample_vec_cbd(c.degree, par.variance, rng));
    ntt_forward(c, b.data());
    poly_sub(c, b.data(), as.data());
    poly_add(c, b.data(), pt.data());
    PolyData ct(2 * sz);
    std::copy(b.begin(), b.end(), ct.begin());
    std::copy(a.begin(), a.end(), ct.begin() + sz);
    return ct;
}

// FheDecrypter::try_decrypt (secret_key.rs:200-282) followed by SIMD decode, or
// by none with `poly` (Encoding::poly_at_level).
static inline std::vector<u64> decrypt(const BfvParams& par, const SecretKey& sk, size_t level,
                                       const u64* ct, size_t nparts, bool poly = false) {
    const Level& L = par.levels[level];
    const Context& c = *L.ctx;
    size_t sz = c.L() * c.degree, n = c.degree;
    PolyData s = poly_from_i64(c, sk.coeffs);
    ntt_forward(c, s.data());
    PolyData si = s;
    PolyData acc(ct, ct + sz);
    for (size_t i = 1; i < nparts; i++) {
        PolyData cis(ct + i * sz, ct + (i + 1) * sz);
        poly_mul(c, cis.data(), si.data());
        poly_add(c, acc.data(), cis.data());
        if (i + 1 < nparts) poly_mul(c, si.data(), s.data());
    }
    ntt_backward(c, acc.data());
    PolyData d = L.decrypt_scaler->scale(acc.data(), false);
    Modulus q0(par.moduli[0]);
    std::vector<u64> w(n);
    for (size_t j = 0; j < n; j++) w[j] = par.t_mod.reduce(q0.reduce(d[j] + par.t));
    return poly ? w : decode_simd(par, w);
}

// ---------------------------------------------------------------------------
// KeySwitchingKey — fhe/src/bfv/keys/key_switching_key.rs:22-289
// c0/c1 are NttShoup polynomials in the key context; the Shoup companions are
// recomputed where needed.
// ---------------------------------------------------------------------------
// With a single-modulus key context the key uses the power-of-two
// decomposition (key_switching_key.rs:70-86, 292-329): digits of log_base =
// log_modulus / 2 bits, log_modulus = log2(next_power_of_two(q)).
static inline size_t decomposition_log_modulus(u64 q) {
    size_t b = 0;
    while (b < 64 && (((u64)1 << b) < q)) b++;
    return b;
}

struct KeySwitchingKey {
    CtxPtr ctx_ct, ctx_ksk;
    size_t ciphertext_level = 0, ksk_level = 0;
    size_t log_base = 0;  // power-of-two decomposition (one-modulus key context), else 0
    std::vector<PolyData> c0, c1, c0s, c1s;  // values and Shoup companions

    // key_switching_key.rs:50-105 and 126-210
    KeySwitchingKey(const BfvParams& par, const SecretKey& sk, const PolyData& from_pb,
                    size_t ct_level, size_t key_level, Rng& rng)
        : ciphertext_level(ct_level), ksk_level(key_level) {
        ctx_ksk = par.levels[key_level].ctx;
        ctx_ct = par.levels[ct_level].ctx;
        const Context& k = *ctx_ksk;
        if (k.L() == 1) {
            // generate_c0_decomposition (key_switching_key.rs:172-210): c0_i =
            // e - c1_i s + 2^(i log_base) from
            const size_t log_modulus = decomposition_log_modulus(k.moduli[0]);
            log_base = log_modulus / 2;
            const size_t size = (log_modulus + log_base - 1) / log_base, n = k.degree;
            PolyData s = poly_from_i64(k, sk.coeffs);
            ntt_forward(k, s.data());
            for (size_t i = 0; i < size; i++) {
                PolyData c1i = random_poly(k, rng);
                PolyData as = c1i;
                poly_mul(k, as.data(), s.data());
                ntt_backward(k, as.data());
                PolyData b = poly_from_i64(k, sample_vec_cbd(n, par.variance, rng));
                poly_sub(k, b.data(), as.data());
                PolyData gf = from_pb;
                poly_mul_big(k, gf.data(), Big((u64)1 << (i * log_base)));  // i log_base < 62
                poly_add(k, b.data(), gf.data());
                ntt_forward(k, b.data());
                c0.push_back(b);
                c1.push_back(c1i);
            }
            shoups();
            return;
        }
        size_t size = ctx_ct->L(), n = k.degree;
        PolyData s = poly_from_i64(k, sk.coeffs);
        ntt_forward(k, s.data());
        RnsContext rns(std::vector<u64>(par.moduli.begin(), par.moduli.begin() + size));
        for (size_t i = 0; i < size; i++) {
            PolyData c1i = random_poly(k, rng);  // NTT domain, uniform
            PolyData as = c1i;
            poly_mul(k, as.data(), s.data());
            ntt_backward(k, as.data());
            PolyData b = poly_from_i64(k, sample_vec_cbd(n, par.variance, rng));
            poly_sub(k, b.data(), as.data());
            PolyData gf = from_pb;
            poly_mul_big(k, gf.data(), rns.garner[i]);
            poly_add(k, b.data(), gf.data());
            ntt_forward(k, b.data());
            c0.push_back(b);
            c1.push_back(c1i);
        }
        shoups();
    }

    void shoups() {
        const Context& k = *ctx_ksk;
        const size_t n = k.degree, sz = k.L() * n;
        for (size_t i = 0; i < c0.size(); i++) {
            PolyData a(sz), b(sz);
            for (size_t l = 0; l < k.L(); l++)
                for (size_t j = 0; j < n; j++) {
                    a[l * n + j] = k.q[l].shoup(c0[i][l * n + j]);
                    b[l * n + j] = k.q[l].shoup(c1[i][l * n + j]);
                }
            c0s.push_back(a);
            c1s.push_back(b);
        }
    }

    // Number of key rows: one per ciphertext modulus, or the decomposition's
    // ceil(log_modulus / log_base) digits.
    size_t digits() const {
        if (!log_base) return ctx_ct->L();
        const size_t lm = decomposition_log_modulus(ctx_ksk->moduli[0]);
        return (lm + log_base - 1) / log_base;
    }

    // A key given by its NTT-domain rows c0, c1 [digits][L_key][N].
    KeySwitchingKey(const BfvParams& par, size_t ct_level, size_t key_level, const u64* r0, const u64* r1)
        : ciphertext_level(ct_level), ksk_level(key_level) {
        ctx_ksk = par.levels[key_level].ctx;
        ctx_ct = par.levels[ct_level].ctx;
        const Context& k = *ctx_ksk;
        size_t sz = k.L() * k.degree;
        if (k.L() == 1) log_base = decomposition_log_modulus(k.moduli[0]) / 2;
        for (size_t i = 0; i < digits(); i++) {
            c0.emplace_back(r0 + i * sz, r0 + (i + 1) * sz);
            c1.emplace_back(r1 + i * sz, r1 + (i + 1) * sz);
        }
        shoups();
    }

    // key_switching_key.rs:214-241 (and key_switch_decomposition, :292-329,
    // with a one-modulus key context); p is power basis in ctx_ct. Returns
    // NTT polys.
    void key_switch(const u64* p, PolyData& out0, PolyData& out1) const {
        const Context& k = *ctx_ksk;
        size_t n = k.degree, sz = k.L() * n;
        out0.assign(sz, 0);
        out1.assign(sz, 0);
        PolyData lift(sz);
        const u64 mask = log_base ? ((u64)1 << log_base) - 1 : 0;
        for (size_t i = 0; i < digits(); i++) {
            // rq/mod.rs:518-540 digit lift: copy, lazy reduce, forward_vt_lazy;
            // a decomposition digit is bits [i log_base, (i + 1) log_base) of p
            for (size_t l = 0; l < k.L(); l++) {
                u64* row = lift.data() + l * n;
                for (size_t j = 0; j < n; j++) {
                    const u64 d = log_base ? (p[j] >> (i * log_base)) & mask : p[i * n + j];
                    row[j] = k.q[l].lazy_reduce_vec_elem(d);
                }
                k.ops[l].forward_lazy(row);
            }
            for (size_t l = 0; l < k.L(); l++) {
                const Modulus& q = k.q[l];
                for (size_t j = 0; j < n; j++) {
                    size_t x = l * n + j;
                    out0[x] = q.add(out0[x], q.mul_shoup(lift[x], c0[i][x], c0s[i][x]));
                    out1[x] = q.add(out1[x], q.mul_shoup(lift[x], c1[i][x], c1s[i][x]));
                }
            }
        }
    }
};

// After a key switch in a context above the ciphertext's: INTT, switch_down_to,
// NTT (mul.rs:199-206, relinearization_key.rs:84-91, galois_key.rs:70-77).
static inline PolyData bring_down(const BfvParams& par, size_t key_level, size_t ct_level, PolyData p) {
    for (size_t lvl = key_level; lvl < ct_level; lvl++) {
        const Context& c = par.ctx(lvl);
        if (lvl == key_level) ntt_backward(c, p.data());
        p = switch_down(c, p.data());
    }
    if (key_level != ct_level) ntt_forward(par.ctx(ct_level), p.data());
    return p;
}

// RelinearizationKey::new_leveled_internal (relinearization_key.rs:43-67)
static inline KeySwitchingKey* relin_key_new(const BfvParams& par, const SecretKey& sk, size_t ct_level,
                                             size_t key_level, Rng& rng) {
    const Level& Lc = par.levels[ct_level];
    const Level& Lk = par.levels[key_level];
    if (Lk.ctx->L() == 1) throw std::runtime_error("These parameters do not support key switching");
    const Context& c = *Lc.ctx;
    PolyData s = poly_from_i64(c, sk.coeffs);
    ntt_forward(c, s.data());
    PolyData s2 = s;
    poly_mul(c, s2.data(), s.data());
    ntt_backward(c, s2.data());
    PolyScaler up(Lc.ctx, Lk.ctx, ScalingFactor(Lk.ctx->rns->product, Lc.ctx->rns->product));
    PolyData s2up = up.scale(s2.data(), false);
    return new KeySwitchingKey(par, sk, s2up, ct_level, key_level, rng);
}

// GaloisKey::new (galois_key.rs:26-60)
static inline KeySwitchingKey* galois_key_new(const BfvParams& par, const SecretKey& sk, size_t exponent,
                                              size_t ct_level, size_t key_level, Rng& rng) {
    const Level& Lc = par.levels[ct_level];
    const Level& Lk = par.levels[key_level];
    const Context& c = *Lc.ctx;
    PolyData s = poly_from_i64(c, sk.coeffs);
    PolyData ss = substitute(c, exponent, s.data(), false);
    PolyScaler up(Lc.ctx, Lk.ctx, ScalingFactor(Lk.ctx->rns->product, Lc.ctx->rns->product));
    PolyData sup = up.scale(ss.data(), false);
    return new KeySwitchingKey(par, sk, sup, ct_level, key_level, rng);
}

// GaloisKey::relinearize_into (galois_key.rs:90-124). ct: [2][L][N] NTT.
static inline PolyData galois_apply(const BfvParams& par, const KeySwitchingKey& gk, size_t exponent,
                                    const u64* ct) {
    const Context& c = *gk.ctx_ct;
    size_t sz = c.L() * c.degree;
    PolyData c2 = substitute(c, exponent, ct + sz, true);
    ntt_backward(c, c2.data());
    PolyData o0, o1;
    gk.key_switch(c2.data(), o0, o1);
    o0 = bring_down(par, gk.ksk_level, gk.ciphertext_level, o0);
    o1 = bring_down(par, gk.ksk_level, gk.ciphertext_level, o1);
    PolyData s0 = substitute(c, exponent, ct, true);
    poly_add(c, o0.data(), s0.data());
    PolyData out(2 * sz);
    std::copy(o0.begin(), o0.end(), out.begin());
    std::copy(o1.begin(), o1.end(), out.begin() + sz);
    return out;
}

// ---------------------------------------------------------------------------
// Plaintext-ciphertext dot product and oblivious expansion (the PIR server path)
// ---------------------------------------------------------------------------
// dot_product_scalar (fhe/src/bfv/ops/dot_product.rs:56-143). cts: `count`
// ciphertexts [parts][L][N] every `ct_stride` words; pts: `count` NTT-domain
// plaintext polynomials [L][N] (Plaintext::poly_ntt) every `pt_stride` words.
// Up to min_q 2^(2 lz(q)) terms one u128 accumulator per coefficient takes the
// exact sum (dot_product.rs:110-127); above that each part goes through
// rq::dot_product (math/src/rq/ops.rs:448-550), which reduces a limb's
// accumulators whenever its count of terms reaches 2^(2 lz(q)).
static inline PolyData dot_product_scalar(const Context& c, const u64* cts, size_t ct_stride, const u64* pts,
                                          size_t pt_stride, size_t count, size_t parts) {
    if (count == 0) throw std::runtime_error("At least one iterator is empty");
    const size_t n = c.degree, L = c.L();
    std::vector<u128> max_acc(L);
    u128 min_of_max = ~(u128)0;
    for (size_t j = 0; j < L; j++) {
        max_acc[j] = (u128)1 << (2 * lzcnt64(c.moduli[j]));
        min_of_max = std::min(min_of_max, max_acc[j]);
    }
    PolyData out(parts * L * n);
    std::vector<u128> acc(L * n);
    if ((u128)count > min_of_max) {
        for (size_t part = 0; part < parts; part++) {
            std::fill(acc.begin(), acc.end(), 0);
            std::vector<u128> num_acc(L, 1);
            for (size_t i = 0; i < count; i++) {
                const u64* x = cts + i * ct_stride + part * L * n;
                const u64* y = pts + i * pt_stride;
                for (size_t k = 0; k < L * n; k++) acc[k] += (u128)x[k] * y[k];
                for (size_t j = 0; j < L; j++) {
                    if (++num_acc[j] == max_acc[j]) {
                        for (size_t k = j * n; k < (j + 1) * n; k++) acc[k] = c.q[j].reduce_u128(acc[k]);
                        num_acc[j] = 1;
                    }
                }
            }
            for (size_t j = 0; j < L; j++)
                for (size_t k = j * n; k < (j + 1) * n; k++) out[part * L * n + k] = c.q[j].reduce_u128(acc[k]);
        }
    } else {
        for (size_t part = 0; part < parts; part++) {
            std::fill(acc.begin(), acc.end(), 0);
            for (size_t i = 0; i < count; i++) {
                const u64* x = cts + i * ct_stride + part * L * n;
                const u64* y = pts + i * pt_stride;
                for (size_t k = 0; k < L * n; k++) acc[k] += (u128)x[k] * y[k];
            }
            for (size_t j = 0; j < L; j++)
                for (size_t k = j * n; k < (j + 1) * n; k++) out[part * L * n + k] = c.q[j].reduce_u128(acc[k]);
        }
    }
    return out;
}

// The expansion monomial of level l: -X^(N - 2^l) in the NTT domain
// (EvaluationKeyBuilder::build, fhe/src/bfv/keys/evaluation_key.rs:355-363).
static inline PolyData expansion_monomial(const Context& c, size_t l) {
    std::vector<i64> m(c.degree, 0);
    m[c.degree - ((size_t)1 << l)] = -1;
    PolyData p = poly_from_i64(c, m);
    ntt_forward(c, p.data());
    return p;
}

// EvaluationKey::expands (fhe/src/bfv/keys/evaluation_key.rs:153-193): the
// oblivious expansion of a [2][L][N] ciphertext into `size` ciphertexts.
// `gk(l)` returns the Galois key of exponent (N >> l) + 1 (or null).
template <class KeyOf>
static inline std::vector<PolyData> expands(const BfvParams& par, const Context& c, KeyOf gk, const u64* ct,
                                            size_t size) {
    size_t level = 0;
    while (((size_t)1 << level) < size) level++;
    const size_t sz = c.L() * c.degree;
    if (level == 0) return {PolyData(ct, ct + 2 * sz)};
    for (size_t l = 0; l < level; l++)
        if (!gk(l)) throw std::runtime_error("This key does not support expansion at this level");
    std::vector<PolyData> out((size_t)1 << level, PolyData(2 * sz, 0));
    std::copy(ct, ct + 2 * sz, out[0].begin());
    for (size_t l = 0; l < level; l++) {
        const PolyData mono = expansion_monomial(c, l);
        const size_t step = (size_t)1 << l, exponent = (c.degree >> l) + 1;
        for (size_t i = 0; i < step; i++) {
            PolyData sub = galois_apply(par, *gk(l), exponent, out[i].data());
            const size_t j = step | i;
            if (j < size) {
                PolyData t = out[i];
                poly_sub(c, t.data(), sub.data());
                poly_sub(c, t.data() + sz, sub.data() + sz);
                poly_mul(c, t.data(), mono.data());
                poly_mul(c, t.data() + sz, mono.data());
                out[j] = t;
            }
            poly_add(c, out[i].data(), sub.data());
            poly_add(c, out[i].data() + sz, sub.data() + sz);
        }
    }
    out.resize(size);
    return out;
}

// ---------------------------------------------------------------------------
// Wire format of Rq coefficients (math/src/rq/convert.rs:15-39, :42-87)
// ---------------------------------------------------------------------------
// transcode_to_bytes / transcode_from_bytes (fhe-util/src/lib.rs:60-133):
// values of nbits bits packed least significant bit first.
static inline std::vector<uint8_t> transcode_to_bytes(const u64* a, size_t n, size_t nbits) {
    const u128 mask = (u128)(~0ULL >> (64 - nbits));
    std::vector<uint8_t> out;
    u128 cur = 0;
    size_t have = 0, i = 0;
    while (i < n) {
        if (have < 8) {
            cur |= ((u128)a[i] & mask) << have;
            have += nbits;
            i++;
        }
        while (have >= 8) {
            out.push_back((uint8_t)cur);
            cur >>= 8;
            have -= 8;
        }
    }
    if (have > 0) out.push_back((uint8_t)cur);
    return out;
}
static inline std::vector<u64> transcode_from_bytes(const uint8_t* b, size_t nb, size_t nbits) {
    const u128 mask = (u128)(~0ULL >> (64 - nbits));
    std::vector<u64> out;
    u128 cur = 0;
    size_t have = 0, i = 0;
    while (i < nb) {
        if (have < nbits) {
            cur |= (u128)b[i] << have;
            have += 8;
            i++;
        }
        while (have >= nbits) {
            out.push_back((u64)(cur & mask));
            cur >>= nbits;
            have -= nbits;
        }
    }
    if (have > 0) out.push_back((u64)cur);
    return out;
}
// Modulus::serialization_length's bit width (zq/mod.rs:773-793): bits of p - 1.
static inline size_t coeff_bits(u64 p) { return (size_t)(64 - lzcnt64(p - 1)); }

// The Rq coefficients payload of a polynomial [L][N] (NTT polynomials are
// serialized in power basis): each limb's coefficients packed with its width.
static inline std::vector<uint8_t> serialize_poly(const Context& c, const u64* p, bool ntt) {
    PolyData pb(p, p + c.L() * c.degree);
    if (ntt) ntt_backward(c, pb.data());
    std::vector<uint8_t> out;
    for (size_t i = 0; i < c.L(); i++) {
        std::vector<uint8_t> v = transcode_to_bytes(pb.data() + i * c.degree, c.degree, coeff_bits(c.moduli[i]));
        out.insert(out.end(), v.begin(), v.end());
    }
    return out;
}
static inline PolyData deserialize_poly(const Context& c, const uint8_t* b, bool ntt) {
    PolyData p;
    for (size_t i = 0; i < c.L(); i++) {
        const size_t nbits = coeff_bits(c.moduli[i]), size = nbits * c.degree / 8;
        std::vector<u64> v = transcode_from_bytes(b, size, nbits);
        p.insert(p.end(), v.begin(), v.end());
        b += size;
    }
    if (ntt) ntt_forward(c, p.data());
    return p;
}

// ---------------------------------------------------------------------------
// RGSWCiphertext (fhe/src/bfv/rgsw_ciphertext.rs)
// ---------------------------------------------------------------------------
// SecretKey::try_encrypt into an RGSW ciphertext (rgsw_ciphertext.rs:97-117):
// key-switching keys at the plaintext's level from m (which == 0) or from m s
// (which == 1), m = the plaintext's poly_ntt [L][N].
static inline KeySwitchingKey* rgsw_key_new(const BfvParams& par, const SecretKey& sk, size_t level,
                                            const u64* pt_ntt, int which, Rng& rng) {
    const Context& c = par.ctx(level);
    PolyData from(pt_ntt, pt_ntt + c.L() * c.degree);
    if (which) {
        PolyData s = poly_from_i64(c, sk.coeffs);
        ntt_forward(c, s.data());
        poly_mul(c, from.data(), s.data());
    }
    ntt_backward(c, from.data());
    return new KeySwitchingKey(par, sk, from, level, level, rng);
}

// Ciphertext x RGSWCiphertext (rgsw_ciphertext.rs:123-160): both parts to
// power basis, key switched by ksk0 and ksk1, the results added.
static inline PolyData rgsw_external_product(const KeySwitchingKey& k0, const KeySwitchingKey& k1, const u64* ct) {
    const Context& c = *k0.ctx_ct;
    size_t sz = c.L() * c.degree;
    PolyData c0(ct, ct + sz), c1(ct + sz, ct + 2 * sz);
    ntt_backward(c, c0.data());
    ntt_backward(c, c1.data());
    PolyData a0, a1, b0, b1;
    k0.key_switch(c0.data(), a0, a1);
    k1.key_switch(c1.data(), b0, b1);
    const Context& k = *k0.ctx_ksk;
    poly_add(k, a0.data(), b0.data());
    poly_add(k, a1.data(), b1.data());
    PolyData out(a0);
    out.insert(out.end(), a1.begin(), a1.end());
    return out;
}

// ---------------------------------------------------------------------------
// Multiplicator — fhe/src/bfv/ops/mul.rs:22-227
// ---------------------------------------------------------------------------
struct Multiplicator {
    const BfvParams* par;
    const KeySwitchingKey* rk;  // may be null
    size_t level;
    CtxPtr base_ctx, mul_ctx;
    std::unique_ptr<PolyScaler> extender, extender_rhs, down_scaler;  // extender_rhs null: same as lhs

    // Multiplicator::new_leveled (mul.rs:37-98): an explicit extended basis and
    // the lhs / rhs extension and post-multiplication scaling factors.
    Multiplicator(const BfvParams& p, const KeySwitchingKey* rk_, size_t lvl, const std::vector<u64>& ext,
                  const ScalingFactor& lhs, const ScalingFactor& rhs, const ScalingFactor& post)
        : par(&p), rk(rk_), level(lvl) {
        base_ctx = p.levels[lvl].ctx;
        mul_ctx = std::make_shared<Context>(ext, p.degree);
        extender.reset(new PolyScaler(base_ctx, mul_ctx, lhs));
        extender_rhs.reset(new PolyScaler(base_ctx, mul_ctx, rhs));
        down_scaler.reset(new PolyScaler(mul_ctx, base_ctx, post));
    }

    // Multiplicator::default (mul.rs:101-130)
    Multiplicator(const BfvParams& p, const KeySwitchingKey* rk_, size_t lvl) : par(&p), rk(rk_), level(lvl) {
        base_ctx = p.levels[lvl].ctx;
        size_t msize = 0;
        for (size_t i = 0; i < base_ctx->L(); i++) msize += p.moduli_sizes[i];
        size_t n_moduli = (msize + 60 + 61) / 62;
        std::vector<u64> ext = base_ctx->moduli;
        u64 ub = 1ULL << 62;
        while (ext.size() != base_ctx->L() + n_moduli) {
            ub = generate_prime(62, 2 * (u64)p.degree, ub);
            if (!ub) throw std::runtime_error("generate_prime failed");
            if (std::find(ext.begin(), ext.end(), ub) == ext.end() &&
                std::find(base_ctx->moduli.begin(), base_ctx->moduli.end(), ub) == base_ctx->moduli.end())
                ext.push_back(ub);
        }
        mul_ctx = std::make_shared<Context>(ext, p.degree);
        extender.reset(new PolyScaler(base_ctx, mul_ctx, ScalingFactor()));
        down_scaler.reset(new PolyScaler(mul_ctx, base_ctx, ScalingFactor(Big(p.t), base_ctx->rns->product)));
    }

    // Extend, tensor, scale (mul.rs:175-190). lhs/rhs: [2][L][N] NTT. Returns [3][L][N] NTT.
    PolyData tensor_and_scale(const u64* lhs, const u64* rhs) const {
        const Context& b = *base_ctx;
        const Context& m = *mul_ctx;
        size_t sz = b.L() * b.degree, msz = m.L() * m.degree;
        PolyData c00 = extender->scale(lhs, true), c01 = extender->scale(lhs + sz, true);
        const PolyScaler& er = extender_rhs ? *extender_rhs : *extender;
        PolyData c10 = er.scale(rhs, true), c11 = er.scale(rhs + sz, true);
        PolyData c0 = c00;
        poly_mul(m, c0.data(), c10.data());
        PolyData c1 = c00;
        poly_mul(m, c1.data(), c11.data());
        PolyData t = c01;
        poly_mul(m, t.data(), c10.data());
        poly_add(m, c1.data(), t.data());
        PolyData c2 = c01;
        poly_mul(m, c2.data(), c11.data());
        (void)msz;
        PolyData out(3 * sz);
        PolyData d0 = down_scaler->scale(c0.data(), true);
        PolyData d1 = down_scaler->scale(c1.data(), true);
        PolyData d2 = down_scaler->scale(c2.data(), true);
        std::copy(d0.begin(), d0.end(), out.begin());
        std::copy(d1.begin(), d1.end(), out.begin() + sz);
        std::copy(d2.begin(), d2.end(), out.begin() + 2 * sz);
        return out;
    }

    // RelinearizationKey::relinearizes on a 3-part ct (relinearization_key.rs:71-98).
    static PolyData relinearize(const BfvParams& par, const KeySwitchingKey& rk, const u64* ct3) {
        const Context& c = *rk.ctx_ct;
        size_t sz = c.L() * c.degree;
        PolyData c2(ct3 + 2 * sz, ct3 + 3 * sz);
        ntt_backward(c, c2.data());
        PolyData r0, r1;
        rk.key_switch(c2.data(), r0, r1);
        r0 = bring_down(par, rk.ksk_level, rk.ciphertext_level, r0);
        r1 = bring_down(par, rk.ksk_level, rk.ciphertext_level, r1);
        PolyData out(ct3, ct3 + 2 * sz);
        poly_add(c, out.data(), r0.data());
        poly_add(c, out.data() + sz, r1.data());
        return out;
    }

    // Multiplicator::multiply (mul.rs:158-227) ; returns [2][L][N] (or [3] without rk)
    PolyData multiply(const u64* lhs, const u64* rhs) const {
        PolyData c = tensor_and_scale(lhs, rhs);
        if (!rk) return c;
        return relinearize(*par, *rk, c.data());
    }
};

}  // namespace oracle
