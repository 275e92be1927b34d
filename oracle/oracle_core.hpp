// ORACLE — test infrastructure only.
//
// A plain C++ restatement of the fhe.rs CPU algorithms on the RNS-BFV evaluation
// path. It exists to check the HIP library (tests/, the smoke entry and the CPU
// baseline leg of bench.py may load it); the product library never links,
// includes or calls anything under oracle/.
//
// Citations are to the fhe.rs sources (crates/fhe-math = math/, crates/fhe = fhe/).
//
// Pinning (what checks each part, see tests/test_oracle_*.py):
//   Modulus ops          vs Python int arithmetic (u128 %)            -> pinned
//   generate_prime       vs NFLlib KAT math/src/zq/primes.rs:67-102    -> pinned
//   generate_moduli      vs KAT fhe/src/bfv/parameters.rs:740-773      -> pinned
//   NttOperator          vs closed form A[i] = a(psi^(2 brv(i)+1)),
//                        O(N^2) negacyclic brute force, round trip    -> pinned
//   primitive_root       fhe.rs's draw from ChaCha8Rng::seed_from_u64(0)
//                        (oracle_rng.hpp); ChaCha and PCG pinned, the
//                        uniform draw (rand 0.9) unpinned              -> parity unpinned
//                        for the NTT-domain representation only; every
//                        power-basis result is independent of the root.
//   RnsContext           vs KAT math/src/rns/mod.rs:205-242            -> pinned
//   RnsScaler            vs round(x*n/d) closed form with the sign
//                        convention of math/src/rns/scaler.rs:397-414  -> pinned
//   switch_down          vs round(x / q_last) closed form              -> pinned
//   substitute           vs x -> x^k closed form in power basis        -> pinned
//   key_switch           vs Python big-int recomputation (tiny N)      -> pinned
//   Multiplicator        vs decrypt(ct*ct) == m1 (.) m2 mod t           -> pinned
//   GaloisKey            vs decrypt(rotate(ct)) == rotated slots        -> pinned
#pragma once

#include <algorithm>
#include <cassert>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "oracle_rng.hpp"

namespace oracle {

typedef unsigned __int128 u128;
typedef uint64_t u64;
typedef int64_t i64;

static inline int lzcnt64(u64 x) { return x == 0 ? 64 : __builtin_clzll(x); }

// ---------------------------------------------------------------------------
// Minimal unsigned big integer (base 2^32, little endian). Exactness is all that
// matters here; speed does not (setup only).
// ---------------------------------------------------------------------------
struct Big {
    std::vector<uint32_t> w;  // no trailing zero words

    Big() {}
    Big(u64 v) {
        while (v) { w.push_back((uint32_t)v); v >>= 32; }
    }
    void trim() { while (!w.empty() && w.back() == 0) w.pop_back(); }
    bool is_zero() const { return w.empty(); }
    size_t bits() const {
        if (w.empty()) return 0;
        return 32 * (w.size() - 1) + (32 - __builtin_clz(w.back()));
    }
    bool bit(size_t i) const {
        size_t k = i / 32;
        return k < w.size() && ((w[k] >> (i % 32)) & 1);
    }
    u64 low64() const {
        u64 r = 0;
        if (w.size() > 0) r |= w[0];
        if (w.size() > 1) r |= (u64)w[1] << 32;
        return r;
    }
    static int cmp(const Big& a, const Big& b) {
        if (a.w.size() != b.w.size()) return a.w.size() < b.w.size() ? -1 : 1;
        for (size_t i = a.w.size(); i-- > 0;) {
            if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
        }
        return 0;
    }
    static Big add(const Big& a, const Big& b) {
        Big r;
        size_t n = std::max(a.w.size(), b.w.size());
        r.w.resize(n + 1);
        u64 c = 0;
        for (size_t i = 0; i < n; i++) {
            u64 s = c;
            if (i < a.w.size()) s += a.w[i];
            if (i < b.w.size()) s += b.w[i];
            r.w[i] = (uint32_t)s;
            c = s >> 32;
        }
        r.w[n] = (uint32_t)c;
        r.trim();
        return r;
    }
    // requires a >= b
    static Big sub(const Big& a, const Big& b) {
        if (cmp(a, b) < 0) throw std::runtime_error("Big::sub underflow");
        Big r;
        r.w.resize(a.w.size());
        int64_t br = 0;
        for (size_t i = 0; i < a.w.size(); i++) {
            int64_t s = (int64_t)a.w[i] - br - (i < b.w.size() ? (int64_t)b.w[i] : 0);
            br = s < 0;
            if (s < 0) s += ((int64_t)1 << 32);
            r.w[i] = (uint32_t)s;
        }
        r.trim();
        return r;
    }
    static Big mul(const Big& a, const Big& b) {
        if (a.is_zero() || b.is_zero()) return Big();
        Big r;
        r.w.assign(a.w.size() + b.w.size(), 0);
        for (size_t i = 0; i < a.w.size(); i++) {
            u64 c = 0;
            for (size_t j = 0; j < b.w.size(); j++) {
                u64 t = (u64)a.w[i] * b.w[j] + r.w[i + j] + c;
                r.w[i + j] = (uint32_t)t;
                c = t >> 32;
            }
            size_t k = i + b.w.size();
            while (c) {
                u64 t = (u64)r.w[k] + c;
                r.w[k] = (uint32_t)t;
                c = t >> 32;
                k++;
            }
        }
        r.trim();
        return r;
    }
    static Big shl(const Big& a, size_t s) {
        if (a.is_zero()) return Big();
        Big r;
        size_t ws = s / 32, bs = s % 32;
        r.w.assign(a.w.size() + ws + 1, 0);
        for (size_t i = 0; i < a.w.size(); i++) {
            u64 v = (u64)a.w[i] << bs;
            r.w[i + ws] |= (uint32_t)v;
            r.w[i + ws + 1] |= (uint32_t)(v >> 32);
        }
        r.trim();
        return r;
    }
    static Big shr(const Big& a, size_t s) {
        size_t ws = s / 32, bs = s % 32;
        if (ws >= a.w.size()) return Big();
        Big r;
        r.w.assign(a.w.size() - ws, 0);
        for (size_t i = 0; i < r.w.size(); i++) {
            u64 v = a.w[i + ws];
            if (i + ws + 1 < a.w.size()) v |= (u64)a.w[i + ws + 1] << 32;
            r.w[i] = (uint32_t)(v >> bs);
        }
        r.trim();
        return r;
    }
    // Binary long division: q = a / b, r = a % b.
    static void divmod(const Big& a, const Big& b, Big& q, Big& r) {
        if (b.is_zero()) throw std::runtime_error("Big::divmod by zero");
        q = Big();
        r = Big();
        if (cmp(a, b) < 0) { r = a; return; }
        size_t n = a.bits();
        q.w.assign((n + 31) / 32, 0);
        for (size_t i = n; i-- > 0;) {
            r = shl(r, 1);
            if (a.bit(i)) {
                if (r.w.empty()) r.w.push_back(1); else r.w[0] |= 1;
            }
            if (cmp(r, b) >= 0) {
                r = sub(r, b);
                q.w[i / 32] |= (1u << (i % 32));
            }
        }
        q.trim();
    }
    static Big div(const Big& a, const Big& b) { Big q, r; divmod(a, b, q, r); return q; }
    static Big mod(const Big& a, const Big& b) { Big q, r; divmod(a, b, q, r); return r; }
    u64 mod_u64(u64 m) const {
        u128 r = 0;
        for (size_t i = w.size(); i-- > 0;) r = ((r << 32) | w[i]) % m;
        return (u64)r;
    }
};

// ---------------------------------------------------------------------------
// Primality (fhe-util/src/lib.rs:16-18 uses num-bigint-dig probably_prime(p, 0),
// i.e. Baillie-PSW, which is exact below 2^64; a deterministic Miller-Rabin
// with the first twelve primes as bases is also exact below 3.3e24).
// ---------------------------------------------------------------------------
static inline u64 powmod_u64(u64 a, u64 e, u64 m) {
    u128 r = 1 % m, b = a % m;
    while (e) {
        if (e & 1) r = (r * b) % m;
        b = (b * b) % m;
        e >>= 1;
    }
    return (u64)r;
}

static inline bool is_prime(u64 n) {
    if (n < 2) return false;
    static const u64 small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (u64 p : small) {
        if (n % p == 0) return n == p;
    }
    u64 d = n - 1;
    int s = 0;
    while ((d & 1) == 0) { d >>= 1; s++; }
    for (u64 a : small) {
        u64 x = powmod_u64(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int r = 1; r < s; r++) {
            x = (u64)(((u128)x * x) % n);
            if (x == n - 1) { comp = false; break; }
        }
        if (comp) return false;
    }
    return true;
}

// math/src/zq/primes.rs:10-24
static inline bool supports_opt(u64 p) {
    int lz = lzcnt64(p);
    if (lz < 1) return false;
    Big middle = Big::shl(Big(1), 3 * (size_t)lz);
    Big left = Big::shl(Big::add(middle, Big(1)), 64);
    middle = Big::mul(middle, Big((1ULL << lz) + 1));
    middle = Big::mul(middle, Big(p));
    return Big::cmp(left, middle) < 0;
}

// math/src/zq/primes.rs:30-59 ; returns 0 for None
static inline u64 generate_prime(size_t num_bits, u64 modulo, u64 upper_bound) {
    if (num_bits < 10 || num_bits > 62) return 0;
    int leading_zeros = 64 - (int)num_bits;
    u64 t = upper_bound - 1;
    while (t % modulo != 1 && lzcnt64(t) == leading_zeros) t -= 1;
    while (lzcnt64(t) == leading_zeros && !is_prime(t) && t >= modulo) t -= modulo;
    if (lzcnt64(t) == leading_zeros && is_prime(t)) return t;
    return 0;
}

// Modular inverse for coprime (not necessarily prime) moduli, extended Euclid.
static inline u64 inv_mod_euclid(u64 a, u64 m) {
    __int128 t = 0, nt = 1, r = m, nr = a % m;
    while (nr != 0) {
        __int128 q = r / nr, tmp;
        tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    if (r != 1) throw std::runtime_error("not invertible");
    if (t < 0) t += m;
    return (u64)t;
}

// ---------------------------------------------------------------------------
// Modulus — math/src/zq/mod.rs:31-98 and the primitives defined at :103-767.
// ---------------------------------------------------------------------------
struct Modulus {
    u64 p = 0;
    u64 barrett_hi = 0, barrett_lo = 0;
    int leading_zeros = 0;
    bool supports_opt_ = false;

    Modulus() {}
    explicit Modulus(u64 p_) : p(p_) {
        // math/src/zq/mod.rs:83-98
        if (p < 2 || (p >> 62) != 0) throw std::runtime_error("InvalidModulus");
        u128 all = ~(u128)0;
        u128 b = all / p;
        if (all % p == p - 1) b += 1;  // floor(2^128 / p)
        barrett_hi = (u64)(b >> 64);
        barrett_lo = (u64)b;
        leading_zeros = lzcnt64(p);
        supports_opt_ = supports_opt(p);
    }
    // :659-668
    static inline u64 reduce1(u64 x, u64 p) { return x >= p ? x - p : x; }
    // :693-707
    inline u64 lazy_reduce_u128(u128 a) const {
        u64 a_lo = (u64)a, a_hi = (u64)(a >> 64);
        u128 p_lo_lo = ((u128)a_lo * barrett_lo) >> 64;
        u128 p_hi_lo = (u128)a_hi * barrett_lo;
        u128 p_lo_hi = (u128)a_lo * barrett_hi;
        u128 q = ((p_lo_hi + p_hi_lo + p_lo_lo) >> 64) + (u128)a_hi * barrett_hi;
        return (u64)(a - q * p);
    }
    // :712-722
    inline u64 lazy_reduce(u64 a) const {
        u128 p_lo_lo = ((u128)a * barrett_lo) >> 64;
        u128 p_lo_hi = (u128)a * barrett_hi;
        u128 q = (p_lo_hi + p_lo_lo) >> 64;
        return (u64)((u128)a - q * p);
    }
    // :730-740
    inline u64 lazy_reduce_opt_u128(u128 a) const {
        u128 q = (((u128)barrett_lo * (a >> 64)) + (a << leading_zeros)) >> 64;
        return (u64)(a - q * p);
    }
    // :744-752
    inline u64 lazy_reduce_opt(u64 a) const {
        u64 q = a >> (64 - leading_zeros);
        return (u64)((u128)a - (u128)q * p);
    }
    inline u64 reduce_u128(u128 a) const { return reduce1(lazy_reduce_u128(a), p); }
    inline u64 reduce(u64 a) const { return reduce1(lazy_reduce(a), p); }
    inline u64 reduce_opt_u128(u128 a) const { return reduce1(lazy_reduce_opt_u128(a), p); }
    inline u64 reduce_i64(i64 a) const {
        // :476-478
        return reduce_u128((u128)((((__int128)p) << 64) + (__int128)a));
    }
    inline u64 add(u64 a, u64 b) const { return reduce1(a + b, p); }
    inline u64 sub(u64 a, u64 b) const { return reduce1(a + p - b, p); }
    inline u64 neg(u64 a) const { return reduce1(p - a, p); }
    inline u64 mul(u64 a, u64 b) const { return reduce_u128((u128)a * b); }
    inline u64 mul_opt(u64 a, u64 b) const { return reduce_opt_u128((u128)a * b); }
    // mul_vec picks mul_opt when supported (:332-344); both are exact.
    inline u64 mul_auto(u64 a, u64 b) const { return supports_opt_ ? mul_opt(a, b) : mul(a, b); }
    inline u64 shoup(u64 a) const { return (u64)(((u128)a << 64) / p); }
    // :224-234
    inline u64 lazy_mul_shoup(u64 a, u64 b, u64 b_shoup) const {
        u128 q = ((u128)a * b_shoup) >> 64;
        return (u64)((u128)a * b - q * p);
    }
    inline u64 mul_shoup(u64 a, u64 b, u64 b_shoup) const {
        return reduce1(lazy_mul_shoup(a, b, b_shoup), p);
    }
    // :756-762
    inline u64 lazy_reduce_vec_elem(u64 a) const {
        return supports_opt_ ? lazy_reduce_opt(a) : lazy_reduce(a);
    }
    // :447-456
    inline i64 center(u64 a) const {
        u64 threshold = p >> 1;
        return a >= threshold ? (i64)(a - p) : (i64)a;
    }
    // :531-550
    u64 pow(u64 a, u64 n) const {
        if (n == 0) return 1;
        if (n == 1) return a;
        u64 r = a;
        for (int i = 62 - lzcnt64(n); i >= 0; i--) {
            r = mul(r, r);
            if ((n >> i) & 1) r = mul(r, a);
        }
        return r;
    }
    // :556-565 (None -> 0)
    u64 inv(u64 a) const {
        if (!is_prime(p) || a == 0) return 0;
        return pow(a, p - 2);
    }
};

// ---------------------------------------------------------------------------
// NttOperator — math/src/ntt/native.rs
// ---------------------------------------------------------------------------
static inline size_t bitrev(size_t i, int logn) {
    size_t r = 0;
    for (int b = 0; b < logn; b++) r |= ((i >> b) & 1) << (logn - 1 - b);
    return r;
}
static inline int ilog2(size_t n) { return 63 - __builtin_clzll((u64)n); }

// math/src/ntt/mod.rs:19-23
static inline bool supports_ntt(u64 p, size_t n) {
    return n >= 8 && (n & (n - 1)) == 0 && p % ((u64)n << 1) == 1 && is_prime(p);
}

// Primitive 2n-th root of unity mod p (math/src/ntt/native.rs:320-336): up
// to 100 candidates drawn with random_range(0..p) from
// ChaCha8Rng::seed_from_u64(0), each raised to the power (p-1)/(2n), the
// first of order 2n (native.rs:341-348) kept.
static inline u64 primitive_root(size_t n, const Modulus& p) {
    const u64 lambda = (p.p - 1) / (2 * (u64)n);
    oracle_rng::ChaCha8Rng rng(oracle_rng::pcg32_seed(0));
    for (int i = 0; i < 100; i++) {
        const u64 root = p.pow(oracle_rng::uniform_single(rng, 0, p.p), lambda);
        if (p.pow(root, 2 * (u64)n) == 1 && p.pow(root, (u64)n) != 1) return root;
    }
    throw std::runtime_error("no primitive root");
}

struct NttOperator {
    Modulus p;
    u64 p_twice = 0;
    size_t size = 0;
    int logn = 0;
    u64 psi = 0;
    std::vector<u64> omegas, omegas_shoup, zetas_inv, zetas_inv_shoup;
    u64 size_inv = 0, size_inv_shoup = 0;

    // native.rs:35-73 ; psi == 0 selects primitive_root()
    NttOperator(const Modulus& m, size_t n, u64 psi_in = 0) : p(m) {
        if (!supports_ntt(m.p, n)) throw std::runtime_error("NTT unsupported");
        size = n;
        logn = ilog2(n);
        p_twice = 2 * m.p;
        size_inv = m.inv((u64)n);
        psi = psi_in ? psi_in : primitive_root(n, m);
        u64 omega_inv = m.inv(psi);
        std::vector<u64> powers(n), powers_inv(n);
        powers[0] = 1;
        for (size_t i = 1; i < n; i++) powers[i] = m.mul(powers[i - 1], psi);
        powers_inv[0] = omega_inv;
        for (size_t i = 1; i < n; i++) powers_inv[i] = m.mul(powers_inv[i - 1], omega_inv);
        omegas.resize(n); zetas_inv.resize(n); omegas_shoup.resize(n); zetas_inv_shoup.resize(n);
        for (size_t i = 0; i < n; i++) {
            size_t j = bitrev(i, logn);
            omegas[i] = powers[j];
            zetas_inv[i] = powers_inv[j];
            omegas_shoup[i] = m.shoup(omegas[i]);
            zetas_inv_shoup[i] = m.shoup(zetas_inv[i]);
        }
        size_inv_shoup = m.shoup(size_inv);
    }
    inline u64 reduce3(u64 a) const {
        return Modulus::reduce1(Modulus::reduce1(a, p_twice), p.p);
    }
    // native.rs:256-269
    inline void butterfly(u64& x, u64& y, u64 w, u64 ws) const {
        x = Modulus::reduce1(x, p_twice);
        u64 t = p.lazy_mul_shoup(y, w, ws);
        y = x + p_twice - t;
        x += t;
    }
    // native.rs:288-300
    inline void inv_butterfly(u64& x, u64& y, u64 z, u64 zs) const {
        u64 t = x;
        x = Modulus::reduce1(y + t, p_twice);
        y = p.lazy_mul_shoup(p_twice + t - y, z, zs);
    }
    // native.rs:77-102 (forward); forward_vt is forward_vt_lazy + reduce3 (:183-189),
    // which produces the same reduced output.
    void forward(u64* a) const {
        size_t l = size >> 1, k = 1;
        while (l > 0) {
            for (size_t s = 0; s < size; s += 2 * l) {
                u64 w = omegas[k], ws = omegas_shoup[k];
                k++;
                if (l == 1) {
                    butterfly(a[s], a[s + 1], w, ws);
                    a[s] = reduce3(a[s]);
                    a[s + 1] = reduce3(a[s + 1]);
                } else {
                    for (size_t j = s; j < s + l; j++) butterfly(a[j], a[j + l], w, ws);
                }
            }
            l >>= 1;
        }
    }
    // native.rs:142-175 : output < 4p
    void forward_lazy(u64* a) const {
        size_t l = size >> 1, m = 1, k = 1;
        while (l > 0) {
            for (size_t i = 0; i < m; i++) {
                u64 w = omegas[k], ws = omegas_shoup[k];
                k++;
                size_t s = 2 * i * l;
                for (size_t j = s; j < s + l; j++) butterfly(a[j], a[j + l], w, ws);
            }
            l >>= 1;
            m <<= 1;
        }
    }
    void forward_vt(u64* a) const {
        forward_lazy(a);
        for (size_t i = 0; i < size; i++) a[i] = reduce3(a[i]);
    }
    // native.rs:106-132
    void backward(u64* a) const {
        size_t k = 0, l = 1;
        while (l < size) {
            for (size_t s = 0; s < size; s += 2 * l) {
                u64 z = zetas_inv[k], zs = zetas_inv_shoup[k];
                k++;
                for (size_t j = s; j < s + l; j++) inv_butterfly(a[j], a[j + l], z, zs);
            }
            l <<= 1;
        }
        for (size_t i = 0; i < size; i++) a[i] = p.mul_shoup(a[i], size_inv, size_inv_shoup);
    }
};

// ---------------------------------------------------------------------------
// RnsContext — math/src/rns/mod.rs:24-147
// ---------------------------------------------------------------------------
struct RnsContext {
    std::vector<u64> moduli_u64;
    std::vector<Modulus> moduli;
    std::vector<u64> q_tilde, q_tilde_shoup;
    std::vector<Big> q_star, garner;
    Big product;

    explicit RnsContext(const std::vector<u64>& m) : moduli_u64(m) {
        if (m.empty()) throw std::runtime_error("The list of moduli is empty");
        product = Big(1);
        for (size_t i = 0; i < m.size(); i++) {
            for (size_t j = 0; j < m.size(); j++) {
                if (i != j) {
                    u64 a = m[i], b = m[j];
                    while (b) { u64 t = a % b; a = b; b = t; }
                    if (a != 1) throw std::runtime_error("The moduli are not coprime");
                }
            }
            product = Big::mul(product, Big(m[i]));
        }
        for (size_t i = 0; i < m.size(); i++) {
            Modulus mi(m[i]);
            Big qs = Big::div(product, Big(m[i]));
            u64 qt = inv_mod_euclid(qs.mod_u64(m[i]), m[i]);
            moduli.push_back(mi);
            q_star.push_back(qs);
            q_tilde.push_back(qt);
            q_tilde_shoup.push_back(mi.shoup(qt));
            garner.push_back(Big::mul(qs, Big(qt)));
        }
    }
    std::vector<u64> project(const Big& a) const {
        std::vector<u64> r;
        for (u64 m : moduli_u64) r.push_back(a.mod_u64(m));
        return r;
    }
    Big lift(const u64* rests) const {
        Big r;
        for (size_t i = 0; i < garner.size(); i++) r = Big::add(r, Big::mul(garner[i], Big(rests[i])));
        return Big::mod(r, product);
    }
};

// ---------------------------------------------------------------------------
// Wrapping 256-bit unsigned integer (stands for ethnum::U256 in
// math/src/rns/scaler.rs:249-352).
// ---------------------------------------------------------------------------
struct U256 {
    u64 v[4] = {0, 0, 0, 0};
    static U256 mul_u64_u128(u64 a, u128 b) {
        U256 r;
        u128 lo = (u128)a * (u64)b;
        u128 hi = (u128)a * (u64)(b >> 64);
        r.v[0] = (u64)lo;
        u128 mid = (lo >> 64) + (u64)hi;
        r.v[1] = (u64)mid;
        u128 top = (mid >> 64) + (hi >> 64);
        r.v[2] = (u64)top;
        r.v[3] = (u64)(top >> 64);
        return r;
    }
    U256 wadd(const U256& o) const {
        U256 r;
        u128 c = 0;
        for (int i = 0; i < 4; i++) {
            u128 s = (u128)v[i] + o.v[i] + c;
            r.v[i] = (u64)s;
            c = s >> 64;
        }
        return r;
    }
    U256 wsub(const U256& o) const {
        U256 r;
        u64 br = 0;
        for (int i = 0; i < 4; i++) {
            u64 a = v[i], b = o.v[i];
            u64 d = a - b - br;
            br = (a < b) || (a - b < br) ? 1 : 0;
            r.v[i] = d;
        }
        return r;
    }
    U256 bnot() const {
        U256 r;
        for (int i = 0; i < 4; i++) r.v[i] = ~v[i];
        return r;
    }
    U256 shr(unsigned s) const {
        U256 r;
        unsigned ws = s / 64, bs = s % 64;
        for (unsigned i = 0; i + ws < 4; i++) {
            u64 lo = v[i + ws] >> bs;
            u64 hi = (bs && i + ws + 1 < 4) ? (v[i + ws + 1] << (64 - bs)) : 0;
            r.v[i] = lo | hi;
        }
        return r;
    }
    bool nonzero() const { return v[0] | v[1] | v[2] | v[3]; }
    u128 as_u128() const { return ((u128)v[1] << 64) | v[0]; }
};

// ---------------------------------------------------------------------------
// RnsScaler — math/src/rns/scaler.rs:52-352
// ---------------------------------------------------------------------------
struct ScalingFactor {
    Big numerator, denominator;
    bool is_one = true;
    ScalingFactor() : numerator(1), denominator(1), is_one(true) {}
    ScalingFactor(const Big& n, const Big& d) : numerator(n), denominator(d) {
        if (d.is_zero()) throw std::runtime_error("zero denominator");
        is_one = Big::cmp(n, d) == 0;
    }
};

struct RnsScaler {
    const RnsContext* from;
    const RnsContext* to;
    ScalingFactor sf;
    std::vector<u64> gamma, gamma_shoup;
    u64 theta_gamma_lo = 0, theta_gamma_hi = 0;
    bool theta_gamma_sign = false;
    std::vector<std::vector<u64>> omega, omega_shoup;  // [to][from]
    std::vector<u64> theta_omega_lo, theta_omega_hi;
    std::vector<bool> theta_omega_sign;
    std::vector<u64> theta_garner_lo, theta_garner_hi;
    size_t theta_garner_shift = 0;

    // scaler.rs:183-229
    static void extract_projection_and_theta(const RnsContext& ctx, const Big& input,
                                             const Big& num, const Big& den, bool round_up,
                                             std::vector<u64>& projected, u64& lo, u64& hi,
                                             bool& sign) {
        Big ni = Big::mul(num, input);
        Big gamma = Big::div(Big::add(ni, Big::shr(den, 1)), den);
        projected = ctx.project(gamma);
        Big theta = Big::mod(ni, den);
        sign = false;
        if (Big::cmp(den, Big(1)) > 0) {
            Big half = Big::shr(den, 1);
            if (den.bit(0)) {
                if (Big::cmp(theta, half) > 0) { sign = true; theta = Big::sub(den, theta); }
            } else {
                if (Big::cmp(theta, half) >= 0) { sign = true; theta = Big::sub(den, theta); }
            }
        }
        Big t127 = Big::shl(theta, 127);
        Big dm1 = Big::sub(den, Big(1));
        bool up = round_up ? !sign : sign;
        theta = up ? Big::div(Big::add(t127, dm1), den) : Big::div(t127, den);
        Big th = Big::shr(theta, 64);
        lo = theta.low64();
        hi = th.low64();
    }

    // scaler.rs:79-175
    RnsScaler(const RnsContext* f, const RnsContext* t, const ScalingFactor& s)
        : from(f), to(t), sf(s) {
        extract_projection_and_theta(*to, from->product, sf.numerator, sf.denominator, false,
                                     gamma, theta_gamma_lo, theta_gamma_hi, theta_gamma_sign);
        for (size_t i = 0; i < gamma.size(); i++) gamma_shoup.push_back(to->moduli[i].shoup(gamma[i]));
        size_t nf = from->moduli.size(), nt = to->moduli.size();
        omega.assign(nt, std::vector<u64>(nf));
        omega_shoup.assign(nt, std::vector<u64>(nf));
        for (size_t i = 0; i < nf; i++) {
            std::vector<u64> om;
            u64 lo, hi;
            bool sg;
            extract_projection_and_theta(*to, from->garner[i], sf.numerator, sf.denominator, true,
                                         om, lo, hi, sg);
            theta_omega_lo.push_back(lo);
            theta_omega_hi.push_back(hi);
            theta_omega_sign.push_back(sg);
            for (size_t j = 0; j < nt; j++) {
                omega[j][i] = to->moduli[j].reduce(om[j]);
                omega_shoup[j][i] = to->moduli[j].shoup(omega[j][i]);
            }
        }
        // scaler.rs:130-142
        size_t shift = 127;
        for (u64 qi : from->moduli_u64) {
            u128 x = (u128)qi * (u128)from->moduli_u64.size();
            // next_power_of_two().ilog2()
            int lg = 0;
            while (((u128)1 << lg) < x) lg++;
            size_t c = 192 - 1 - (size_t)lg;
            shift = std::min(shift, c);
        }
        theta_garner_shift = shift;
        for (size_t i = 0; i < nf; i++) {
            Big th = Big::div(Big::add(Big::shl(from->garner[i], shift), Big::shr(from->product, 1)),
                              from->product);
            theta_garner_lo.push_back(th.low64());
            theta_garner_hi.push_back(Big::shr(th, 64).low64());
        }
    }

    // scaler.rs:249-352
    void scale(const u64* rests, u64* out, size_t out_len, size_t starting_index,
               size_t out_stride = 1) const {
        size_t nf = from->moduli.size();
        U256 sum;
        for (size_t j = 0; j < nf; j++) {
            u128 th = ((u128)theta_garner_hi[j] << 64) | theta_garner_lo[j];
            sum = sum.wadd(U256::mul_u64_u128(rests[j], th));
        }
        sum = sum.shr((unsigned)(theta_garner_shift - 1));
        u128 vv = sum.as_u128();
        u128 v = (vv >> 1) + (vv & 1);  // div_ceil(2)

        bool w_sign = false;
        u128 w = 0;
        if (!sf.is_one) {
            U256 so;
            for (size_t j = 0; j < nf; j++) {
                u128 th = ((u128)theta_omega_hi[j] << 64) | theta_omega_lo[j];
                U256 prod = U256::mul_u64_u128(rests[j], th);
                so = theta_omega_sign[j] ? so.wsub(prod) : so.wadd(prod);
            }
            // v * theta_gamma: v < 2^128, theta_gamma < 2^128 ; full product fits 256 bits
            u128 tg = ((u128)theta_gamma_hi << 64) | theta_gamma_lo;
            U256 a = U256::mul_u64_u128((u64)v, tg);
            U256 b = U256::mul_u64_u128((u64)(v >> 64), tg);
            U256 bsh;  // b << 64
            bsh.v[1] = b.v[0]; bsh.v[2] = b.v[1]; bsh.v[3] = b.v[2];
            U256 vtg = a.wadd(bsh);
            so = theta_gamma_sign ? so.wadd(vtg) : so.wsub(vtg);
            w_sign = so.shr(63 + 128).nonzero();
            if (w_sign) {
                w = so.bnot().shr(126).as_u128() + 1;
                w /= 2;
            } else {
                w = so.shr(126).as_u128();
                w = (w >> 1) + (w & 1);
            }
        }
        for (size_t i = 0; i < out_len; i++) {
            const Modulus& qi = to->moduli[starting_index + i];
            const std::vector<u64>& om = omega[starting_index + i];
            const std::vector<u64>& oms = omega_shoup[starting_index + i];
            u128 yi = (u128)(qi.p * 2 - qi.lazy_mul_shoup(qi.reduce_u128(v), gamma[starting_index + i],
                                                        gamma_shoup[starting_index + i]));
            if (!sf.is_one) {
                u64 wi = qi.lazy_reduce_u128(w);
                yi += (u128)(w_sign ? qi.p * 2 - wi : wi);
            }
            for (size_t j = 0; j < nf; j++) yi += (u128)qi.lazy_mul_shoup(rests[j], om[j], oms[j]);
            out[i * out_stride] = qi.reduce_u128(yi);
        }
    }
};

}  // namespace oracle

