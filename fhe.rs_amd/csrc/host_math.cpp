#include "host_math.h"

#include <algorithm>
#include <stdexcept>

namespace fhehip {

u64 mulmod64(u64 a, u64 b, u64 m) { return (u64)(((u128)a * b) % m); }

u64 powmod64(u64 a, u64 e, u64 m) {
    u64 r = 1 % m;
    a %= m;
    while (e) {
        if (e & 1) r = mulmod64(r, a, m);
        a = mulmod64(a, a, m);
        e >>= 1;
    }
    return r;
}

u64 invmod64(u64 a, u64 m) {
    // Extended Euclid on signed 128-bit to cover composite moduli too.
    __int128 r0 = m, r1 = a % m, s0 = 0, s1 = 1;
    while (r1) {
        __int128 q = r0 / r1, t;
        t = r0 - q * r1; r0 = r1; r1 = t;
        t = s0 - q * s1; s0 = s1; s1 = t;
    }
    if (r0 != 1) throw std::invalid_argument("value not invertible");
    if (s0 < 0) s0 += m;
    return (u64)s0;
}

u64 shoup64(u64 a, u64 p) { return (u64)(((u128)a << 64) / p); }

bool is_prime64(u64 n) {
    // Deterministic Miller-Rabin; the base set is exact for n < 3.3e24.
    static const u64 bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (u64 b : bases)
        if (n % b == 0) return n == b;
    u64 d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; ++s; }
    for (u64 b : bases) {
        u64 x = powmod64(b, d, n);
        if (x == 1 || x == n - 1) continue;
        bool witness = true;
        for (int i = 1; i < s && witness; i++) {
            x = mulmod64(x, x, n);
            if (x == n - 1) witness = false;
        }
        if (witness) return false;
    }
    return true;
}

bool supports_ntt64(u64 p, u64 n) {
    // math/src/ntt/mod.rs:19-23
    if (n < 8 || (n & (n - 1))) return false;
    return p % (n << 1) == 1 && is_prime64(p);
}

static int leading_zeros64(u64 x) { return x ? __builtin_clzll(x) : 64; }

u64 generate_prime64(u32 num_bits, u64 modulo, u64 upper_bound) {
    if (num_bits < 10 || num_bits > 62) return 0;
    const int lz = 64 - (int)num_bits;
    u64 c = upper_bound - 1;
    while (c % modulo != 1 && leading_zeros64(c) == lz) --c;
    while (leading_zeros64(c) == lz && !is_prime64(c) && c >= modulo) c -= modulo;
    return (leading_zeros64(c) == lz && is_prime64(c)) ? c : 0;
}

u64 primitive_root64(u64 n, u64 p) {
    const u64 e = (p - 1) / (2 * n);
    for (u64 g = 2; g < p; ++g) {
        u64 r = powmod64(g, e, p);
        if (powmod64(r, n, p) == p - 1) return r;  // order exactly 2n
    }
    throw std::invalid_argument("no primitive root");
}

void barrett128(u64 p, u64& lo, u64& hi) {
    u128 all = ~(u128)0;
    u128 q = all / p;
    if (all % p == p - 1) q += 1;
    lo = (u64)q;
    hi = (u64)(q >> 64);
}

int ilog2_u64(u64 x) { return 63 - __builtin_clzll(x); }

u32 bit_reverse(u32 x, int bits) {
    u32 r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

// ---------------------------------------------------------------------------
void BigUint::normalize() {
    while (!limbs_.empty() && limbs_.back() == 0) limbs_.pop_back();
}

BigUint BigUint::from_words32(const uint32_t* w, size_t n) {
    BigUint r;
    r.limbs_.assign((n + 1) / 2, 0);
    for (size_t i = 0; i < n; i++) r.limbs_[i / 2] |= (u64)w[i] << (32 * (i % 2));
    r.normalize();
    return r;
}

std::vector<uint32_t> BigUint::words32() const {
    std::vector<uint32_t> w;
    for (u64 l : limbs_) {
        w.push_back((uint32_t)l);
        w.push_back((uint32_t)(l >> 32));
    }
    while (!w.empty() && w.back() == 0) w.pop_back();
    return w;
}

size_t BigUint::bits() const {
    if (limbs_.empty()) return 0;
    return 64 * (limbs_.size() - 1) + (64 - __builtin_clzll(limbs_.back()));
}

bool BigUint::test_bit(size_t i) const { return (limb(i / 64) >> (i % 64)) & 1; }

u64 BigUint::mod_u64(u64 m) const {
    u128 r = 0;
    for (size_t i = limbs_.size(); i-- > 0;) r = ((r << 64) | limbs_[i]) % m;
    return (u64)r;
}

int BigUint::compare(const BigUint& o) const {
    if (limbs_.size() != o.limbs_.size()) return limbs_.size() < o.limbs_.size() ? -1 : 1;
    for (size_t i = limbs_.size(); i-- > 0;)
        if (limbs_[i] != o.limbs_[i]) return limbs_[i] < o.limbs_[i] ? -1 : 1;
    return 0;
}

BigUint BigUint::operator+(const BigUint& o) const {
    BigUint r;
    size_t n = std::max(limbs_.size(), o.limbs_.size());
    r.limbs_.resize(n + 1);
    u128 c = 0;
    for (size_t i = 0; i < n; i++) {
        c += (u128)limb(i) + o.limb(i);
        r.limbs_[i] = (u64)c;
        c >>= 64;
    }
    r.limbs_[n] = (u64)c;
    r.normalize();
    return r;
}

BigUint BigUint::operator-(const BigUint& o) const {
    if (compare(o) < 0) throw std::invalid_argument("BigUint subtraction underflow");
    BigUint r;
    r.limbs_.resize(limbs_.size());
    u64 borrow = 0;
    for (size_t i = 0; i < limbs_.size(); i++) {
        u64 a = limbs_[i], b = o.limb(i);
        u64 d = a - b - borrow;
        borrow = (a < b || (a - b) < borrow) ? 1 : 0;
        r.limbs_[i] = d;
    }
    r.normalize();
    return r;
}

BigUint BigUint::operator*(const BigUint& o) const {
    BigUint r;
    if (is_zero() || o.is_zero()) return r;
    r.limbs_.assign(limbs_.size() + o.limbs_.size(), 0);
    for (size_t i = 0; i < limbs_.size(); i++) {
        u128 c = 0;
        for (size_t j = 0; j < o.limbs_.size(); j++) {
            c += (u128)limbs_[i] * o.limbs_[j] + r.limbs_[i + j];
            r.limbs_[i + j] = (u64)c;
            c >>= 64;
        }
        r.limbs_[i + o.limbs_.size()] = (u64)c;
    }
    r.normalize();
    return r;
}

BigUint BigUint::operator<<(size_t s) const {
    BigUint r;
    if (is_zero()) return r;
    size_t ls = s / 64, bs = s % 64;
    r.limbs_.assign(limbs_.size() + ls + 1, 0);
    for (size_t i = 0; i < limbs_.size(); i++) {
        r.limbs_[i + ls] |= limbs_[i] << bs;
        if (bs) r.limbs_[i + ls + 1] |= limbs_[i] >> (64 - bs);
    }
    r.normalize();
    return r;
}

BigUint BigUint::operator>>(size_t s) const {
    BigUint r;
    size_t ls = s / 64, bs = s % 64;
    if (ls >= limbs_.size()) return r;
    r.limbs_.assign(limbs_.size() - ls, 0);
    for (size_t i = 0; i < r.limbs_.size(); i++) {
        r.limbs_[i] = limbs_[i + ls] >> bs;
        if (bs && i + ls + 1 < limbs_.size()) r.limbs_[i] |= limbs_[i + ls + 1] << (64 - bs);
    }
    r.normalize();
    return r;
}

void BigUint::divmod(const BigUint& a, const BigUint& b, BigUint& q, BigUint& r) {
    if (b.is_zero()) throw std::invalid_argument("BigUint division by zero");
    q = BigUint();
    r = BigUint();
    if (a.compare(b) < 0) {
        r = a;
        return;
    }
    // Restoring division one bit at a time; operands are a few thousand bits at most.
    const size_t nb = a.bits();
    q.limbs_.assign((nb + 63) / 64, 0);
    for (size_t i = nb; i-- > 0;) {
        r = r << 1;
        if (a.test_bit(i)) {
            if (r.limbs_.empty()) r.limbs_.push_back(1);
            else r.limbs_[0] |= 1;
        }
        if (r.compare(b) >= 0) {
            r = r - b;
            q.limbs_[i / 64] |= (u64)1 << (i % 64);
        }
    }
    q.normalize();
}

BigUint BigUint::operator/(const BigUint& o) const {
    BigUint q, r;
    divmod(*this, o, q, r);
    return q;
}

BigUint BigUint::operator%(const BigUint& o) const {
    BigUint q, r;
    divmod(*this, o, q, r);
    return r;
}

// ---------------------------------------------------------------------------
void RnsBasis::init(const std::vector<u64>& m) {
    if (m.empty()) throw std::invalid_argument("The list of moduli is empty");
    moduli = m;
    product = BigUint(1);
    for (size_t i = 0; i < m.size(); i++) {
        for (size_t j = i + 1; j < m.size(); j++) {
            u64 a = m[i], b = m[j];
            while (b) { u64 t = a % b; a = b; b = t; }
            if (a != 1) throw std::invalid_argument("The moduli are not coprime");
        }
        product = product * BigUint(m[i]);
    }
    garner.clear();
    for (u64 qi : m) {
        BigUint qstar = product / BigUint(qi);
        u64 qtilde = invmod64(qstar.mod_u64(qi), qi);
        garner.push_back(qstar * BigUint(qtilde));
    }
}

// extract_projection_and_theta, math/src/rns/scaler.rs:183-229
static void projection_and_theta(const RnsBasis& to, const BigUint& input, const BigUint& num, const BigUint& den,
                                 bool round_up, std::vector<u64>& proj, u64& lo, u64& hi, bool& sign) {
    BigUint ni = num * input;
    BigUint half = den >> 1;
    BigUint gamma = (ni + half) / den;
    proj.clear();
    for (u64 q : to.moduli) proj.push_back(gamma.mod_u64(q));
    BigUint theta = ni % den;
    sign = false;
    if (BigUint(1) < den) {
        bool odd = den.test_bit(0);
        bool flip = odd ? (half < theta) : !(theta < half);
        if (flip) {
            sign = true;
            theta = den - theta;
        }
    }
    BigUint shifted = theta << 127;
    bool ceil_div = round_up ? !sign : sign;
    theta = ceil_div ? (shifted + den - BigUint(1)) / den : shifted / den;
    lo = theta.limb(0);
    hi = theta.limb(1);
}

void ScalerConsts::build(const RnsBasis& from, const RnsBasis& to, const BigUint& num, const BigUint& den) {
    if (den.is_zero()) throw std::invalid_argument("zero denominator");
    is_one = (num == den);
    const size_t nf = from.moduli.size(), nt = to.moduli.size();
    projection_and_theta(to, from.product, num, den, false, gamma, theta_gamma_lo, theta_gamma_hi, theta_gamma_sign);
    gamma_shoup.resize(nt);
    for (size_t j = 0; j < nt; j++) gamma_shoup[j] = shoup64(gamma[j], to.moduli[j]);
    omega.assign(nt * nf, 0);
    omega_shoup.assign(nt * nf, 0);
    theta_omega_lo.resize(nf);
    theta_omega_hi.resize(nf);
    theta_omega_sign.resize(nf);
    for (size_t i = 0; i < nf; i++) {
        std::vector<u64> proj;
        bool sg;
        projection_and_theta(to, from.garner[i], num, den, true, proj, theta_omega_lo[i], theta_omega_hi[i], sg);
        theta_omega_sign[i] = sg;
        for (size_t j = 0; j < nt; j++) {
            u64 q = to.moduli[j];
            u64 o = proj[j] % q;
            omega[j * nf + i] = o;
            omega_shoup[j * nf + i] = shoup64(o, q);
        }
    }
    // shift = min(191 - log2(next_pow2(q_i * |from|)), 127)  (scaler.rs:130-142)
    u32 shift = 127;
    for (u64 q : from.moduli) {
        u128 x = (u128)q * nf;
        u32 lg = 0;
        while (((u128)1 << lg) < x) ++lg;
        shift = std::min<u32>(shift, 191 - lg);
    }
    theta_garner_shift = shift;
    theta_garner_lo.resize(nf);
    theta_garner_hi.resize(nf);
    for (size_t i = 0; i < nf; i++) {
        BigUint th = ((from.garner[i] << shift) + (from.product >> 1)) / from.product;
        theta_garner_lo[i] = th.limb(0);
        theta_garner_hi[i] = th.limb(1);
    }
}

}  // namespace fhehip
