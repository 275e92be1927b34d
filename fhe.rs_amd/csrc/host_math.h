// Host-side exact arithmetic used to build the resident tables: 64-bit
// modular helpers, primality, NTT-friendly prime generation, an arbitrary
// precision unsigned integer for CRT constants, and the RnsScaler constants
// of math/src/rns/scaler.rs:79-175.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace fhehip {

typedef uint64_t u64;
typedef uint32_t u32;
typedef unsigned __int128 u128;

u64 mulmod64(u64 a, u64 b, u64 m);
u64 powmod64(u64 a, u64 e, u64 m);
u64 invmod64(u64 a, u64 m);  // gcd(a, m) == 1 required; extended Euclid
u64 shoup64(u64 a, u64 p);   // floor(a * 2^64 / p)
bool is_prime64(u64 n);
bool supports_ntt64(u64 p, u64 n);
// math/src/zq/primes.rs:30-59 ; 0 when no prime exists
u64 generate_prime64(u32 num_bits, u64 modulo, u64 upper_bound);
// Deterministic primitive 2n-th root (first candidate g = 2, 3, ... whose
// power g^((p-1)/2n) has order exactly 2n).
u64 primitive_root64(u64 n, u64 p);
void barrett128(u64 p, u64& lo, u64& hi);  // floor(2^128 / p)
int ilog2_u64(u64 x);
u32 bit_reverse(u32 x, int bits);

// Arbitrary precision unsigned integer, little-endian 64-bit limbs.
class BigUint {
public:
    BigUint() {}
    BigUint(u64 v) { if (v) limbs_.push_back(v); }
    static BigUint from_words32(const uint32_t* w, size_t n);
    bool is_zero() const { return limbs_.empty(); }
    size_t bits() const;
    bool test_bit(size_t i) const;
    u64 limb(size_t i) const { return i < limbs_.size() ? limbs_[i] : 0; }
    u64 mod_u64(u64 m) const;
    int compare(const BigUint& o) const;
    bool operator==(const BigUint& o) const { return compare(o) == 0; }
    bool operator<(const BigUint& o) const { return compare(o) < 0; }
    BigUint operator+(const BigUint& o) const;
    BigUint operator-(const BigUint& o) const;  // requires *this >= o
    BigUint operator*(const BigUint& o) const;
    BigUint operator<<(size_t s) const;
    BigUint operator>>(size_t s) const;
    BigUint operator/(const BigUint& o) const;
    BigUint operator%(const BigUint& o) const;
    static void divmod(const BigUint& a, const BigUint& b, BigUint& q, BigUint& r);
    std::vector<uint32_t> words32() const;

private:
    std::vector<u64> limbs_;
    void normalize();
};

// CRT data of a basis (math/src/rns/mod.rs:24-113).
struct RnsBasis {
    std::vector<u64> moduli;
    BigUint product;
    std::vector<BigUint> garner;  // (Q / q_i) * ((Q / q_i)^-1 mod q_i)
    void init(const std::vector<u64>& m);
};

// Constants of an RnsScaler from `from` to `to` by numerator/denominator
// (math/src/rns/scaler.rs:79-175).
struct ScalerConsts {
    bool is_one = true;
    std::vector<u64> gamma, gamma_shoup;             // [to]
    u64 theta_gamma_lo = 0, theta_gamma_hi = 0;
    bool theta_gamma_sign = false;
    std::vector<u64> omega, omega_shoup;             // [to][from]
    std::vector<u64> theta_omega_lo, theta_omega_hi; // [from]
    std::vector<uint8_t> theta_omega_sign;           // [from]
    std::vector<u64> theta_garner_lo, theta_garner_hi;
    u32 theta_garner_shift = 0;
    void build(const RnsBasis& from, const RnsBasis& to, const BigUint& num, const BigUint& den);
};

}  // namespace fhehip
