// ORACLE — test infrastructure only (see oracle_core.hpp).
//
// The pseudo-random streams fhe.rs draws from, restated from the published
// algorithms of the crates it pins (Cargo.lock: sha2 0.10.9, rand_chacha
// 0.9.0, rand_core 0.9.3, rand 0.9.2; none of them is vendored under the
// reference):
//   sha256            FIPS 180-4; the round constants and initial hash are
//                     derived here from the primes (cube and square roots)
//                                                    -> pinned vs hashlib
//   chacha_block      Bernstein's ChaCha, 64-bit block counter in words 12-13,
//                     stream 0 in words 14-15; ChaCha8 = 4 double rounds
//                                                    -> pinned (20 rounds) vs
//                                                       RFC 7539 2.3.2 / A.1
//   ChaCha8Rng        rand_core's BlockRng over 4 blocks (64 words):
//                     next_u32 / next_u64 / fill_bytes
//   pcg32_seed        rand_core SeedableRng::seed_from_u64 (PCG32 XSH-RR with
//                     the state advanced before each output)
//                                                    -> output function and
//                                                       LCG step pinned vs the
//                                                       PCG demo sequence
//   uniform_sample    rand 0.9 UniformInt<u64>::sample (widening multiply,
//                     reject lo < 2^64 mod range)    -> parity unpinned
//   uniform_single    rand 0.9 UniformInt<u64>::sample_single (Canon's method,
//                     one extra draw when lo > -range) -> parity unpinned
// Users in fhe.rs: Poly::random_from_seed (math/src/rq/mod.rs:241-257),
// KeySwitchingKey::generate_c1 (fhe/src/bfv/keys/key_switching_key.rs:108-123),
// NttOperator::primitive_root (math/src/ntt/native.rs:320-336).
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <vector>

namespace oracle_rng {

using u8 = uint8_t;
using u32 = uint32_t;
using u64 = uint64_t;
using u128 = unsigned __int128;

// ---- SHA-256 ---------------------------------------------------------------
inline bool small_prime(u32 x) {
    if (x < 2) return false;
    for (u32 d = 2; d * d <= x; d++)
        if (x % d == 0) return false;
    return true;
}

// floor(cbrt(v)) and floor(sqrt(v)) of 128-bit values, by bisection.
inline u128 icbrt128(u128 v) {
    u128 lo = 0, hi = (u128)1 << 43;
    while (lo < hi) {
        u128 mid = (lo + hi + 1) / 2;
        if (mid * mid * mid <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
inline u128 isqrt128(u128 v) {
    u128 lo = 0, hi = (u128)1 << 64;
    while (lo < hi) {
        u128 mid = (lo + hi + 1) / 2;
        if (mid * mid <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct Sha256Consts {
    u32 k[64], h0[8];
    Sha256Consts() {
        // first 32 bits of the fractional parts of the cube roots of the first
        // 64 primes, and of the square roots of the first 8
        int found = 0;
        for (u32 x = 2; found < 64; x++) {
            if (!small_prime(x)) continue;
            k[found] = (u32)icbrt128((u128)x << 96);
            if (found < 8) h0[found] = (u32)isqrt128((u128)x << 64);
            found++;
        }
    }
};

inline u32 rotr(u32 x, int n) { return (x >> n) | (x << (32 - n)); }

inline std::array<u8, 32> sha256(const u8* data, size_t len) {
    static const Sha256Consts C;
    u32 h[8];
    std::memcpy(h, C.h0, sizeof h);
    std::vector<u8> m(data, data + len);
    const u64 bits = (u64)len * 8;
    m.push_back(0x80);
    while (m.size() % 64 != 56) m.push_back(0);
    for (int i = 7; i >= 0; i--) m.push_back((u8)(bits >> (8 * i)));
    for (size_t off = 0; off < m.size(); off += 64) {
        u32 w[64];
        for (int i = 0; i < 16; i++)
            w[i] = ((u32)m[off + 4 * i] << 24) | ((u32)m[off + 4 * i + 1] << 16) | ((u32)m[off + 4 * i + 2] << 8) |
                   m[off + 4 * i + 3];
        for (int i = 16; i < 64; i++) {
            const u32 s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
            const u32 s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        u32 a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
        for (int i = 0; i < 64; i++) {
            const u32 t1 = hh + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + C.k[i] + w[i];
            const u32 t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
            hh = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = b;
            b = a;
            a = t1 + t2;
        }
        h[0] += a, h[1] += b, h[2] += c, h[3] += d, h[4] += e, h[5] += f, h[6] += g, h[7] += hh;
    }
    std::array<u8, 32> out;
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 4; j++) out[4 * i + j] = (u8)(h[i] >> (24 - 8 * j));
    return out;
}

// ---- ChaCha ----------------------------------------------------------------
inline u32 rotl(u32 x, int n) { return (x << n) | (x >> (32 - n)); }

inline void quarter(u32* s, int a, int b, int c, int d) {
    s[a] += s[b], s[d] ^= s[a], s[d] = rotl(s[d], 16);
    s[c] += s[d], s[b] ^= s[c], s[b] = rotl(s[b], 12);
    s[a] += s[b], s[d] ^= s[a], s[d] = rotl(s[d], 8);
    s[c] += s[d], s[b] ^= s[c], s[b] = rotl(s[b], 7);
}

// One block of the keystream: the 16 state words after `rounds` rounds,
// added to the input state.
inline void chacha_block(const u32 in[16], int rounds, u32 out[16]) {
    u32 s[16];
    std::memcpy(s, in, sizeof s);
    for (int r = 0; r < rounds; r += 2) {
        quarter(s, 0, 4, 8, 12), quarter(s, 1, 5, 9, 13), quarter(s, 2, 6, 10, 14), quarter(s, 3, 7, 11, 15);
        quarter(s, 0, 5, 10, 15), quarter(s, 1, 6, 11, 12), quarter(s, 2, 7, 8, 13), quarter(s, 3, 4, 9, 14);
    }
    for (int i = 0; i < 16; i++) out[i] = s[i] + in[i];
}

// "expand 32-byte k" as little-endian words
inline void chacha_constants(u32* w) {
    const char* sigma = "expand 32-byte k";
    for (int i = 0; i < 4; i++)
        w[i] = (u32)(u8)sigma[4 * i] | ((u32)(u8)sigma[4 * i + 1] << 8) | ((u32)(u8)sigma[4 * i + 2] << 16) |
               ((u32)(u8)sigma[4 * i + 3] << 24);
}

struct ChaCha8Rng {
    u32 key[8];
    u64 counter = 0;  // next block
    u32 results[64];
    u32 index = 64;   // empty buffer

    explicit ChaCha8Rng(const std::array<u8, 32>& seed) {
        for (int i = 0; i < 8; i++)
            key[i] = (u32)seed[4 * i] | ((u32)seed[4 * i + 1] << 8) | ((u32)seed[4 * i + 2] << 16) |
                     ((u32)seed[4 * i + 3] << 24);
    }
    void generate() {
        for (int b = 0; b < 4; b++) {
            u32 st[16];
            chacha_constants(st);
            std::memcpy(st + 4, key, sizeof key);
            st[12] = (u32)counter, st[13] = (u32)(counter >> 32);
            st[14] = st[15] = 0;
            chacha_block(st, 8, results + 16 * b);
            counter++;
        }
    }
    u32 next_u32() {
        if (index >= 64) generate(), index = 0;
        return results[index++];
    }
    u64 next_u64() {
        if (index < 63) {
            const u64 v = (u64)results[index] | ((u64)results[index + 1] << 32);
            index += 2;
            return v;
        }
        if (index >= 64) {
            generate();
            index = 2;
            return (u64)results[0] | ((u64)results[1] << 32);
        }
        const u64 lo = results[63];
        generate();
        index = 1;
        return lo | ((u64)results[0] << 32);
    }
    void fill_bytes(u8* dst, size_t len) {
        size_t done = 0;
        while (done < len) {
            if (index >= 64) generate(), index = 0;
            const u32 w = results[index++];
            for (int j = 0; j < 4 && done < len; j++) dst[done++] = (u8)(w >> (8 * j));
        }
    }
};

// rand_core SeedableRng::seed_from_u64: PCG32 fills the 32-byte seed.
inline u32 pcg32_output(u64 state) {
    const u32 xorshifted = (u32)(((state >> 18) ^ state) >> 27);
    const u32 rot = (u32)(state >> 59);
    return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
}
constexpr u64 kPcgMul = 6364136223846793005ull;
constexpr u64 kPcgInc = 11634580027462260723ull;

inline std::array<u8, 32> pcg32_seed(u64 state) {
    std::array<u8, 32> seed{};
    for (int i = 0; i < 8; i++) {
        state = state * kPcgMul + kPcgInc;
        const u32 x = pcg32_output(state);
        for (int j = 0; j < 4; j++) seed[4 * i + j] = (u8)(x >> (8 * j));
    }
    return seed;
}

// ---- rand 0.9 uniform integers ------------------------------------------------
// Uniform::new(0, range) then sample: hi of the widening product, rejecting
// lo < 2^64 mod range.
inline u64 uniform_sample(ChaCha8Rng& rng, u64 range) {
    const u64 thresh = (0 - range) % range;
    for (;;) {
        const u128 m = (u128)rng.next_u64() * range;
        if ((u64)m >= thresh) return (u64)(m >> 64);
    }
}

// random_range(low..high): Canon's method.
inline u64 uniform_single(ChaCha8Rng& rng, u64 low, u64 high) {
    const u64 range = high - low;
    const u128 m = (u128)rng.next_u64() * range;
    u64 result = (u64)(m >> 64);
    const u64 lo = (u64)m;
    if (lo > 0 - range) {
        const u64 hi2 = (u64)(((u128)rng.next_u64() * range) >> 64);
        result += (lo + hi2 < lo) ? 1 : 0;
    }
    return low + result;
}

// Poly::random_from_seed: the seed hashed with SHA-256 seeds ChaCha8Rng, then
// limb by limb, N uniform values below q_i (NTT representation).
inline std::vector<u64> random_from_seed(const std::vector<u64>& moduli, size_t n, const std::array<u8, 32>& seed) {
    ChaCha8Rng rng(sha256(seed.data(), 32));
    std::vector<u64> out(moduli.size() * n);
    for (size_t l = 0; l < moduli.size(); l++)
        for (size_t j = 0; j < n; j++) out[l * n + j] = uniform_sample(rng, moduli[l]);
    return out;
}

// KeySwitchingKey::generate_c1: `size` polynomials, each from a 32-byte seed
// filled from ChaCha8Rng(seed).
inline std::vector<u64> generate_c1(const std::vector<u64>& moduli, size_t n, const std::array<u8, 32>& seed,
                                    size_t size) {
    ChaCha8Rng rng(seed);
    std::vector<u64> out;
    out.reserve(size * moduli.size() * n);
    for (size_t i = 0; i < size; i++) {
        std::array<u8, 32> si;
        rng.fill_bytes(si.data(), 32);
        std::vector<u64> p = random_from_seed(moduli, n, si);
        out.insert(out.end(), p.begin(), p.end());
    }
    return out;
}

}  // namespace oracle_rng
