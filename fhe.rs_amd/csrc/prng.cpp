// Host pseudo-random streams of fhe.rs's seeded values (see prng.h).
#include "prng.h"

#include <cstring>

#include "host_math.h"

namespace fhehip {

namespace {

using u128 = unsigned __int128;

// Integer k-th root (k = 2 or 3) of v < 2^105: the largest r with r^k <= v,
// by bisection over r < 2^43.
u64 iroot(u128 v, int k) {
    u64 lo = 0, hi = (u64)1 << 43;
    while (lo < hi) {
        const u64 mid = lo + (hi - lo + 1) / 2;
        u128 p = mid;
        bool over = false;
        for (int i = 1; i < k && !over; i++) {
            if (p > (~(u128)0) / mid) over = true;
            else p *= mid;
        }
        if (!over && p <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct ShaTables {
    uint32_t k[64];
    uint32_t h[8];
    ShaTables() {
        // SHA-256's constants: the first 32 fractional bits of the cube roots
        // (k) and square roots (h) of the leading primes.
        int n = 0;
        for (uint32_t c = 2; n < 64; c++) {
            bool prime = true;
            for (uint32_t d = 2; d * d <= c; d++) prime = prime && (c % d != 0);
            if (!prime) continue;
            k[n] = (uint32_t)iroot((u128)c << 96, 3);
            if (n < 8) h[n] = (uint32_t)iroot((u128)c << 64, 2);
            n++;
        }
    }
};

inline uint32_t ror32(uint32_t x, unsigned r) { return (x >> r) | (x << (32 - r)); }
inline uint32_t rol32(uint32_t x, unsigned r) { return (x << r) | (x >> (32 - r)); }

inline uint32_t load_le32(const uint8_t* b) {
    return (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
}

void sha_compress(uint32_t st[8], const uint8_t blk[64], const uint32_t* K) {
    uint32_t w[64];
    for (int t = 0; t < 16; t++)
        w[t] = ((uint32_t)blk[4 * t] << 24) | ((uint32_t)blk[4 * t + 1] << 16) | ((uint32_t)blk[4 * t + 2] << 8) |
               (uint32_t)blk[4 * t + 3];
    for (int t = 16; t < 64; t++)
        w[t] = (ror32(w[t - 2], 17) ^ ror32(w[t - 2], 19) ^ (w[t - 2] >> 10)) + w[t - 7] +
               (ror32(w[t - 15], 7) ^ ror32(w[t - 15], 18) ^ (w[t - 15] >> 3)) + w[t - 16];
    uint32_t v[8];
    std::memcpy(v, st, sizeof v);
    for (int t = 0; t < 64; t++) {
        const uint32_t ch = (v[4] & v[5]) ^ (~v[4] & v[6]);
        const uint32_t maj = (v[0] & v[1]) ^ (v[0] & v[2]) ^ (v[1] & v[2]);
        const uint32_t x = v[7] + (ror32(v[4], 6) ^ ror32(v[4], 11) ^ ror32(v[4], 25)) + ch + K[t] + w[t];
        const uint32_t y = (ror32(v[0], 2) ^ ror32(v[0], 13) ^ ror32(v[0], 22)) + maj;
        std::memmove(v + 1, v, 7 * sizeof(uint32_t));
        v[4] += x;
        v[0] = x + y;
    }
    for (int i = 0; i < 8; i++) st[i] += v[i];
}

}  // namespace

Seed32 sha256(const uint8_t* data, size_t len) {
    static const ShaTables T;
    uint32_t st[8];
    std::memcpy(st, T.h, sizeof st);
    size_t off = 0;
    for (; off + 64 <= len; off += 64) sha_compress(st, data + off, T.k);
    // final block(s): the rest, 0x80, zeros and the bit length (big-endian)
    uint8_t tail[128] = {};
    const size_t rest = len - off;
    std::memcpy(tail, data + off, rest);
    tail[rest] = 0x80;
    const size_t total = rest + 9 <= 64 ? 64 : 128;
    const uint64_t bits = (uint64_t)len * 8;
    for (int i = 0; i < 8; i++) tail[total - 1 - i] = (uint8_t)(bits >> (8 * i));
    for (size_t b = 0; b < total; b += 64) sha_compress(st, tail + b, T.k);
    Seed32 out;
    for (int i = 0; i < 32; i++) out[i] = (uint8_t)(st[i / 4] >> (24 - 8 * (i % 4)));
    return out;
}

void chacha_block(const uint32_t in[16], int rounds, uint32_t out[16]) {
    uint32_t x[16];
    std::memcpy(x, in, sizeof x);
    auto qr = [&](int a, int b, int c, int d) {
        x[a] += x[b];
        x[d] = rol32(x[d] ^ x[a], 16);
        x[c] += x[d];
        x[b] = rol32(x[b] ^ x[c], 12);
        x[a] += x[b];
        x[d] = rol32(x[d] ^ x[a], 8);
        x[c] += x[d];
        x[b] = rol32(x[b] ^ x[c], 7);
    };
    for (int r = 0; r < rounds / 2; r++) {
        for (int c = 0; c < 4; c++) qr(c, c + 4, c + 8, c + 12);         // columns
        for (int c = 0; c < 4; c++) qr(c, 4 + (c + 1) % 4, 8 + (c + 2) % 4, 12 + (c + 3) % 4);  // diagonals
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + in[i];
}

ChaCha8Stream::ChaCha8Stream(const Seed32& key) {
    for (int i = 0; i < 8; i++) key_[i] = load_le32(key.data() + 4 * i);
}

ChaCha8Stream ChaCha8Stream::from_u64(uint64_t state) {
    // PCG32 (XSH-RR output of the advanced state) fills the seed 4 bytes at a time.
    Seed32 seed;
    for (int i = 0; i < 8; i++) {
        state = state * 6364136223846793005ull + 11634580027462260723ull;
        const uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
        const unsigned rot = (unsigned)(state >> 59);
        const uint32_t v = rot ? ror32(xs, rot) : xs;
        for (int j = 0; j < 4; j++) seed[4 * i + j] = (uint8_t)(v >> (8 * j));
    }
    return ChaCha8Stream(seed);
}

void ChaCha8Stream::refill() {
    static const uint8_t sigma[] = "expand 32-byte k";
    uint32_t st[16];
    for (int i = 0; i < 4; i++) st[i] = load_le32(sigma + 4 * i);
    std::memcpy(st + 4, key_, sizeof key_);
    st[14] = st[15] = 0;
    for (int b = 0; b < 4; b++, block_++) {
        st[12] = (uint32_t)block_;
        st[13] = (uint32_t)(block_ >> 32);
        chacha_block(st, 8, words_ + 16 * b);
    }
    pos_ = 0;
}

uint32_t ChaCha8Stream::next_u32() {
    if (pos_ == 64) refill();
    return words_[pos_++];
}

uint64_t ChaCha8Stream::next_u64() {
    // two consecutive keystream words, the first low (across a refill too)
    const uint64_t lo = next_u32();
    return lo | ((uint64_t)next_u32() << 32);
}

void ChaCha8Stream::fill(uint8_t* dst, size_t len) {
    for (size_t i = 0; i < len; i += 4) {
        const uint32_t w = next_u32();
        for (size_t j = 0; j < 4 && i + j < len; j++) dst[i + j] = (uint8_t)(w >> (8 * j));
    }
}

uint64_t ChaCha8Stream::uniform(uint64_t range) {
    // widening multiply; the low half below 2^64 mod range is rejected
    const uint64_t reject_below = (0 - range) % range;
    for (;;) {
        const u128 m = (u128)next_u64() * range;
        if ((uint64_t)m >= reject_below) return (uint64_t)(m >> 64);
    }
}

uint64_t ChaCha8Stream::uniform_single(uint64_t range) {
    // one widening multiply; a second draw decides a carry into the high
    // half when the low half lies in the top `range` values
    const u128 m = (u128)next_u64() * range;
    uint64_t hi = (uint64_t)(m >> 64);
    const uint64_t lo = (uint64_t)m;
    if (lo > 0 - range) {
        const uint64_t hi2 = (uint64_t)(((u128)next_u64() * range) >> 64);
        if ((uint64_t)(lo + hi2) < lo) hi++;
    }
    return hi;
}

void random_from_seed(const std::vector<uint64_t>& moduli, uint32_t n, const Seed32& seed, uint64_t* out) {
    ChaCha8Stream rng(sha256(seed.data(), seed.size()));
    for (size_t l = 0; l < moduli.size(); l++)
        for (uint32_t j = 0; j < n; j++) out[l * n + j] = rng.uniform(moduli[l]);
}

void generate_c1(const std::vector<uint64_t>& moduli, uint32_t n, const Seed32& seed, uint32_t count, uint64_t* out) {
    ChaCha8Stream rng(seed);
    const size_t per = moduli.size() * (size_t)n;
    for (uint32_t i = 0; i < count; i++) {
        Seed32 si;
        rng.fill(si.data(), si.size());
        random_from_seed(moduli, n, si, out + i * per);
    }
}

uint64_t fhers_primitive_root(uint32_t n, uint64_t p) {
    const uint64_t lambda = (p - 1) / (2 * (uint64_t)n);
    ChaCha8Stream rng = ChaCha8Stream::from_u64(0);
    for (int i = 0; i < 100; i++) {
        const uint64_t root = powmod64(rng.uniform_single(p), lambda, p);
        if (powmod64(root, 2 * (uint64_t)n, p) == 1 && powmod64(root, n, p) != 1) return root;
    }
    return 0;
}

}  // namespace fhehip
