// Pseudo-random streams of fhe.rs's seeded values, on the host: SHA-256
// (sha2), ChaCha8Rng (rand_chacha 0.9, BlockRng of rand_core 0.9),
// SeedableRng::seed_from_u64 (rand_core 0.9) and the uniform integer draws of
// rand 0.9 (Uniform::sample, random_range). Used for Poly::random_from_seed
// (math/src/rq/mod.rs:241-257), KeySwitchingKey::generate_c1
// (fhe/src/bfv/keys/key_switching_key.rs:108-123) and the default root of
// NttOperator::new (math/src/ntt/native.rs:320-336).
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace fhehip {

using Seed32 = std::array<uint8_t, 32>;

Seed32 sha256(const uint8_t* data, size_t len);

// ChaCha with 8 rounds, 64-bit block counter, stream 0, consumed as
// little-endian 32-bit words in keystream order.
class ChaCha8Stream {
  public:
    explicit ChaCha8Stream(const Seed32& key);
    static ChaCha8Stream from_u64(uint64_t state);  // SeedableRng::seed_from_u64
    uint32_t next_u32();
    uint64_t next_u64();
    void fill(uint8_t* dst, size_t len);  // whole words, little-endian bytes
    // Uniform::new(0, range).sample(rng) and rng.random_range(0..range).
    uint64_t uniform(uint64_t range);
    uint64_t uniform_single(uint64_t range);

  private:
    void refill();
    uint32_t key_[8];
    uint64_t block_ = 0;
    uint32_t words_[64];
    unsigned pos_ = 64;
};

// One ChaCha block (rounds even) of the 16-word input state.
void chacha_block(const uint32_t in[16], int rounds, uint32_t out[16]);

// Poly::random_from_seed into rows [L][n] (NTT-domain values below q_l).
void random_from_seed(const std::vector<uint64_t>& moduli, uint32_t n, const Seed32& seed, uint64_t* out);
// KeySwitchingKey::generate_c1: `count` polynomials [count][L][n].
void generate_c1(const std::vector<uint64_t>& moduli, uint32_t n, const Seed32& seed, uint32_t count, uint64_t* out);
// NttOperator's primitive 2n-th root of p (0 when 100 draws find none).
uint64_t fhers_primitive_root(uint32_t n, uint64_t p);

}  // namespace fhehip
