// Host-side generation of the Poseidon2 KoalaBear round constants exactly as
// `Poseidon2KoalaBear<W>::new_from_rng_128(&mut StdRng::seed_from_u64(seed))`
// (crates/whir/src/test.rs:195-196, crates/common/src/test.rs:1-4):
//   * StdRng (rand 0.10) = ChaCha12, key from rand_core's PCG32 seed_from_u64
//     expansion, 64-bit block counter, stream 0, keystream read as LE u32 words;
//   * StandardUniform<KoalaBear | BabyBear>: v = next_u32() >> 1, rejected unless v < p,
//     taken as the Montgomery representative;
//   * order: 4 initial external rounds (W each), 4 terminal, R_P internal.
#pragma once
#include <cstdint>
#include <cstring>

namespace whir_host {

class ChaChaRng {
 public:
  explicit ChaChaRng(uint64_t seed, int rounds = 12) : rounds_(rounds) {
    const uint64_t MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
    uint64_t st = seed;
    for (int i = 0; i < 8; ++i) {
      st = st * MUL + INC;
      const uint32_t xs = (uint32_t)(((st >> 18) ^ st) >> 27);
      const uint32_t rot = (uint32_t)(st >> 59);
      key_[i] = (xs >> rot) | (xs << ((32 - rot) & 31));
    }
  }
  uint32_t next_u32() {
    if (pos_ == 16) refill();
    return buf_[pos_++];
  }
  // StandardUniform of the build's field (kb::P): the Montgomery representative
  uint32_t next_field() {
    for (;;) {
      const uint32_t v = next_u32() >> 1;
      if (v < kb::P) return v;
    }
  }

 private:
  static uint32_t rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }
  void refill() {
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key_[0], key_[1], key_[2], key_[3],
                      key_[4], key_[5], key_[6], key_[7], (uint32_t)ctr_, (uint32_t)(ctr_ >> 32), 0u, 0u};
    uint32_t x[16];
    memcpy(x, s, sizeof(x));
    auto qr = [&](int a, int b, int c, int d) {
      x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16);
      x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 12);
      x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8);
      x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 7);
    };
    for (int r = 0; r < rounds_; r += 2) {
      qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
      qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
    }
    for (int i = 0; i < 16; ++i) buf_[i] = x[i] + s[i];
    ++ctr_;
    pos_ = 0;
  }
  uint32_t key_[8];
  uint64_t ctr_ = 0;
  uint32_t buf_[16];
  int pos_ = 16;
  int rounds_;
};

}  // namespace whir_host
