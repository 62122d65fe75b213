// ORACLE — test infrastructure only (see field.h).
//
// Seeded randomness of the reference: `common::test::rng(seed)` =
// `StdRng::seed_from_u64(seed)` (crates/common/src/test.rs:1-4), rand 0.10 whose
// StdRng is ChaCha12 (crate chacha20 0.10, Cargo.lock:66,686).
//   * seed_from_u64: rand_core's PCG32 expansion (mul 6364136223846793005,
//     inc 11634580027462260723, xorshift + rotate output) fills the 32-byte key.
//   * ChaCha block: djb layout, 64-bit block counter in words 12-13, stream 0.
//     The keystream is consumed as consecutive little-endian u32 words.
//   * Sampling a field element (Plonky3 StandardUniform for MontyField31):
//     loop { v = next_u32() >> 1; if v < p return from_monty(v) }.
// Pinning: the ChaCha block function (20-round form) is checked against the
// RFC 8439 / RFC 7539 known-answer vectors, and the PCG32 multiplier and XSH-RR
// output function against the PCG reference implementation's demo outputs
// (seed 42, sequence 54), in tests/test_oracle_rng.py. The 12-round variant,
// seed_from_u64's increment and step-before-output order, and the sampling loop
// are restated from the crates' documented algorithms: parity unpinned (no
// reference vectors exist and the crates are absent).
#pragma once
#include <cstdint>
#include "field.h"

namespace oracle {

void chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream, int rounds,
                  uint32_t out[16]);

struct StdRng {
  uint32_t key[8];
  uint64_t counter = 0;
  uint32_t buf[16];
  int idx = 16;
  int rounds = 12;

  static StdRng seed_from_u64(uint64_t seed);
  uint32_t next_u32();
  F next_f();                // StandardUniform<KoalaBear>
  E next_e();                // StandardUniform<BinomialExtensionField<_,4>>
};

}  // namespace oracle
