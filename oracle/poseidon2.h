// ORACLE — test infrastructure only (see field.h).
//
// Poseidon2 over KoalaBear, widths 16 and 24, S-box x^3, as instantiated by the
// reference: `Poseidon2KoalaBear<W>::new_from_rng_128(&mut rng(1000))`
// (crates/whir/src/test.rs:195-196), hasher `PaddingFreeSponge<P24,24,16,8>`
// (:188) and compressor `TruncatedPermutation<P16,2,8,16>` (:189).
// Restated from Plonky3 p3-poseidon2 / p3-monty-31 / p3-koala-bear @ e52636ec:
//   * rounds: R_F = 8 (4 initial + 4 terminal), R_P = 20 (W16) / 23 (W24);
//   * constants drawn from the rng in the order: 4 initial external rounds
//     (W values each), 4 terminal external rounds, R_P internal constants;
//   * external linear layer "MDS light": M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],
//     [3,1,1,2]] on each 4-chunk, then x_i += sum_j x_{4j + i%4}; applied once
//     before the first round and after every external round;
//   * internal linear layer 1*1^T + diag(V) with V from Plonky3's KoalaBear
//     parameters (INTERNAL_V16 / INTERNAL_V24 below, canonical fractions).
// Pinning: the linear layers are checked against their explicit matrices and the
// permutation against an independent inverse (tests/test_oracle_poseidon2.py);
// the internal vectors V, the round counts and the constant stream are recalled
// from Plonky3 and cannot be imported here: parity unpinned.
#pragma once
#include <cstdint>
#include <vector>
#include "field.h"

namespace oracle {

struct Poseidon2 {
  int width = 0;
  int rounds_p = 0;
  std::vector<F> ext_initial;   // 4 * width
  std::vector<F> ext_terminal;  // 4 * width
  std::vector<F> internal;      // rounds_p
  std::vector<F> diag;          // width (V)

  static Poseidon2 make(int width, uint64_t seed);
  void permute(F* state) const;
  void mds_light(F* state) const;
  void internal_layer(F* state) const;
};

// Canonical diagonal entries as numerator/2^shift with sign.
void internal_diag_canonical(int width, std::vector<F>& out);

struct Hashers {
  Poseidon2 p16, p24;
  static const Hashers& get();  // seed 1000, the reference's instantiation
};

// PaddingFreeSponge<P24, 24, 16, 8>::hash_iter over `n` base elements.
void sponge24(const Poseidon2& p24, const F* in, size_t n, F out[8]);
// TruncatedPermutation<P16, 2, 8, 16>::compress([l, r]).
void compress16(const Poseidon2& p16, const F l[8], const F r[8], F out[8]);

}  // namespace oracle
