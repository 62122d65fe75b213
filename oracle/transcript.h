// ORACLE — test infrastructure only (see field.h).
//
// Fiat–Shamir transcript of the reference: PoseidonWriter / PoseidonReader over
// DuplexChallenger<KoalaBear, Poseidon2KoalaBear<16>, 16, 8>
// (crates/transcript/src/poseidon.rs:13-164; crates/whir/src/test.rs:190).
//   * write(x): append bincode bytes, then observe (Ext: its 4 coefficients in
//     order; digest: its 8 elements). write_hint(x): bytes only.
//   * DuplexChallenger (Plonky3 p3-challenger @ e52636ec, recalled):
//       observe(v): clear output buffer; push v; if input has RATE=8 -> duplex.
//       sample(): if input non-empty or output empty -> duplex; pop() from the END
//                 of the output buffer.
//       duplex: overwrite state[0..len(input)] with the input, permute,
//               output = state[0..8].
//       sample_algebra_element: 4 samples, coefficient 0 first.
//       sample_bits(b): canonical(sample()) & (2^b - 1).
// Pinning: prover/verifier self-consistency only; the buffer semantics are
// recalled from Plonky3: parity unpinned.
#pragma once
#include <cstdint>
#include <vector>
#include "field.h"
#include "poseidon2.h"

namespace oracle {

struct Challenger {
  const Poseidon2* perm;
  F state[16];
  std::vector<F> in, out;

  explicit Challenger(const Poseidon2* p) : perm(p) {
    for (auto& s : state) s = F{0};
  }
  void duplex();
  void observe(F v);
  F sample();
  E sample_ext();
  uint64_t sample_bits(unsigned bits);
};

enum Status { OK = 0, ERR_TRANSCRIPT = 1, ERR_VERIFY = 2, ERR_IO = 3 };

struct Writer {
  Challenger ch;
  std::vector<uint8_t> bytes;
  explicit Writer(const Poseidon2* p) : ch(p) {}

  void hint_f(F x);
  void write_f(F x);
  void hint_e(const E& x);
  void write_e(const E& x);
  void hint_digest(const F d[8]);
  void write_digest(const F d[8]);
  E draw_e() { return ch.sample_ext(); }
  F draw_f() { return ch.sample(); }
  uint64_t draw_bits(unsigned b) { return ch.sample_bits(b); }
};

struct Reader {
  Challenger ch;
  const uint8_t* data;
  size_t len, pos = 0;
  Reader(const Poseidon2* p, const uint8_t* d, size_t n) : ch(p), data(d), len(n) {}

  // Each returns OK or ERR_TRANSCRIPT (out of bytes / non-canonical value).
  int hint_f(F& x);
  int read_f(F& x);
  int hint_e(E& x);
  int read_e(E& x);
  int hint_digest(F d[8]);
  int read_digest(F d[8]);
  E draw_e() { return ch.sample_ext(); }
  F draw_f() { return ch.sample(); }
  uint64_t draw_bits(unsigned b) { return ch.sample_bits(b); }
};

}  // namespace oracle
