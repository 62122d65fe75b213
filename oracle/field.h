// ORACLE — test infrastructure only. Imported by tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg as the checker; never part of the product path.
//
// KoalaBear prime field (p = 2^31 - 2^24 + 1) in Montgomery form (R = 2^32) and
// its degree-4 binomial extension F[x]/(x^4 - 3).
//
// Restates Plonky3 p3-monty-31 / p3-koala-bear / p3-field BinomialExtensionField
// (pinned dependency Plonky3 @ e52636ec, Cargo.lock:345-535), which the reference
// re-exports at crates/common/src/field.rs:1-12.
//   * Elements are stored as the Montgomery representative x*2^32 mod p, in [0,p).
//     That is the value Plonky3 serializes and the value `StandardUniform`
//     samples (`new_monty(next_u32() >> 1)`, rejection above p).
//   * Pinning: arithmetic is pinned by closed form (tests/test_oracle_field.py
//     checks add/mul/inv/pow against Python big integers); W = 3 is checked to be
//     a quadratic non-residue so x^4 - 3 is irreducible; two-adic generators are
//     checked against 3^((p-1)/2^b).
#pragma once
#include <cstdint>
#include <cstddef>

namespace oracle {

#if defined(ORACLE_BABYBEAR)
// BabyBear (p = 15 * 2^27 + 1), the reference's second field
// (crates/transcript/src/lib.rs:389-396): GENERATOR 31, two-adicity 27, x^4 = 11.
constexpr uint32_t P = 0x78000001u;
constexpr uint32_t GENERATOR_CANON = 31;
#else
constexpr uint32_t P = 0x7f000001u;
constexpr uint32_t GENERATOR_CANON = 3;
#endif
// -p^{-1} mod 2^32, used by the Montgomery reduction below.
constexpr uint32_t P_NEG_INV = []() constexpr {
  uint32_t inv = 1;
  for (int i = 0; i < 5; ++i) inv *= 2u - P * inv;  // Newton: inv = p^{-1} mod 2^32
  return 0u - inv;
}();

inline uint32_t mont_reduce(uint64_t x) {
  // x < p * 2^32.  m = x * (-p^{-1}) mod 2^32 makes x + m*p divisible by 2^32.
  uint32_t m = (uint32_t)x * P_NEG_INV;
  uint64_t t = (x + (uint64_t)m * P) >> 32;  // < 2p, fits (x < 2^63)
  return t >= P ? (uint32_t)(t - P) : (uint32_t)t;
}

struct F {
  uint32_t v;  // Montgomery representative
  static F from_monty(uint32_t m) { return F{m}; }
  static F from_u32(uint32_t c) {  // canonical -> Montgomery
    return F{(uint32_t)(((uint64_t)(c % P) << 32) % P)};
  }
  uint32_t canonical() const { return mont_reduce(v); }
  bool operator==(const F& o) const { return v == o.v; }
  bool operator!=(const F& o) const { return v != o.v; }
};

inline F fadd(F a, F b) {
  uint32_t s = a.v + b.v;
  return F{s >= P ? s - P : s};
}
inline F fsub(F a, F b) { return F{a.v >= b.v ? a.v - b.v : a.v + P - b.v}; }
inline F fneg(F a) { return F{a.v == 0 ? 0 : P - a.v}; }
inline F fmul(F a, F b) { return F{mont_reduce((uint64_t)a.v * b.v)}; }
inline F fdbl(F a) { return fadd(a, a); }
inline F fsq(F a) { return fmul(a, a); }
inline F fpow(F a, uint64_t e) {
  F r = F::from_u32(1);
  while (e) {
    if (e & 1) r = fmul(r, a);
    a = fmul(a, a);
    e >>= 1;
  }
  return r;
}
inline F finv(F a) { return fpow(a, P - 2); }
inline F fzero() { return F{0}; }
inline F fone() { return F::from_u32(1); }

// Plonky3 KoalaBear: GENERATOR = 3, TWO_ADICITY = 24,
// two_adic_generator(bits) = 3^((p-1)/2^bits).
F two_adic_generator(unsigned bits);

// ---- degree-4 extension, x^4 = W with W = 3 (BabyBear: 11) ----
#if defined(ORACLE_BABYBEAR)
constexpr uint32_t W_CANON = 11;
#else
constexpr uint32_t W_CANON = 3;
#endif

struct E {
  F c[4];
  bool operator==(const E& o) const {
    return c[0] == o.c[0] && c[1] == o.c[1] && c[2] == o.c[2] && c[3] == o.c[3];
  }
  bool operator!=(const E& o) const { return !(*this == o); }
};

inline E ezero() { return E{{F{0}, F{0}, F{0}, F{0}}}; }
inline E efrom(F a) { return E{{a, F{0}, F{0}, F{0}}}; }
inline E eone() { return efrom(fone()); }
inline E eadd(const E& a, const E& b) {
  return E{{fadd(a.c[0], b.c[0]), fadd(a.c[1], b.c[1]), fadd(a.c[2], b.c[2]), fadd(a.c[3], b.c[3])}};
}
inline E esub(const E& a, const E& b) {
  return E{{fsub(a.c[0], b.c[0]), fsub(a.c[1], b.c[1]), fsub(a.c[2], b.c[2]), fsub(a.c[3], b.c[3])}};
}
inline E eneg(const E& a) { return E{{fneg(a.c[0]), fneg(a.c[1]), fneg(a.c[2]), fneg(a.c[3])}}; }
inline E edbl(const E& a) { return eadd(a, a); }
inline E emul_base(const E& a, F b) {
  return E{{fmul(a.c[0], b), fmul(a.c[1], b), fmul(a.c[2], b), fmul(a.c[3], b)}};
}
E emul(const E& a, const E& b);
inline E esq(const E& a) { return emul(a, a); }
E epow(E a, uint64_t e);

// bincode bytes (Montgomery u32, little endian): 4 B per base element, 16 B per
// extension element (crates/transcript/src/field.rs:6-25).
inline void put_u32(uint8_t* out, uint32_t v) {
  out[0] = (uint8_t)v;
  out[1] = (uint8_t)(v >> 8);
  out[2] = (uint8_t)(v >> 16);
  out[3] = (uint8_t)(v >> 24);
}
inline uint32_t get_u32(const uint8_t* in) {
  return (uint32_t)in[0] | ((uint32_t)in[1] << 8) | ((uint32_t)in[2] << 16) | ((uint32_t)in[3] << 24);
}

}  // namespace oracle
