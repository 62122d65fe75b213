// KoalaBear field (p = 2^31 - 2^24 + 1) and its quartic extension x^4 = 3, for
// device (and host) code; built with WHIR_FIELD_BABYBEAR, BabyBear
// (p = 15 * 2^27 + 1) and x^4 = 11 (the reference's second field,
// crates/transcript/src/lib.rs:389-396). Elements are Montgomery
// representatives in [0, p), the representation the reference serializes
// (Plonky3 MontyField31).
//
// Costs on gfx950 VALU (one lane): mul = 2 wide multiplies + 1 mul_lo + 2 ALU;
// add/sub = 3 ALU (add, sub, min). Ext mul = 16 base mul.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "job.h"

#define KB_HD __host__ __device__ __forceinline__

namespace kb {

#if defined(WHIR_FIELD_BABYBEAR)
constexpr uint32_t P = 0x78000001u;      // 15 * 2^27 + 1
constexpr uint32_t GENERATOR = 31;       // Plonky3 BabyBear GENERATOR (two-adicity 27)
constexpr uint32_t EXT_W = 11;           // x^4 = 11 (the smallest non-residue)
#else
constexpr uint32_t P = 0x7f000001u;      // 2^31 - 2^24 + 1
constexpr uint32_t GENERATOR = 3;        // Plonky3 KoalaBear GENERATOR (two-adicity 24)
constexpr uint32_t EXT_W = 3;            // x^4 = 3 (the smallest non-residue)
#endif
constexpr uint32_t neg_inv_p() {  // -p^{-1} mod 2^32 by Newton's iteration
  uint32_t inv = 1;
  for (int i = 0; i < 5; ++i) inv *= 2u - P * inv;
  return 0u - inv;
}
constexpr uint32_t NPINV = neg_inv_p();
constexpr uint32_t ONE = (uint32_t)((1ull << 32) % P);                   // Montgomery form of 1
constexpr uint32_t TWO = (uint32_t)((2ull << 32) % P);                   // Montgomery form of 2
constexpr uint32_t HALF = (uint32_t)(((uint64_t)(P + 1) / 2 << 32) % P);  // Montgomery form of 1/2
constexpr uint32_t R2MOD = (uint32_t)((uint64_t)ONE * ONE % P);  // 2^64 mod p (mulr(x, R2MOD) = x R)
static_assert(P != 0x7f000001u || (NPINV == 0x7effffffu && ONE == 0x01fffffeu && HALF == 0x00ffffffu),
              "KoalaBear constants");

KB_HD uint32_t reduce64(uint64_t x) {
  // Montgomery reduction, x < p * 2^32 -> x * 2^-32 mod p in [0, p).
  uint32_t m = (uint32_t)x * NPINV;
  uint32_t t = (uint32_t)((x + (uint64_t)m * P) >> 32);
  uint32_t u = t - P;
  return t < u ? t : u;  // t in [0, 2p): min(t, t-p) as unsigned
}
// Reduction of a sum of up to four products (x < 4 p^2 < 2p * 2^32): bring the
// high word below p first so that reduce64's bound x < p * 2^32 holds.
KB_HD uint32_t reduce_sum4(uint64_t x) {
  uint32_t hi = (uint32_t)(x >> 32), lo = (uint32_t)x;
  uint32_t h2 = hi - P;
  hi = hi < h2 ? hi : h2;
  return reduce64(((uint64_t)hi << 32) | lo);
}
KB_HD uint32_t mul(uint32_t a, uint32_t b) { return reduce64((uint64_t)a * b); }
KB_HD uint32_t add(uint32_t a, uint32_t b) {
  uint32_t s = a + b, t = s - P;
  return s < t ? s : t;
}
KB_HD uint32_t sub(uint32_t a, uint32_t b) {
  uint32_t d = a - b, t = d + P;
  return d < t ? d : t;
}
KB_HD uint32_t neg(uint32_t a) { return sub(0u, a); }
KB_HD uint32_t dbl(uint32_t a) { return add(a, a); }
KB_HD uint32_t sqr(uint32_t a) { return mul(a, a); }
KB_HD uint32_t pow(uint32_t a, uint64_t e) {
  uint32_t r = ONE;
  while (e) {
    if (e & 1) r = mul(r, a);
    a = mul(a, a);
    e >>= 1;
  }
  return r;
}
KB_HD uint32_t inv(uint32_t a) { return pow(a, P - 2); }
// Montgomery -> canonical integer in [0, p).
KB_HD uint32_t canonical(uint32_t m) { return reduce64((uint64_t)m); }
// canonical integer -> Montgomery
KB_HD uint32_t from_canonical(uint32_t c) { return (uint32_t)(((uint64_t)c << 32) % P); }

struct Ext {
  uint32_t c[4];
};

KB_HD Ext ezero() { return Ext{{0u, 0u, 0u, 0u}}; }
KB_HD Ext eone() { return Ext{{ONE, 0u, 0u, 0u}}; }
KB_HD Ext efrom(uint32_t a) { return Ext{{a, 0u, 0u, 0u}}; }
KB_HD Ext eadd(const Ext& a, const Ext& b) {
  return Ext{{add(a.c[0], b.c[0]), add(a.c[1], b.c[1]), add(a.c[2], b.c[2]), add(a.c[3], b.c[3])}};
}
KB_HD Ext esub(const Ext& a, const Ext& b) {
  return Ext{{sub(a.c[0], b.c[0]), sub(a.c[1], b.c[1]), sub(a.c[2], b.c[2]), sub(a.c[3], b.c[3])}};
}
KB_HD Ext eneg(const Ext& a) { return Ext{{neg(a.c[0]), neg(a.c[1]), neg(a.c[2]), neg(a.c[3])}}; }
KB_HD Ext edbl(const Ext& a) { return eadd(a, a); }
// W t for the extension's W (3: two additions; 11 = 8 + 2 + 1: five).
static_assert(EXT_W == 3 || EXT_W == 11, "extension W");
KB_HD uint32_t mul_w(uint32_t t) {
  if constexpr (EXT_W == 3) {
    return add(dbl(t), t);
  } else {
    const uint32_t t2 = dbl(t);
    return add(add(dbl(dbl(t2)), t2), t);
  }
}
KB_HD Ext emul_base(const Ext& a, uint32_t b) {
  return Ext{{mul(a.c[0], b), mul(a.c[1], b), mul(a.c[2], b), mul(a.c[3], b)}};
}
KB_HD bool eeq(const Ext& a, const Ext& b) {
  return a.c[0] == b.c[0] && a.c[1] == b.c[1] && a.c[2] == b.c[2] && a.c[3] == b.c[3];
}

// Product in F[x]/(x^4 - W). The four output coefficients are accumulated as
// 64-bit sums of 32x32 products (each < p^2 < 2^62) and reduced once each; the
// x^4 wrap-around terms are reduced, multiplied by W and folded in.
KB_HD Ext emul(const Ext& a, const Ext& b) {
  const uint64_t a0 = a.c[0], a1 = a.c[1], a2 = a.c[2], a3 = a.c[3];
  const uint64_t b0 = b.c[0], b1 = b.c[1], b2 = b.c[2], b3 = b.c[3];
  // wrap terms t4 = a1b3 + a2b2 + a3b1, t5 = a2b3 + a3b2, t6 = a3b3
  uint32_t t4 = reduce_sum4(a1 * b3 + a2 * b2 + a3 * b1);
  uint32_t t5 = reduce64(a2 * b3 + a3 * b2);
  uint32_t t6 = reduce64(a3 * b3);
  // W t by additions (cheaper than a multiply by the Montgomery form of W)
  uint32_t w4 = mul_w(t4), w5 = mul_w(t5), w6 = mul_w(t6);
  uint32_t r0 = add(reduce64(a0 * b0), w4);
  uint32_t r1 = add(reduce64(a0 * b1 + a1 * b0), w5);
  uint32_t r2 = add(reduce_sum4(a0 * b2 + a1 * b1 + a2 * b0), w6);
  uint32_t r3 = reduce_sum4(a0 * b3 + a1 * b2 + a2 * b1 + a3 * b0);
  return Ext{{r0, r1, r2, r3}};
}
// Square: the ten distinct products a_i a_j, the cross terms taken once with
// a doubled factor (2 a_i < 2p < 2^32), so every sum stays below 4 p^2.
KB_HD Ext esqr(const Ext& a) {
  const uint64_t a0 = a.c[0], a1 = a.c[1], a2 = a.c[2], a3 = a.c[3];
  const uint64_t d0 = 2 * a0, d1 = 2 * a1, d2 = 2 * a2;
  uint32_t t4 = reduce_sum4(d1 * a3 + a2 * a2);
  uint32_t t5 = reduce_sum4(d2 * a3);
  uint32_t t6 = reduce64(a3 * a3);
  uint32_t w4 = mul_w(t4), w5 = mul_w(t5), w6 = mul_w(t6);
  uint32_t r0 = add(reduce64(a0 * a0), w4);
  uint32_t r1 = add(reduce_sum4(d0 * a1), w5);
  uint32_t r2 = add(reduce_sum4(d0 * a2 + a1 * a1), w6);
  uint32_t r3 = reduce_sum4(d0 * a3 + d1 * a2);
  return Ext{{r0, r1, r2, r3}};
}
KB_HD Ext epow(Ext a, uint64_t e) {
  Ext r = eone();
  while (e) {
    if (e & 1) r = emul(r, a);
    e >>= 1;
    if (e) a = esqr(a);
  }
  return r;
}

// Two-adic generator of order 2^bits: GENERATOR^((p-1)/2^bits) (Plonky3
// KoalaBear / BabyBear two_adic_generator).
KB_HD uint32_t two_adic_generator(unsigned bits) {
  return pow(from_canonical(GENERATOR), (uint64_t)(P - 1) >> bits);
}

}  // namespace kb
