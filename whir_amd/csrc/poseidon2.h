// Poseidon2 over KoalaBear, widths 16 and 24, S-box x^3: one permutation per
// lane with the whole state in VGPRs. Round constants sit in __constant__
// memory (uniform -> scalar loads); the internal diagonal V is applied with
// shifts/adds (multiplication by +-2^-n uses p = 127 * 2^24 + 1) instead of
// Montgomery multiplies.
//
// Structure (Plonky3 p3-poseidon2 as instantiated by the reference,
// crates/whir/src/test.rs:185-199): MDS-light; 4 external rounds (+RC, x^3,
// MDS-light); R_P internal rounds (+RC on s0, x^3 on s0, 1*1^T + diag(V));
// 4 external rounds.
#pragma once
#include "field.h"

namespace kb {

struct Poseidon2Consts {
  uint32_t ext16_init[4][16];
  uint32_t ext16_term[4][16];
  uint32_t int16[20];
  uint32_t ext24_init[4][24];
  uint32_t ext24_term[4][24];
  uint32_t int24[23];
};

__constant__ Poseidon2Consts c_p2;

__device__ __forceinline__ uint32_t cube(uint32_t x) { return mul(sqr(x), x); }

// x * 2^-n for 1 <= n <= 24: since p = 1 (mod 2^24), k = (-x) mod 2^n makes
// x + k*p divisible by 2^n, and (x + k*p) / 2^n = ((x + k) >> n) + 127*k*2^(24-n).
template <int N>
__device__ __forceinline__ uint32_t mul_2exp_neg(uint32_t x) {
  const uint32_t mask = (1u << N) - 1u;
  uint32_t k = (0u - x) & mask;
  uint32_t r = ((x + k) >> N) + ((k * 127u) << (24 - N));
  uint32_t t = r - P;
  return r < t ? r : t;
}

// Plonky3 MDSMat4: [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]].
__device__ __forceinline__ void mat4(uint32_t& x0, uint32_t& x1, uint32_t& x2, uint32_t& x3) {
  uint32_t t01 = add(x0, x1);
  uint32_t t23 = add(x2, x3);
  uint32_t t0123 = add(t01, t23);
  uint32_t t01123 = add(t0123, x1);
  uint32_t t01233 = add(t0123, x3);
  uint32_t y3 = add(t01233, dbl(x0));
  uint32_t y1 = add(t01123, dbl(x2));
  uint32_t y0 = add(t01123, t01);
  uint32_t y2 = add(t01233, t23);
  x0 = y0;
  x1 = y1;
  x2 = y2;
  x3 = y3;
}

template <int W>
__device__ __forceinline__ void mds_light(uint32_t (&s)[W]) {
#pragma unroll
  for (int c = 0; c < W; c += 4) mat4(s[c], s[c + 1], s[c + 2], s[c + 3]);
  uint32_t sums[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    sums[i] = s[i];
#pragma unroll
    for (int c = 4; c < W; c += 4) sums[i] = add(sums[i], s[c + i]);
  }
#pragma unroll
  for (int i = 0; i < W; ++i) s[i] = add(s[i], sums[i & 3]);
}

// V_i * x + sum for the KoalaBear internal diagonals
//   V16 = [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, 2^-8, 2^-3, 2^-24, -2^-8, -2^-3, -2^-4, -2^-24]
//   V24 = [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, 2^-8, 2^-2, 2^-3, 2^-4, 2^-5, 2^-6, 2^-24,
//          -2^-8, -2^-3, -2^-4, -2^-5, -2^-6, -2^-7, -2^-9, -2^-24]
template <int W, int I>
__device__ __forceinline__ uint32_t diag_apply(uint32_t x, uint32_t sum) {
  // common prefix for both widths
  if constexpr (I == 0) return sub(sub(sum, x), x);
  else if constexpr (I == 1) return add(sum, x);
  else if constexpr (I == 2) return add(sum, dbl(x));
  else if constexpr (I == 3) return add(sum, mul_2exp_neg<1>(x));
  else if constexpr (I == 4) return add(sum, add(dbl(x), x));
  else if constexpr (I == 5) return add(sum, dbl(dbl(x)));
  else if constexpr (I == 6) return sub(sum, mul_2exp_neg<1>(x));
  else if constexpr (I == 7) return sub(sum, add(dbl(x), x));
  else if constexpr (I == 8) return sub(sum, dbl(dbl(x)));
  else if constexpr (W == 16) {
    if constexpr (I == 9) return add(sum, mul_2exp_neg<8>(x));
    else if constexpr (I == 10) return add(sum, mul_2exp_neg<3>(x));
    else if constexpr (I == 11) return add(sum, mul_2exp_neg<24>(x));
    else if constexpr (I == 12) return sub(sum, mul_2exp_neg<8>(x));
    else if constexpr (I == 13) return sub(sum, mul_2exp_neg<3>(x));
    else if constexpr (I == 14) return sub(sum, mul_2exp_neg<4>(x));
    else return sub(sum, mul_2exp_neg<24>(x));
  } else {
    if constexpr (I == 9) return add(sum, mul_2exp_neg<8>(x));
    else if constexpr (I == 10) return add(sum, mul_2exp_neg<2>(x));
    else if constexpr (I == 11) return add(sum, mul_2exp_neg<3>(x));
    else if constexpr (I == 12) return add(sum, mul_2exp_neg<4>(x));
    else if constexpr (I == 13) return add(sum, mul_2exp_neg<5>(x));
    else if constexpr (I == 14) return add(sum, mul_2exp_neg<6>(x));
    else if constexpr (I == 15) return add(sum, mul_2exp_neg<24>(x));
    else if constexpr (I == 16) return sub(sum, mul_2exp_neg<8>(x));
    else if constexpr (I == 17) return sub(sum, mul_2exp_neg<3>(x));
    else if constexpr (I == 18) return sub(sum, mul_2exp_neg<4>(x));
    else if constexpr (I == 19) return sub(sum, mul_2exp_neg<5>(x));
    else if constexpr (I == 20) return sub(sum, mul_2exp_neg<6>(x));
    else if constexpr (I == 21) return sub(sum, mul_2exp_neg<7>(x));
    else if constexpr (I == 22) return sub(sum, mul_2exp_neg<9>(x));
    else return sub(sum, mul_2exp_neg<24>(x));
  }
}

template <int W, int I = 0>
__device__ __forceinline__ void diag_all(uint32_t (&s)[W], uint32_t sum) {
  if constexpr (I < W) {
    s[I] = diag_apply<W, I>(s[I], sum);
    diag_all<W, I + 1>(s, sum);
  }
}

template <int W>
__device__ __forceinline__ void internal_layer(uint32_t (&s)[W]) {
  uint32_t sum = s[0];
#pragma unroll
  for (int i = 1; i < W; ++i) sum = add(sum, s[i]);
  diag_all<W>(s, sum);
}

template <int W>
__device__ __forceinline__ uint32_t rc_init(int r, int i) {
  if constexpr (W == 16) return c_p2.ext16_init[r][i];
  else return c_p2.ext24_init[r][i];
}
template <int W>
__device__ __forceinline__ uint32_t rc_term(int r, int i) {
  if constexpr (W == 16) return c_p2.ext16_term[r][i];
  else return c_p2.ext24_term[r][i];
}
template <int W>
__device__ __forceinline__ uint32_t rc_int(int r) {
  if constexpr (W == 16) return c_p2.int16[r];
  else return c_p2.int24[r];
}

template <int W>
__device__ __forceinline__ void permute(uint32_t (&s)[W]) {
  constexpr int RP = W == 16 ? 20 : 23;
  mds_light<W>(s);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int i = 0; i < W; ++i) s[i] = cube(add(s[i], rc_init<W>(r, i)));
    mds_light<W>(s);
  }
#pragma unroll
  for (int r = 0; r < RP; ++r) {
    s[0] = cube(add(s[0], rc_int<W>(r)));
    internal_layer<W>(s);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int i = 0; i < W; ++i) s[i] = cube(add(s[i], rc_term<W>(r, i)));
    mds_light<W>(s);
  }
}

// TruncatedPermutation<P16, 2, 8, 16>: first 8 words of P16(l || r).
__device__ __forceinline__ void compress(const uint32_t (&l)[8], const uint32_t (&r)[8], uint32_t (&out)[8]) {
  uint32_t s[16];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    s[i] = l[i];
    s[8 + i] = r[i];
  }
  permute<16>(s);
#pragma unroll
  for (int i = 0; i < 8; ++i) out[i] = s[i];
}

}  // namespace kb
