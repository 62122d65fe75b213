// Poseidon2 over KoalaBear, widths 16 and 24, S-box x^3: one permutation per
// lane with the whole state in VGPRs.
//
// Structure (Plonky3 p3-poseidon2 as instantiated by the reference,
// crates/whir/src/test.rs:185-199): MDS-light; 4 external rounds (+RC, x^3,
// MDS-light); R_P internal rounds (+RC on s0, x^3 on s0, 1*1^T + diag(V));
// 4 external rounds. The results are the reference's words, bit for bit; the
// arithmetic inside is organised for gfx950's VALU:
//
// * Lazy reduction. Every linear layer is accumulated exactly in 64 bits with
//   v_mad_u64_u32 (32x32+64 in one full-rate instruction) and reduced once per
//   output by a Montgomery REDC (T -> T * 2^-32 mod p: one mul_lo and one mad),
//   never through chains of modular additions. Outputs stay unreduced where
//   the next consumer's bounds allow (see the bounds on each function).
// * Scaled state. REDC multiplies by R^-1 (R = 2^32), so between layers the
//   state holds x = w * a (mod p) for the reference words w and a known scale
//   a that depends only on the round position. Linear layers with integer
//   coefficients and the S-box commute with this (x^3 -> scale a^3), so the
//   only corrections are: round constants folded into the preceding linear
//   layer's 64-bit sum already multiplied by the scale there; one extra
//   Montgomery multiply of s0 per internal round (by kappa = a^-2 R) to keep all
//   lanes on one scale; and one multiply per output word at the end.
//   Poseidon2Lazy holds these per-round constants, made on the host from the
//   plain round constants by lazy_consts_build().
// Bounds (p = 2^31 - 2^24 + 1, R = 2^32, REDC(T) < T / R + p, valid for
// T < p R): MDS-light sums of inputs < 2p stay below 99p, so their REDC is
// < p + 2^7 ("near-canonical", safe as both factors of a square); an internal
// layer sum + d_i x_i with d_i < p and x_i < 2p stays below 2p^2 + 48p < p R.
#pragma once
#include <cstring>
#include "field.h"

namespace kb {

KB_HD uint64_t mad64(uint32_t a, uint32_t b, uint64_t c) { return (uint64_t)a * b + c; }

// a * K + c in one v_mad_u64_u32 with K an inline constant (1 or 2); written
// as inline ALU asm so that the compiler keeps one instruction per term instead
// of factoring the sums into 64 x 32-bit products.
template <int K>
KB_HD uint64_t mad_k(uint32_t a, uint64_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t d;
  asm("v_mad_u64_u32 %0, vcc, %1, %2, %3" : "=v"(d) : "v"(a), "i"(K), "v"(c) : "vcc");
  return d;
#else
  return (uint64_t)a * K + c;
#endif
}
// a * K (zero-extended to 64 bits) in one instruction.
template <int K>
KB_HD uint64_t mul_k(uint32_t a) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t d;
  asm("v_mad_u64_u32 %0, vcc, %1, %2, 0" : "=v"(d) : "v"(a), "i"(K) : "vcc");
  return d;
#else
  return (uint64_t)a * K;
#endif
}
// c + u for a wave-uniform u (an SGPR operand), one instruction.
KB_HD uint64_t add_u(uint64_t c, uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t d;
  asm("v_mad_u64_u32 %0, vcc, %1, 1, %2" : "=v"(d) : "s"(u), "v"(c) : "vcc");
  return d;
#else
  return c + u;
#endif
}
// a * b + c for a wave-uniform b (an SGPR operand), one instruction.
KB_HD uint64_t mad_u(uint32_t a, uint32_t b, uint64_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t d;
  asm("v_mad_u64_u32 %0, vcc, %1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c) : "vcc");
  return d;
#else
  return (uint64_t)a * b + c;
#endif
}
// T * R^-1 mod p, in [0, T / R + p) for T < p R.
KB_HD uint32_t redc(uint64_t T) {
  const uint32_t m = (uint32_t)T * NPINV;
  return (uint32_t)(mad64(m, P, T) >> 32);
}
KB_HD uint32_t mulr(uint32_t a, uint32_t b) { return redc((uint64_t)a * b); }  // needs a * b < p R
KB_HD uint32_t canon2(uint32_t t) {  // [0, 2p) -> [0, p)
  const uint32_t u = t - P;
  return t < u ? t : u;
}

// The same reference, passed through an empty asm statement: loads through it
// cannot be hoisted out of a loop around the caller. A persistent kernel that
// runs one permutation per loop iteration uses it for the round constants,
// which LLVM would otherwise load once before the loop and keep in more
// scalar registers than there are (spilling them).
// The pointer stays in the constant address space (the reference is to
// __constant__ memory), so the loads remain scalar loads.
template <class T>
__device__ __forceinline__ const T& opaque_ref(const T& r) {
  typedef const __attribute__((address_space(4))) T CT;
  CT* q = (CT*)&r;
  asm volatile("" : "+s"(q));
  return *(const T*)q;
}

}  // namespace kb

#if defined(WHIR_FIELD_BABYBEAR)
#include "poseidon2_bb.h"
#else
namespace kb {

// The plain round constants (Montgomery words, the reference's draw order);
// host side only: the kernels use the derived Poseidon2Lazy tables.
struct Poseidon2Consts {
  uint32_t ext16_init[4][16];
  uint32_t ext16_term[4][16];
  uint32_t int16[20];
  uint32_t ext24_init[4][24];
  uint32_t ext24_term[4][24];
  uint32_t int24[23];
};

// Per-round constants of the lazily reduced permutation of width W (values in
// [0, p)).
//   lin_add[0]     : ext round 0 RC, folded into the first MDS-light
//   lin_add[1..3]  : ext rounds 1..3 RC, folded into ext rounds 0..2's MDS
//   lin_add[4][0]  : internal round 0 RC (lane 0 only), into ext round 3's MDS
//   int_add[r]     : internal round r RC (r >= 1), into round r-1's layer
//   lin_add[5]     : terminal round 0 RC, into the last internal layer
//   lin_add[6..8]  : terminal rounds 1..3 RC, into terminal rounds 0..2's MDS
//   kappa[r]       : s0 scale correction of internal round r
//   fix            : a_final^-1 * R, turns the final state back into words
// diag(i) is V_i as an integer in [0, p) (Plonky3 KoalaBear INTERNAL_V16 /
// INTERNAL_V24: sign * num / 2^sh).
struct DiagTerm {
  int sign;
  uint32_t num;
  int sh;
};
constexpr DiagTerm kV16[16] = {{-1, 2, 0}, {1, 1, 0}, {1, 2, 0}, {1, 1, 1}, {1, 3, 0},  {1, 4, 0},
                               {-1, 1, 1}, {-1, 3, 0}, {-1, 4, 0}, {1, 1, 8}, {1, 1, 3}, {1, 1, 24},
                               {-1, 1, 8}, {-1, 1, 3}, {-1, 1, 4}, {-1, 1, 24}};
constexpr DiagTerm kV24[24] = {{-1, 2, 0}, {1, 1, 0},  {1, 2, 0},  {1, 1, 1},  {1, 3, 0},  {1, 4, 0},
                               {-1, 1, 1}, {-1, 3, 0}, {-1, 4, 0}, {1, 1, 8},  {1, 1, 2},  {1, 1, 3},
                               {1, 1, 4},  {1, 1, 5},  {1, 1, 6},  {1, 1, 24}, {-1, 1, 8}, {-1, 1, 3},
                               {-1, 1, 4}, {-1, 1, 5}, {-1, 1, 6}, {-1, 1, 7}, {-1, 1, 9}, {-1, 1, 24}};
// num * 2^-sh mod p with sign; 2^-k = p - (p - 1) / 2^k since p = 1 (mod 2^24).
constexpr uint32_t diag_value(DiagTerm d) {
  const uint64_t inv = d.sh == 0 ? 1 : (uint64_t)P - ((uint64_t)(P - 1) >> d.sh);
  const uint32_t v = (uint32_t)((uint64_t)d.num * inv % P);
  return d.sign < 0 ? (v ? P - v : 0u) : v;
}

template <int W>
struct Poseidon2Lazy {
  static constexpr int RP = W == 16 ? 20 : 23;
  uint32_t lin_add[9][W];
  uint32_t int_add[RP];
  uint32_t kappa[RP];
  uint32_t fix;
  KB_HD static constexpr uint32_t diag(int i) { return diag_value(W == 16 ? kV16[i] : kV24[i]); }
};

struct Poseidon2LazyConsts {
  Poseidon2Lazy<16> w16;
  Poseidon2Lazy<24> w24;
};

__constant__ Poseidon2LazyConsts c_p2l;

// Per-lane coefficient columns of the one-wavefront P16 (permute16_wave in
// tx.h), whose 20 internal rounds run as one chain of S-boxes over a linear
// recurrence. With the state u entering the internal rounds, x_r the S-box
// input of round r and y_r = x_r^3 its output, and A = M_I with column 0
// zeroed, b = column 0 of M_I (M_I = 1 1^T + diag(V)):
//   u_{r+1} = A u_r + b y_r,  x_r = e0^T A^r u + rc_r + sum_{j<r} e0^T A^{r-1-j} b y_j,
//   u_20 = A^20 u + sum_j A^{19-j} b y_j.
// Lane 16 + r (r = 1..19) accumulates R * (scaled x_r); lane i < 16 accumulates
// R * (scaled u_20[i] + the first terminal round constant). Column layout,
// lane-major: col[l][j] (j < 20) multiplies y_j, col[l][20 + k - 1] (k = 1..15)
// multiplies input word k, col[l][35] is the accumulator's starting value.
struct Poseidon2Wave16 {
  uint32_t col[64][36];
};
__constant__ Poseidon2Wave16 c_p2w;

template <int W>
__device__ __forceinline__ const Poseidon2Lazy<W>& lazy_consts() {
  if constexpr (W == 16) return c_p2l.w16;
  else return c_p2l.w24;
}

// internal_lazy's sum without x[0] as two interleaved chains, sa = x1 + x3 +
// ... + x[W-1] and sb = x2 + x4 + ..., in one asm block on the device: the
// same v_mad_u64_u32 sequence as separate mad_k<1> calls, without the wait
// state the compiler puts after every inline asm statement (consecutive
// dependent v_mad_u64_u32 need none: compiled code issues them back to back).
template <int W>
KB_HD void sum_chains(const uint32_t (&x)[W], uint64_t& sa, uint64_t& sb) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (W == 16) {
    asm("v_mad_u64_u32 %0, vcc, %2, 1, 0\n"
        "v_mad_u64_u32 %1, vcc, %3, 1, 0\n"
        "v_mad_u64_u32 %0, vcc, %4, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %5, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %6, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %7, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %8, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %9, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %10, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %11, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %12, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %13, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %14, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %15, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %16, 1, %0"
        : "=&v"(sa), "=&v"(sb)
        : "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7]), "v"(x[8]), "v"(x[9]), "v"(x[10]), "v"(x[11]), "v"(x[12]), "v"(x[13]), "v"(x[14]), "v"(x[15])
        : "vcc");
    return;
  }
  else if constexpr (W == 24) {
    asm("v_mad_u64_u32 %0, vcc, %2, 1, 0\n"
        "v_mad_u64_u32 %1, vcc, %3, 1, 0\n"
        "v_mad_u64_u32 %0, vcc, %4, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %5, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %6, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %7, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %8, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %9, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %10, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %11, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %12, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %13, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %14, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %15, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %16, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %17, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %18, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %19, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %20, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %21, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %22, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %23, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %24, 1, %0"
        : "=&v"(sa), "=&v"(sb)
        : "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7]), "v"(x[8]), "v"(x[9]), "v"(x[10]), "v"(x[11]), "v"(x[12]), "v"(x[13]), "v"(x[14]), "v"(x[15]), "v"(x[16]), "v"(x[17]), "v"(x[18]), "v"(x[19]), "v"(x[20]), "v"(x[21]), "v"(x[22]), "v"(x[23])
        : "vcc");
    return;
  }
#endif
  sa = x[1];
  sb = x[2];
  for (int j = 3; j + 1 < W; j += 2) {
    sa += x[j];
    sb += x[j + 1];
  }
  sa += x[W - 1];
}

// One 4-chunk of MDS-light, y_a = S + x_a + 2 x_{a+1} (S the chunk sum), as
// one asm block of twelve v_mad_u64_u32 (the mad_k chain of mds_lazy).
KB_HD void m4_chunk(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3, uint64_t (&y)[4]) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t s;
  asm("v_mad_u64_u32 %4, vcc, %5, 1, 0\n"
      "v_mad_u64_u32 %4, vcc, %6, 1, %4\n"
      "v_mad_u64_u32 %4, vcc, %7, 1, %4\n"
      "v_mad_u64_u32 %4, vcc, %8, 1, %4\n"
      "v_mad_u64_u32 %0, vcc, %5, 1, %4\n"
      "v_mad_u64_u32 %1, vcc, %6, 1, %4\n"
      "v_mad_u64_u32 %2, vcc, %7, 1, %4\n"
      "v_mad_u64_u32 %3, vcc, %8, 1, %4\n"
      "v_mad_u64_u32 %0, vcc, %6, 2, %0\n"
      "v_mad_u64_u32 %1, vcc, %7, 2, %1\n"
      "v_mad_u64_u32 %2, vcc, %8, 2, %2\n"
      "v_mad_u64_u32 %3, vcc, %5, 2, %3"
      : "=&v"(y[0]), "=&v"(y[1]), "=&v"(y[2]), "=&v"(y[3]), "=&v"(s)
      : "v"(x0), "v"(x1), "v"(x2), "v"(x3)
      : "vcc");
#else
  const uint64_t s = (uint64_t)x0 + x1 + x2 + x3;
  y[0] = s + x0 + 2ull * x1;
  y[1] = s + x1 + 2ull * x2;
  y[2] = s + x2 + 2ull * x3;
  y[3] = s + x3 + 2ull * x0;
#endif
}

// t[k] = x[I0 + k] * V_{I0 + k} + sum for k < K (K <= 8), the internal
// layer's diagonal terms, as one asm block on the device (see sum_chains).
template <int W, int I0, int K>
KB_HD void diag_mads(const uint32_t (&x)[W], uint64_t sum, uint64_t (&t)[K]) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (K == 8) {
    asm("v_mad_u64_u32 %0, vcc, %8, %16, %24\n"
        "v_mad_u64_u32 %1, vcc, %9, %17, %24\n"
        "v_mad_u64_u32 %2, vcc, %10, %18, %24\n"
        "v_mad_u64_u32 %3, vcc, %11, %19, %24\n"
        "v_mad_u64_u32 %4, vcc, %12, %20, %24\n"
        "v_mad_u64_u32 %5, vcc, %13, %21, %24\n"
        "v_mad_u64_u32 %6, vcc, %14, %22, %24\n"
        "v_mad_u64_u32 %7, vcc, %15, %23, %24"
        : "=&v"(t[0]), "=&v"(t[1]), "=&v"(t[2]), "=&v"(t[3]), "=&v"(t[4]), "=&v"(t[5]), "=&v"(t[6]), "=&v"(t[7])
        : "v"(x[I0]), "v"(x[I0 + 1]), "v"(x[I0 + 2]), "v"(x[I0 + 3]), "v"(x[I0 + 4]), "v"(x[I0 + 5]),
          "v"(x[I0 + 6]), "v"(x[I0 + 7]), "s"(Poseidon2Lazy<W>::diag(I0)), "s"(Poseidon2Lazy<W>::diag(I0 + 1)),
          "s"(Poseidon2Lazy<W>::diag(I0 + 2)), "s"(Poseidon2Lazy<W>::diag(I0 + 3)),
          "s"(Poseidon2Lazy<W>::diag(I0 + 4)), "s"(Poseidon2Lazy<W>::diag(I0 + 5)),
          "s"(Poseidon2Lazy<W>::diag(I0 + 6)), "s"(Poseidon2Lazy<W>::diag(I0 + 7)), "v"(sum)
        : "vcc");
    return;
  } else if constexpr (K == 7) {
    asm("v_mad_u64_u32 %0, vcc, %7, %14, %21\n"
        "v_mad_u64_u32 %1, vcc, %8, %15, %21\n"
        "v_mad_u64_u32 %2, vcc, %9, %16, %21\n"
        "v_mad_u64_u32 %3, vcc, %10, %17, %21\n"
        "v_mad_u64_u32 %4, vcc, %11, %18, %21\n"
        "v_mad_u64_u32 %5, vcc, %12, %19, %21\n"
        "v_mad_u64_u32 %6, vcc, %13, %20, %21"
        : "=&v"(t[0]), "=&v"(t[1]), "=&v"(t[2]), "=&v"(t[3]), "=&v"(t[4]), "=&v"(t[5]), "=&v"(t[6])
        : "v"(x[I0]), "v"(x[I0 + 1]), "v"(x[I0 + 2]), "v"(x[I0 + 3]), "v"(x[I0 + 4]), "v"(x[I0 + 5]),
          "v"(x[I0 + 6]), "s"(Poseidon2Lazy<W>::diag(I0)), "s"(Poseidon2Lazy<W>::diag(I0 + 1)),
          "s"(Poseidon2Lazy<W>::diag(I0 + 2)), "s"(Poseidon2Lazy<W>::diag(I0 + 3)),
          "s"(Poseidon2Lazy<W>::diag(I0 + 4)), "s"(Poseidon2Lazy<W>::diag(I0 + 5)),
          "s"(Poseidon2Lazy<W>::diag(I0 + 6)), "v"(sum)
        : "vcc");
    return;
  }
#endif
  for (int k = 0; k < K; ++k) t[k] = (uint64_t)x[I0 + k] * Poseidon2Lazy<W>::diag(I0 + k) + sum;
}

enum { LAZY_ADD_NONE = 0, LAZY_ADD_FIRST = 1, LAZY_ADD_ALL = 2 };

// x <- REDC(MDS-light(x) + add), x_i < 2p on entry, < p + 2^7 on exit.
// MDS-light = (J + I) (x) M4 (J all-ones over the 4-chunks); per chunk
// M4 x = S + (x_a + 2 x_{a+1}) with S the chunk sum (Plonky3's M4 rows).
template <int W, int ADD>
KB_HD void mds_lazy(uint32_t (&x)[W], const uint32_t* add) {
  uint64_t y[W];
#pragma unroll
  for (int c = 0; c < W; c += 4) {
    uint64_t yc[4];
    m4_chunk(x[c], x[c + 1], x[c + 2], x[c + 3], yc);
#pragma unroll
    for (int a = 0; a < 4; ++a) y[c + a] = yc[a];
  }
  uint64_t sums[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    sums[a] = y[a];
#pragma unroll
    for (int c = 4; c < W; c += 4) sums[a] += y[c + a];
  }
#pragma unroll
  for (int i = 0; i < W; ++i) {
    uint64_t t = y[i] + sums[i & 3];
    if constexpr (ADD == LAZY_ADD_ALL) t = add_u(t, add[i]);
    if constexpr (ADD == LAZY_ADD_FIRST) {
      if (i == 0) t = add_u(t, add[0]);
    }
    x[i] = redc(t);
  }
}

// External round: x^3 on every lane (x near-canonical, result < 2p), then MDS.
template <int W, int ADD>
KB_HD void external_lazy(uint32_t (&x)[W], const uint32_t* add) {
#pragma unroll
  for (int i = 0; i < W; ++i) x[i] = mulr(mulr(x[i], x[i]), x[i]);
  mds_lazy<W, ADD>(x, add);
}

// Internal round r: s0 <- kappa_r * s0^3 (x0 near-canonical -> < 1.87p), then
// x_i <- REDC(sum + V_i x_i) with the next round constant folded in; lane 0
// uses V_0 = -2 as 2 (2p - x0) so that its result is near-canonical for the
// next S-box. The last round canonicalises every lane (its outputs feed the
// terminal S-boxes) and folds terminal round 0's constants.
template <int W, bool LAST>
KB_HD void internal_lazy(uint32_t (&x)[W], int r, const Poseidon2Lazy<W>& C) {
  x[0] = mulr(mulr(mulr(x[0], x[0]), x[0]), C.kappa[r]);
  // two interleaved accumulation chains over x[1..W)
  uint64_t sa, sb;
  sum_chains<W>(x, sa, sb);
  const uint64_t sum = mad_k<1>(x[0], sa + sb);
  uint64_t t0 = add_u(mad_k<2>(2u * P - x[0], sum), LAST ? C.lin_add[5][0] : C.int_add[r + 1]);
  x[0] = LAST ? canon2(redc(t0)) : redc(t0);
  // words 1..W-1 in blocks of 8, 8 (W = 24) and 7
  uint64_t t[W];
  {
    uint64_t b[8];
    diag_mads<W, 1, 8>(x, sum, b);
#pragma unroll
    for (int k = 0; k < 8; ++k) t[1 + k] = b[k];
  }
  if constexpr (W == 24) {
    uint64_t b[8];
    diag_mads<W, 9, 8>(x, sum, b);
#pragma unroll
    for (int k = 0; k < 8; ++k) t[9 + k] = b[k];
  }
  {
    uint64_t b[7];
    diag_mads<W, W - 7, 7>(x, sum, b);
#pragma unroll
    for (int k = 0; k < 7; ++k) t[W - 7 + k] = b[k];
  }
#pragma unroll
  for (int i = 1; i < W; ++i) {
    if constexpr (LAST) {
      x[i] = canon2(redc(add_u(t[i], C.lin_add[5][i])));
    } else {
      x[i] = redc(t[i]);
    }
  }
}

// Poseidon2 permutation of width W. In: s = words in [0, p). Out: s[i] for the
// lanes i in OUT (a bit mask) are the reference's output words in [0, p); the
// other lanes hold scaled values and must not be used.
// Internal rounds R..RP-2 by template recursion, so that every round is
// straight-line code with compile-time constant offsets (a counted loop here
// makes the compiler re-load the round constants through indexed scalar loads).
template <int W, int R>
KB_HD void internal_rounds_lazy(uint32_t (&s)[W], const Poseidon2Lazy<W>& C) {
  if constexpr (R < Poseidon2Lazy<W>::RP - 1) {
    internal_lazy<W, false>(s, R, C);
    internal_rounds_lazy<W, R + 1>(s, C);
  }
}

template <int W, uint32_t OUT = (1u << W) - 1u>
KB_HD void permute_lazy(uint32_t (&s)[W], const Poseidon2Lazy<W>& C) {
  constexpr int RP = Poseidon2Lazy<W>::RP;
  mds_lazy<W, LAZY_ADD_ALL>(s, C.lin_add[0]);
#pragma unroll
  for (int r = 0; r < 3; ++r) external_lazy<W, LAZY_ADD_ALL>(s, C.lin_add[r + 1]);
  external_lazy<W, LAZY_ADD_FIRST>(s, C.lin_add[4]);
  internal_rounds_lazy<W, 0>(s, C);
  internal_lazy<W, true>(s, RP - 1, C);
#pragma unroll
  for (int r = 0; r < 3; ++r) external_lazy<W, LAZY_ADD_ALL>(s, C.lin_add[6 + r]);
  external_lazy<W, LAZY_ADD_NONE>(s, nullptr);
#pragma unroll
  for (int i = 0; i < W; ++i)
    if ((OUT >> i) & 1u) s[i] = canon2(mulr(s[i], C.fix));
}

template <int W, uint32_t OUT = (1u << W) - 1u, bool OPAQUE = false>
__device__ __forceinline__ void permute(uint32_t (&s)[W]) {
  if constexpr (OPAQUE) permute_lazy<W, OUT>(s, opaque_ref(lazy_consts<W>()));
  else permute_lazy<W, OUT>(s, lazy_consts<W>());
}

// Host: the constants of permute_lazy from the plain round constants
// (Montgomery words, in the reference's order) by replaying the scale of every
// layer: a = 1 at the input, a -> a / R at each REDC, a -> a^3 at each S-box.
namespace lazy_host {
inline uint32_t mm(uint64_t a, uint64_t b) { return (uint32_t)(a * b % P); }
inline uint32_t mpow(uint32_t a, uint64_t e) {
  uint32_t r = 1;
  while (e) {
    if (e & 1) r = mm(r, a);
    a = mm(a, a);
    e >>= 1;
  }
  return r;
}
inline uint32_t minv(uint32_t a) { return mpow(a, P - 2); }
}  // namespace lazy_host

template <int W>
inline void lazy_consts_build(const uint32_t (&ext_init)[4][W], const uint32_t (&ext_term)[4][W],
                              const uint32_t* internal, Poseidon2Lazy<W>& C) {
  using namespace lazy_host;
  constexpr int RP = Poseidon2Lazy<W>::RP;
  const uint32_t R = (uint32_t)((1ull << 32) % P), Rinv = minv(R);
  memset(&C, 0, sizeof(C));
  uint32_t a = 1;
  auto cube_scale = [&]() { a = mm(mm(a, a), a); };
  for (int i = 0; i < W; ++i) C.lin_add[0][i] = mm(ext_init[0][i], a);
  a = mm(a, Rinv);
  for (int r = 0; r < 4; ++r) {
    cube_scale();
    if (r < 3)
      for (int i = 0; i < W; ++i) C.lin_add[r + 1][i] = mm(ext_init[r + 1][i], a);
    else
      C.lin_add[4][0] = mm(internal[0], a);
    a = mm(a, Rinv);
  }
  for (int r = 0; r < RP; ++r) {
    C.kappa[r] = mm(minv(mm(a, a)), R);
    if (r + 1 < RP)
      C.int_add[r + 1] = mm(internal[r + 1], a);
    else
      for (int i = 0; i < W; ++i) C.lin_add[5][i] = mm(ext_term[0][i], a);
    a = mm(a, Rinv);
  }
  for (int r = 0; r < 4; ++r) {
    cube_scale();
    if (r < 3)
      for (int i = 0; i < W; ++i) C.lin_add[6 + r][i] = mm(ext_term[r + 1][i], a);
    a = mm(a, Rinv);
  }
  C.fix = mm(minv(a), R);
}

// Host: the columns of Poseidon2Wave16 for the lazy scales of
// lazy_consts_build (sigma: the scale of the state entering the internal
// rounds; a_t: the scale leaving them). Words are integers mod p: the
// internal layer is the integer matrix M_I, the S-box w -> w^3 R^-2.
inline void wave16_consts_build(const uint32_t (&ext_term)[4][16], const uint32_t* internal,
                                Poseidon2Wave16& Wc) {
  using namespace lazy_host;
  constexpr int N = 16, RP = 20;
  const uint32_t R = (uint32_t)((1ull << 32) % P), Rinv = minv(R);
  uint32_t a = mm(1, Rinv);
  for (int r = 0; r < 4; ++r) a = mm(mm(mm(a, a), a), Rinv);
  const uint32_t sigma = a;
  for (int r = 0; r < RP; ++r) a = mm(a, Rinv);
  const uint32_t a_t = a;
  const uint32_t si1 = minv(sigma), si2 = mm(si1, si1), si3 = mm(si2, si1);
  typedef uint32_t Mat[N][N];
  static Mat pw[RP + 1];  // A^n
  Mat A;
  uint32_t b[N];
  for (int i = 0; i < N; ++i) {
    for (int k = 0; k < N; ++k) A[i][k] = k == 0 ? 0u : (uint32_t)((1ull + (i == k ? Poseidon2Lazy<16>::diag(i) : 0u)) % P);
    b[i] = (uint32_t)((1ull + (i == 0 ? Poseidon2Lazy<16>::diag(0) : 0u)) % P);
  }
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < N; ++k) pw[0][i][k] = i == k ? 1u : 0u;
  for (int n = 1; n <= RP; ++n)
    for (int i = 0; i < N; ++i)
      for (int k = 0; k < N; ++k) {
        uint64_t s = 0;
        for (int j = 0; j < N; ++j) s = (s + (uint64_t)A[i][j] * pw[n - 1][j][k]) % P;
        pw[n][i][k] = (uint32_t)s;
      }
  auto apply_b = [&](int n, int i) {  // (A^n b)[i]
    uint64_t s = 0;
    for (int j = 0; j < N; ++j) s = (s + (uint64_t)pw[n][i][j] * b[j]) % P;
    return (uint32_t)s;
  };
  memset(&Wc, 0, sizeof(Wc));
  for (int i = 0; i < N; ++i) {
    const uint32_t ra = mm(R, a_t);
    for (int j = 0; j < RP; ++j) Wc.col[i][j] = mm(mm(ra, apply_b(RP - 1 - j, i)), si3);
    for (int k = 1; k < N; ++k) Wc.col[i][RP + k - 1] = mm(mm(ra, pw[RP][i][k]), si1);
    Wc.col[i][35] = mm(ra, ext_term[0][i]);
  }
  for (int r = 1; r < RP; ++r) {
    uint32_t* c = Wc.col[16 + r];
    for (int j = 0; j < r; ++j) c[j] = mm(mm(R, apply_b(r - 1 - j, 0)), si2);
    for (int k = 1; k < N; ++k) c[RP + k - 1] = mm(R, pw[r][0][k]);
    c[35] = mm(mm(R, sigma), internal[r]);
  }
}

// TruncatedPermutation<P16, 2, 8, 16>: first 8 words of P16(l || r).
template <bool OPAQUE = false>
__device__ __forceinline__ void compress(const uint32_t (&l)[8], const uint32_t (&r)[8], uint32_t (&out)[8]) {
  uint32_t s[16];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    s[i] = l[i];
    s[8 + i] = r[i];
  }
  permute<16, 0xFFu, OPAQUE>(s);
#pragma unroll
  for (int i = 0; i < 8; ++i) out[i] = s[i];
}

}  // namespace kb
#endif
