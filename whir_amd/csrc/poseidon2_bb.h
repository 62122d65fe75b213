// Poseidon2 over BabyBear, widths 16 and 24 (Plonky3 Poseidon2BabyBear<W>
// with constants new_from_rng_128(StdRng::seed_from_u64(seed)), the
// permutation the reference's transcript test instantiates,
// crates/transcript/src/lib.rs:389-396): S-box x^7 (7 is the smallest d with
// gcd(d, p - 1) = 1), R_F = 8, R_P = 13 (W16) / 21 (W24) (the 128-bit round
// numbers), MDS-light external layer, internal layer 1 1^T + diag(V) with
// Plonky3's BabyBear V. Plain Montgomery arithmetic on canonical words: one
// permutation per lane, or per 16-lane row (cooperative, tx.h).
#pragma once

namespace kb {

constexpr int BB_SBOX_DEGREE = 7;

// V_i = sign * num / 2^sh (Plonky3 baby-bear INTERNAL_DIAG_MONTY_16 / _24).
struct DiagTerm {
  int sign;
  uint32_t num;
  int sh;
};
constexpr DiagTerm kV16[16] = {{-1, 2, 0}, {1, 1, 0}, {1, 2, 0}, {1, 1, 1}, {1, 3, 0},  {1, 4, 0},
                               {-1, 1, 1}, {-1, 3, 0}, {-1, 4, 0}, {1, 1, 8}, {1, 1, 2}, {1, 1, 3},
                               {1, 1, 27}, {-1, 1, 8}, {-1, 1, 4}, {-1, 1, 27}};
constexpr DiagTerm kV24[24] = {{-1, 2, 0}, {1, 1, 0},  {1, 2, 0},  {1, 1, 1},  {1, 3, 0},  {1, 4, 0},
                               {-1, 1, 1}, {-1, 3, 0}, {-1, 4, 0}, {1, 1, 8},  {1, 1, 2},  {1, 1, 3},
                               {1, 1, 4},  {1, 1, 7},  {1, 1, 9},  {1, 1, 27}, {-1, 1, 8}, {-1, 1, 2},
                               {-1, 1, 3}, {-1, 1, 4}, {-1, 1, 5}, {-1, 1, 6}, {-1, 1, 7}, {-1, 1, 27}};
// num * 2^-sh mod p with sign, canonical; 2^-k = p - (p - 1) / 2^k since p = 1 (mod 2^27).
constexpr uint32_t diag_value(DiagTerm d) {
  const uint64_t inv = d.sh == 0 ? 1 : (uint64_t)P - ((uint64_t)(P - 1) >> d.sh);
  const uint32_t v = (uint32_t)((uint64_t)d.num * inv % P);
  return d.sign < 0 ? (v ? P - v : 0u) : v;
}

// Round constants and the diagonal, Montgomery words.
template <int W>
struct Poseidon2Plain {
  static constexpr int RP = W == 16 ? 13 : 21;
  uint32_t ext_init[4][W];
  uint32_t ext_term[4][W];
  uint32_t internal[RP];
  uint32_t diag[W];
};
struct Poseidon2PlainConsts {
  Poseidon2Plain<16> w16;
  Poseidon2Plain<24> w24;
};

__constant__ Poseidon2PlainConsts c_p2b;

template <int W>
__device__ __forceinline__ const Poseidon2Plain<W>& plain_consts() {
  if constexpr (W == 16) return c_p2b.w16;
  else return c_p2b.w24;
}

KB_HD uint32_t sbox7(uint32_t x) {
  const uint32_t x2 = mul(x, x), x3 = mul(x2, x), x4 = mul(x2, x2);
  return mul(x3, x4);
}

// MDS-light: M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]] on every 4-chunk
// (row a: chunk sum + x_a + 2 x_(a+1)), then every word plus the sum of the
// chunks' words at its position.
template <int W>
KB_HD void mds_light_plain(uint32_t (&s)[W]) {
#pragma unroll
  for (int c = 0; c < W; c += 4) {
    const uint32_t x0 = s[c], x1 = s[c + 1], x2 = s[c + 2], x3 = s[c + 3];
    const uint32_t t = add(add(x0, x1), add(x2, x3));
    s[c] = add(add(t, x0), dbl(x1));
    s[c + 1] = add(add(t, x1), dbl(x2));
    s[c + 2] = add(add(t, x2), dbl(x3));
    s[c + 3] = add(add(t, x3), dbl(x0));
  }
  uint32_t col[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int i = 0; i < W; ++i) col[i & 3] = add(col[i & 3], s[i]);
#pragma unroll
  for (int i = 0; i < W; ++i) s[i] = add(s[i], col[i & 3]);
}

template <int W>
KB_HD void permute_plain(uint32_t (&s)[W], const Poseidon2Plain<W>& C) {
  mds_light_plain<W>(s);
#pragma unroll 1
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int i = 0; i < W; ++i) s[i] = sbox7(add(s[i], C.ext_init[r][i]));
    mds_light_plain<W>(s);
  }
#pragma unroll 1
  for (int r = 0; r < Poseidon2Plain<W>::RP; ++r) {
    s[0] = sbox7(add(s[0], C.internal[r]));
    uint32_t sum = s[0];
#pragma unroll
    for (int i = 1; i < W; ++i) sum = add(sum, s[i]);
#pragma unroll
    for (int i = 0; i < W; ++i) s[i] = add(mul(s[i], C.diag[i]), sum);
  }
#pragma unroll 1
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int i = 0; i < W; ++i) s[i] = sbox7(add(s[i], C.ext_term[r][i]));
    mds_light_plain<W>(s);
  }
}

// The same interface as the KoalaBear build's lazily reduced permutation:
// every output word is the reference's (OUT is accepted and ignored).
template <int W, uint32_t OUT = (1u << W) - 1u, bool OPAQUE = false>
__device__ __forceinline__ void permute(uint32_t (&s)[W]) {
  if constexpr (OPAQUE) permute_plain<W>(s, opaque_ref(plain_consts<W>()));
  else permute_plain<W>(s, plain_consts<W>());
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
  permute<16, (1u << 16) - 1u, OPAQUE>(s);
#pragma unroll
  for (int i = 0; i < 8; ++i) out[i] = s[i];
}

// Host: the diagonal in Montgomery form.
template <int W>
inline void plain_diag_build(Poseidon2Plain<W>& C) {
  for (int i = 0; i < W; ++i) {
    const uint32_t v = diag_value(W == 16 ? kV16[i] : kV24[i]);
    C.diag[i] = (uint32_t)(((uint64_t)v << 32) % P);
  }
}

}  // namespace kb
