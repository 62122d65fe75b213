// Poseidon2 with two lanes per permutation. Lane h = threadIdx.x & 1 of a lane
// pair holds the state words [h H, (h + 1) H), H = W / 2 (whole 4-word chunks).
// The arithmetic, the constants and every 64-bit sum are those of
// permute_lazy (poseidon2.h), so its bounds hold unchanged; the two lanes
// exchange their MDS-light column sums and their internal-layer half sums with
// one DPP move inside the quad. Per lane this is about half of permute_lazy's
// instructions, so a launch over the same work has twice as many, half as long
// instruction streams: it pays where one stream's latency is exposed (trees of
// a few thousand waves, leaf sponges of several chunks).
#pragma once
#include "poseidon2.h"
#include "tx.h"

namespace kb {

__device__ __forceinline__ uint64_t pair_swap64(uint64_t v) {
  const uint32_t lo = dpp<DPP_QUAD_XOR1>((uint32_t)v), hi = dpp<DPP_QUAD_XOR1>((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// t + (h ? b : a) for wave-uniform a, b.
__device__ __forceinline__ uint64_t add_sel(uint64_t t, uint32_t a, uint32_t b, bool h) {
  return mad_k<1>(h ? b : a, t);
}

template <int W>
struct PairLane {
  static constexpr int H = W / 2;
  bool h;         // this lane holds the second half
  uint32_t d[H];  // diagonal of the lane's words; 2 for word 0 (applied to 2p - x0)
  __device__ __forceinline__ void init() {
    h = (threadIdx.x & 1) != 0;
#pragma unroll
    for (int i = 0; i < H; ++i) d[i] = h ? Poseidon2Lazy<W>::diag(H + i) : (i == 0 ? 2u : Poseidon2Lazy<W>::diag(i));
  }
};

// The two interleaved sum chains of a lane's half state, sa = x0 + x2 + ...
// and sb = x1 + x3 + ..., as one asm block on the device (see sum_chains).
template <int H>
__device__ __forceinline__ void pair_sum_chains(const uint32_t (&x)[H], uint64_t& sa, uint64_t& sb) {
  if constexpr (H == 8) {
    asm("v_mad_u64_u32 %0, vcc, %2, 1, 0\n"
        "v_mad_u64_u32 %1, vcc, %3, 1, 0\n"
        "v_mad_u64_u32 %0, vcc, %4, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %5, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %6, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %7, 1, %1\n"
        "v_mad_u64_u32 %0, vcc, %8, 1, %0\n"
        "v_mad_u64_u32 %1, vcc, %9, 1, %1"
        : "=&v"(sa), "=&v"(sb)
        : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7])
        : "vcc");
    return;
  }
  else if constexpr (H == 12) {
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
        "v_mad_u64_u32 %1, vcc, %13, 1, %1"
        : "=&v"(sa), "=&v"(sb)
        : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7]), "v"(x[8]), "v"(x[9]), "v"(x[10]), "v"(x[11])
        : "vcc");
    return;
  }
  sa = x[0];
  sb = x[1];
  for (int j = 2; j < H; j += 2) {
    sa += x[j];
    sb += x[j + 1];
  }
}

// x <- REDC(MDS-light(x) + add) as mds_lazy, on the lane's half of the state.
template <int W, int ADD>
__device__ __forceinline__ void mds_pair(uint32_t (&x)[W / 2], const uint32_t* add, const PairLane<W>& L) {
  constexpr int H = W / 2;
  uint64_t y[H];
#pragma unroll
  for (int c = 0; c < H; c += 4) {
    uint64_t yc[4];
    m4_chunk(x[c], x[c + 1], x[c + 2], x[c + 3], yc);
#pragma unroll
    for (int a = 0; a < 4; ++a) y[c + a] = yc[a];
  }
  uint64_t cs[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    cs[a] = y[a];
#pragma unroll
    for (int c = 4; c < H; c += 4) cs[a] += y[c + a];
    cs[a] += pair_swap64(cs[a]);  // + the other half's chunks
  }
#pragma unroll
  for (int i = 0; i < H; ++i) {
    uint64_t t = y[i] + cs[i & 3];
    if constexpr (ADD == LAZY_ADD_ALL) t = add_sel(t, add[i], add[H + i], L.h);
    if constexpr (ADD == LAZY_ADD_FIRST) {
      if (i == 0) t = add_sel(t, add[0], 0u, L.h);
    }
    x[i] = redc(t);
  }
}

template <int W, int ADD>
__device__ __forceinline__ void external_pair(uint32_t (&x)[W / 2], const uint32_t* add, const PairLane<W>& L) {
#pragma unroll
  for (int i = 0; i < W / 2; ++i) x[i] = mulr(mulr(x[i], x[i]), x[i]);
  mds_pair<W, ADD>(x, add, L);
}

// internal_lazy on the lane's half: word 0 of the first half takes the S-box
// (times kappa) and 2 (2p - x0); the sum is the two halves' sums.
template <int W, bool LAST>
__device__ __forceinline__ void internal_pair(uint32_t (&x)[W / 2], int r, const Poseidon2Lazy<W>& C,
                                              const PairLane<W>& L) {
  constexpr int H = W / 2;
  const uint32_t c0 = mulr(mulr(mulr(x[0], x[0]), x[0]), C.kappa[r]);
  x[0] = L.h ? x[0] : c0;
  uint64_t sa, sb;
  pair_sum_chains<H>(x, sa, sb);
  uint64_t sum = sa + sb;
  sum += pair_swap64(sum);
  const uint32_t y0 = L.h ? x[0] : 2u * P - x[0];
  uint64_t t0 = mad64(y0, L.d[0], sum);
  if constexpr (LAST) {
    t0 = add_sel(t0, C.lin_add[5][0], C.lin_add[5][H], L.h);
    x[0] = canon2(redc(t0));
  } else {
    t0 = add_sel(t0, C.int_add[r + 1], 0u, L.h);
    x[0] = redc(t0);
  }
#pragma unroll
  for (int i = 1; i < H; ++i) {
    uint64_t t = mad64(x[i], L.d[i], sum);
    if constexpr (LAST) {
      t = add_sel(t, C.lin_add[5][i], C.lin_add[5][H + i], L.h);
      x[i] = canon2(redc(t));
    } else {
      x[i] = redc(t);
    }
  }
}

template <int W, int R>
__device__ __forceinline__ void internal_rounds_pair(uint32_t (&s)[W / 2], const Poseidon2Lazy<W>& C,
                                                     const PairLane<W>& L) {
  if constexpr (R < Poseidon2Lazy<W>::RP - 1) {
    internal_pair<W, false>(s, R, C, L);
    internal_rounds_pair<W, R + 1>(s, C, L);
  }
}

// The permutation on a lane pair. Out: the words whose global index is in the
// mask OUT hold the reference's output words in [0, p) (a word is fixed when
// either half wants it; the others hold scaled values and must not be used).
template <int W, uint32_t OUT>
__device__ __forceinline__ void permute_pair(uint32_t (&s)[W / 2], const PairLane<W>& L) {
  constexpr int H = W / 2, RP = Poseidon2Lazy<W>::RP;
  const Poseidon2Lazy<W>& C = lazy_consts<W>();
  mds_pair<W, LAZY_ADD_ALL>(s, C.lin_add[0], L);
#pragma unroll
  for (int r = 0; r < 3; ++r) external_pair<W, LAZY_ADD_ALL>(s, C.lin_add[r + 1], L);
  external_pair<W, LAZY_ADD_FIRST>(s, C.lin_add[4], L);
  internal_rounds_pair<W, 0>(s, C, L);
  internal_pair<W, true>(s, RP - 1, C, L);
#pragma unroll
  for (int r = 0; r < 3; ++r) external_pair<W, LAZY_ADD_ALL>(s, C.lin_add[6 + r], L);
  external_pair<W, LAZY_ADD_NONE>(s, nullptr, L);
#pragma unroll
  for (int i = 0; i < H; ++i)
    if (((OUT >> i) | (OUT >> (H + i))) & 1u) s[i] = canon2(mulr(s[i], C.fix));
}

}  // namespace kb
