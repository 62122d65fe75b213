// Column-batched DFT over KoalaBear (Reed–Solomon encoding), replacing Plonky3
// Radix2DFTSmallBatch::dft_batch / dft_algebra_batch as called at
// crates/whir/src/whir.rs:34,59:
//   out[i][c] = sum_{j < n_in} in[j][c] * w^(i*j),  w = two_adic_generator(log2 N)
// for an N-row, C-column row-major matrix whose rows >= n_in are zero (the
// rate padding, whir.rs:32,57). An extension matrix is the same transform on
// its 4*C base columns.
//
// Decomposition: N = L_1 * ... * L_P (four-step applied recursively). Pass p
// works on blocks of M_{p-1} = L_p * M_p rows; for block b, column q (q spans
// the M_p rows times C columns that follow each other in memory) it computes a
// length-L_p DFT over the stride-M_p rows, then multiplies output i_p by
// w_{M_{p-1}}^(i_p * j') (j' = row offset inside the block). Positions end up in
// mixed-radix digit order; the last pass (M_P = 1) writes each output to its
// natural row i_1 + L_1 i_2 + L_1 L_2 i_3 + ...
//
// Per workgroup: an L x Q tile (Q >= 32 consecutive words per row segment, so
// every global access is a >= 128 B contiguous run) staged in LDS, radix-2 DIT
// butterflies on bit-reversed LDS rows. Pass 1 skips loads of zero rows.
// Twiddles: w_{2^24}^e = T_lo[e & 4095] * T_hi[e >> 12] (two 16 KB tables).
#pragma once
#include "field.h"

namespace kb {

constexpr int NTT_THREADS = 256;
constexpr int NTT_MAX_LDS_WORDS = 16384;  // 64 KiB tile

struct NttPass {
  const uint32_t* src;
  uint32_t* dst;
  const uint32_t* tw_lo;  // w24^i, i < 4096
  const uint32_t* tw_hi;  // w24^(4096 i), i < 4096
  uint64_t n_in_rows;     // rows of src that may be non-zero (first pass)
  uint32_t log_n, log_c;  // N = 2^log_n rows, C = 2^log_c columns
  uint32_t log_l, log_m;  // this pass: L = 2^log_l, remaining M = 2^log_m
  uint32_t log_q;         // tile width (words)
  uint32_t first, last;
  uint32_t n_pass;
  uint32_t logl[4];  // all pass sizes (for the last pass's digit order)
};

__device__ __forceinline__ uint32_t tw24(const uint32_t* lo, const uint32_t* hi, uint32_t e) {
  return mul(lo[e & 4095u], hi[e >> 12]);
}

template <int LOGL>
__global__ void __launch_bounds__(NTT_THREADS) ntt_pass_kernel(NttPass a) {
  extern __shared__ uint32_t lds[];
  constexpr uint32_t L = 1u << LOGL;
  const uint32_t Q = 1u << a.log_q;
  const uint32_t C = 1u << a.log_c;
  const uint32_t tid = threadIdx.x;
  uint32_t* tile = lds;            // [L][Q]
  uint32_t* twl = lds + (L << a.log_q);  // L/2 twiddles of order L

  // inner twiddles w_L^e
  for (uint32_t e = tid; e < L / 2; e += NTT_THREADS) twl[e] = tw24(a.tw_lo, a.tw_hi, e << (24 - LOGL));

  const uint64_t n_elems_tile = (uint64_t)L << a.log_q;
  uint64_t blk_base = 0;       // first element of the block (non-last) / of the tile (last)
  uint64_t q0 = 0;             // first q of the tile inside its block (non-last)
  uint64_t out_row0 = 0;       // last pass: natural row of (t=0, i_P=0)
  if (!a.last) {
    // grid: (number of blocks) x (q tiles per block)
    const uint64_t q_tiles = ((uint64_t)1 << (a.log_m + a.log_c)) >> a.log_q;
    const uint64_t b = blockIdx.x / q_tiles;
    q0 = (uint64_t)(blockIdx.x % q_tiles) << a.log_q;
    blk_base = b << (a.log_l + a.log_m + a.log_c);
  } else {
    // tile = T consecutive values of the first digit i_1 (others fixed);
    // Q = T * C words per transform row.
    const uint32_t log_t = a.log_q - a.log_c;
    const uint32_t log_l1 = a.n_pass > 1 ? a.logl[0] : 0;
    const uint32_t t_tiles = (1u << log_l1) >> log_t;  // tiles over i_1
    const uint64_t w = blockIdx.x;
    const uint64_t i1_0 = (w % (t_tiles ? t_tiles : 1)) << log_t;
    uint64_t rest = w / (t_tiles ? t_tiles : 1);  // digits i_2..i_{P-1}, i_2 least significant
    // b_rest: the same digits with i_2 most significant
    uint64_t b_rest = 0;
    uint32_t log_s1 = 0;  // log of L_2 * ... * L_{P-1}
    for (uint32_t d = 1; d + 1 < a.n_pass; ++d) log_s1 += a.logl[d];
    {
      uint64_t r = rest;
      uint32_t shift = log_s1;
      for (uint32_t d = 1; d + 1 < a.n_pass; ++d) {
        const uint32_t ld = a.logl[d];
        const uint64_t digit = r & (((uint64_t)1 << ld) - 1);
        r >>= ld;
        shift -= ld;
        b_rest |= digit << shift;
      }
    }
    const uint64_t b0 = (i1_0 << log_s1) + b_rest;  // block of t = 0; t advances by 2^log_s1
    blk_base = b0;  // reused below as the block index
    out_row0 = i1_0 + (rest << log_l1);
  }
  __syncthreads();

  // ---- load (into bit-reversed rows) ----
  for (uint64_t e = tid; e < n_elems_tile; e += NTT_THREADS) {
    uint32_t jp, q;
    uint64_t gidx;
    bool zero = false;
    if (!a.last) {
      jp = (uint32_t)(e >> a.log_q);
      q = (uint32_t)(e & (Q - 1));
      const uint64_t qq = q0 + q;
      gidx = blk_base + ((uint64_t)jp << (a.log_m + a.log_c)) + qq;
      if (a.first) {
        const uint64_t row = ((uint64_t)jp << a.log_m) + (qq >> a.log_c);
        zero = row >= a.n_in_rows;
      }
    } else {
      // contiguous walk over T blocks of L rows x C columns
      const uint32_t col = (uint32_t)(e & (C - 1));
      const uint64_t rowt = e >> a.log_c;         // t * L + jp
      jp = (uint32_t)(rowt & (L - 1));
      const uint64_t t = rowt >> LOGL;
      q = (uint32_t)((t << a.log_c) + col);
      uint32_t log_s1 = 0;
      for (uint32_t d = 1; d + 1 < a.n_pass; ++d) log_s1 += a.logl[d];
      const uint64_t b = blk_base + (t << log_s1);
      const uint64_t row = (b << LOGL) + jp;
      gidx = (row << a.log_c) + col;
      if (a.first) zero = row >= a.n_in_rows;
    }
    const uint32_t v = zero ? 0u : a.src[gidx];
    const uint32_t jr = __brev(jp) >> (32 - LOGL);
    tile[(jr << a.log_q) + q] = v;
  }
  __syncthreads();

  // ---- radix-2 DIT stages ----
  const uint32_t n_bfly = (L / 2) << a.log_q;
#pragma unroll 1
  for (int s = 0; s < LOGL; ++s) {
    const uint32_t half = 1u << s;
    for (uint32_t e = tid; e < n_bfly; e += NTT_THREADS) {
      const uint32_t q = e & (Q - 1);
      const uint32_t bi = e >> a.log_q;              // butterfly index in [0, L/2)
      const uint32_t j = bi & (half - 1);
      const uint32_t g = bi >> s;
      const uint32_t i0 = (g << (s + 1)) + j;
      const uint32_t i1 = i0 + half;
      const uint32_t w = twl[j << (LOGL - 1 - s)];
      const uint32_t u = tile[(i0 << a.log_q) + q];
      const uint32_t v = mul(tile[(i1 << a.log_q) + q], w);
      tile[(i0 << a.log_q) + q] = add(u, v);
      tile[(i1 << a.log_q) + q] = sub(u, v);
    }
    __syncthreads();
  }

  // ---- twiddle + store ----
  for (uint64_t e = tid; e < n_elems_tile; e += NTT_THREADS) {
    if (!a.last) {
      const uint32_t ip = (uint32_t)(e >> a.log_q);
      const uint32_t q = (uint32_t)(e & (Q - 1));
      const uint64_t qq = q0 + q;
      uint32_t v = tile[(ip << a.log_q) + q];
      // w_{M_{p-1}}^(ip * j'), j' = qq >> log_c; exponent in the order-2^24 group
      const uint64_t jprime = qq >> a.log_c;
      const uint32_t log_mprev = a.log_l + a.log_m;
      const uint64_t ex = ((uint64_t)ip * jprime) & (((uint64_t)1 << log_mprev) - 1);
      if (ex) v = mul(v, tw24(a.tw_lo, a.tw_hi, (uint32_t)(ex << (24 - log_mprev))));
      a.dst[blk_base + ((uint64_t)ip << (a.log_m + a.log_c)) + qq] = v;
    } else {
      // walk outputs so that consecutive threads write consecutive rows:
      // e = (ip * T + t) * C + col
      const uint32_t col = (uint32_t)(e & (C - 1));
      const uint64_t r = e >> a.log_c;
      const uint32_t log_t = a.log_q - a.log_c;
      const uint64_t t = r & (((uint64_t)1 << log_t) - 1);
      const uint32_t ip = (uint32_t)(r >> log_t);
      const uint32_t q = (uint32_t)((t << a.log_c) + col);
      const uint32_t v = tile[(ip << a.log_q) + q];
      const uint64_t row = out_row0 + t + ((uint64_t)ip << (a.log_n - LOGL));
      a.dst[(row << a.log_c) + col] = v;
    }
  }
}

}  // namespace kb
