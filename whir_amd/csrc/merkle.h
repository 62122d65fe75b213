// Poseidon2 Merkle tree with split-half pairing, replacing
// PackedPoseidonMerkleTree::commit / query (crates/merkle/src/poseidon_packed.rs:
// 96-143, 145-233, 268-304; crates/merkle/src/lib.rs:19-83):
//   leaf[i]     = PaddingFreeSponge<P24, 24, 16, 8>(row i, flattened base words)
//   parent[j]   = P16(layer[j] || layer[j + n/2])[0..8]
// Device layout: all layers concatenated, leaves first; layer l starts at digest
// offset sum_{i<l} H >> i (H = number of rows). Digests are 8 words (32 B).
#pragma once
#include "field.h"
#include "poseidon2.h"
#include "tx.h"
#if !defined(WHIR_FIELD_BABYBEAR)
#include "p2pair.h"
#endif

namespace kb {

constexpr int MERKLE_THREADS = 256;

// Leaf i: the sponge of row i (ROW_WORDS flattened base words), on this lane.
// OPAQUE: the round constants are loaded inside the caller's loop (opaque_ref).
template <int ROW_WORDS, bool OPAQUE = false>
__device__ __forceinline__ void leaf_hash_row(const uint32_t* __restrict__ rows, uint64_t i,
                                              uint32_t* __restrict__ leaves) {
  const uint32_t* row = rows + i * ROW_WORDS;
  uint32_t s[24];
#pragma unroll
  for (int j = 0; j < 24; ++j) s[j] = 0u;
  constexpr int CHUNKS = (ROW_WORDS + 15) / 16;
  // One permutation body for every chunk (the code of a P24 is ~35 KB):
  // it restores both the digest lanes and the capacity lanes carried over.
  constexpr uint32_t OUT = CHUNKS > 1 ? 0xFF00FFu : 0xFFu;
#pragma unroll 1
  for (int c = 0; c < CHUNKS; ++c) {
    constexpr int FULL = ROW_WORDS >= 16 ? 16 : ROW_WORDS;
    if constexpr (FULL % 4 == 0) {
#pragma unroll
      for (int j = 0; j < FULL; j += 4) {
        const uint4 v = *reinterpret_cast<const uint4*>(row + c * 16 + j);
        s[j] = v.x;
        s[j + 1] = v.y;
        s[j + 2] = v.z;
        s[j + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int j = 0; j < FULL; ++j) s[j] = row[c * 16 + j];
    }
    permute<24, OUT, OPAQUE>(s);
  }
  uint4* out = reinterpret_cast<uint4*>(leaves + i * 8);
  out[0] = make_uint4(s[0], s[1], s[2], s[3]);
  out[1] = make_uint4(s[4], s[5], s[6], s[7]);
}

// One sponge per lane. Rows of several chunks run the permutation in a loop,
// out of which LLVM would hoist the round constants into more scalar
// registers than there are (221 spilled through VGPR lanes, ~450 extra
// instructions per permutation), so they read the constants inside the loop
// (OPAQUE): 3 864 instead of 4 307 instructions per loop body and 90 instead
// of 104 VGPRs.
template <int ROW_WORDS>
__global__ void __launch_bounds__(MERKLE_THREADS) merkle_leaves_kernel(JP<const uint32_t> rows_, uint64_t n_rows,
                                                                   JP<uint32_t> leaves_) {
  const uint64_t i = (uint64_t)blockIdx.x * MERKLE_THREADS + threadIdx.x;
  if (i >= n_rows) return;
  leaf_hash_row<ROW_WORDS, (ROW_WORDS > 16)>(rows_, i, leaves_);
}

// parent[j] = C(prev[j], prev[j + half]), on this lane (OPAQUE as for
// leaf_hash_row).
template <bool OPAQUE = false>
__device__ __forceinline__ void layer_node(const uint32_t* __restrict__ prev, uint64_t half, uint64_t j,
                                           uint32_t* __restrict__ next) {
  uint32_t l[8], r[8], o[8];
  const uint4* pl = reinterpret_cast<const uint4*>(prev + j * 8);
  const uint4* pr = reinterpret_cast<const uint4*>(prev + (j + half) * 8);
  uint4 a = pl[0], b = pl[1], c = pr[0], d = pr[1];
  l[0] = a.x; l[1] = a.y; l[2] = a.z; l[3] = a.w; l[4] = b.x; l[5] = b.y; l[6] = b.z; l[7] = b.w;
  r[0] = c.x; r[1] = c.y; r[2] = c.z; r[3] = c.w; r[4] = d.x; r[5] = d.y; r[6] = d.z; r[7] = d.w;
  compress<OPAQUE>(l, r, o);
  uint4* po = reinterpret_cast<uint4*>(next + j * 8);
  po[0] = make_uint4(o[0], o[1], o[2], o[3]);
  po[1] = make_uint4(o[4], o[5], o[6], o[7]);
}

// parent[j] = C(prev[j], prev[j + half]) for j < half, one compression per lane.
__global__ void __launch_bounds__(MERKLE_THREADS) merkle_layer_kernel(JP<const uint32_t> prev_, uint64_t half,
                                                                  JP<uint32_t> next_) {
  const uint64_t j = (uint64_t)blockIdx.x * MERKLE_THREADS + threadIdx.x;
  if (j >= half) return;
  layer_node(prev_, half, j, next_);
}

#if !defined(WHIR_FIELD_BABYBEAR)  // two lanes per permutation: KoalaBear lazy arithmetic (p2pair.h)
// Two lanes per row (p2pair.h): lane 0 takes words 0..11 of each 16-word chunk,
// lane 1 words 12..15 (its words 0..3; its words 4..11 are the capacity).
// ROW_WORDS is a multiple of 16.
template <int ROW_WORDS>
__global__ void __launch_bounds__(MERKLE_THREADS) merkle_leaves_pair_kernel(JP<const uint32_t> rows_,
                                                                        uint64_t n_rows,
                                                                        JP<uint32_t> leaves_) {
  const uint32_t* __restrict__ rows = rows_;
  uint32_t* __restrict__ leaves = leaves_;
  static_assert(ROW_WORDS % 16 == 0, "whole chunks");
  const uint64_t i = ((uint64_t)blockIdx.x * MERKLE_THREADS + threadIdx.x) >> 1;
  if (i >= n_rows) return;  // both lanes of a pair leave together
  PairLane<24> L;
  L.init();
  const uint32_t* row = rows + i * ROW_WORDS;
  uint32_t s[12];
#pragma unroll
  for (int j = 0; j < 12; ++j) s[j] = 0u;
  constexpr int CHUNKS = ROW_WORDS / 16;
  constexpr uint32_t OUT = CHUNKS > 1 ? 0xFF00FFu : 0xFFu;
#pragma unroll 1
  for (int c = 0; c < CHUNKS; ++c) {
    const uint4* q = reinterpret_cast<const uint4*>(row + c * 16);
    if (!L.h) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const uint4 v = q[j];
        s[4 * j] = v.x;
        s[4 * j + 1] = v.y;
        s[4 * j + 2] = v.z;
        s[4 * j + 3] = v.w;
      }
    } else {
      const uint4 v = q[3];
      s[0] = v.x;
      s[1] = v.y;
      s[2] = v.z;
      s[3] = v.w;
    }
    permute_pair<24, OUT>(s, L);
  }
  if (!L.h) {
    uint4* out = reinterpret_cast<uint4*>(leaves + i * 8);
    out[0] = make_uint4(s[0], s[1], s[2], s[3]);
    out[1] = make_uint4(s[4], s[5], s[6], s[7]);
  }
}

// parent[j] = C(prev[j], prev[j + half]) on a lane pair: lane 0 holds the left
// digest (words 0..7), lane 1 the right one (words 8..15).
__global__ void __launch_bounds__(MERKLE_THREADS) merkle_layer_pair_kernel(JP<const uint32_t> prev_,
                                                                       uint64_t half, JP<uint32_t> next_) {
  const uint32_t* __restrict__ prev = prev_;
  uint32_t* __restrict__ next = next_;
  const uint64_t j = ((uint64_t)blockIdx.x * MERKLE_THREADS + threadIdx.x) >> 1;
  if (j >= half) return;
  PairLane<16> L;
  L.init();
  const uint4* src = reinterpret_cast<const uint4*>(prev + (j + (L.h ? half : 0)) * 8);
  const uint4 a = src[0], b = src[1];
  uint32_t s[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  permute_pair<16, 0xFFu>(s, L);
  if (!L.h) {
    uint4* po = reinterpret_cast<uint4*>(next + j * 8);
    po[0] = make_uint4(s[0], s[1], s[2], s[3]);
    po[1] = make_uint4(s[4], s[5], s[6], s[7]);
  }
}

#endif

// Upper layers with cooperative compressions (16 lanes per P16, tx.h), which
// cut the per-layer latency of the narrow top of the tree. Under split-half
// pairing the nodes of a layer that are congruent to w (mod G) are exactly the
// subtree below node w of the G-node layer, so workgroup w builds every layer
// from n0/2 nodes down to G nodes for its residue class without a grid-wide
// barrier: its n0/G nodes stay in LDS from layer to layer (each layer is also
// written out for the openings). Launch with G = 1 and one workgroup for the
// final layers.
// A workgroup is one wave per SIMD (16 compressions in flight), so a level of
// at most 16 compressions per workgroup costs one cooperative P16 latency; the
// host cuts the top of a tree into launches of at most COOP_MAX_NODES nodes
// per residue class (five levels each). A level of at most four compressions
// per workgroup runs them one per wavefront in the lower-latency one-wavefront
// form (permute16_wave, KoalaBear).
constexpr int COOP_THREADS = 256;
constexpr uint32_t COOP_MAX_NODES = 32;  // n0 / G

// Workgroup w's part (COOP_THREADS threads); on return s_node[0..8) holds its
// node of the G-node layer and every thread has passed the last barrier.
__device__ __forceinline__ void coop_subtrees_body(uint32_t* layer, uint32_t n0, uint32_t G, uint32_t w,
                                                   const CoopP16& K, uint32_t* s_node) {
  const uint32_t slot = threadIdx.x >> 4, li = threadIdx.x & 15;
  uint32_t kk = n0 / G;  // nodes of this residue class in the current layer
  for (uint32_t e = threadIdx.x; e < kk * 8; e += COOP_THREADS)
    s_node[e] = layer[(uint64_t)(w + (e >> 3) * G) * 8 + (e & 7)];
  __syncthreads();
  uint32_t* next = layer + (uint64_t)n0 * 8;
  uint32_t n = n0;
  while (kk > 1) {
    const uint32_t h = kk >> 1;
    // compression t reads nodes t and t + h and overwrites node t: no other
    // compression of this layer touches either, so the update is in place
#if !defined(WHIR_FIELD_BABYBEAR)
    if (h <= COOP_THREADS / 64) {
      const uint32_t t = threadIdx.x >> 6, lane = threadIdx.x & 63;
      if (t < h) {  // wavefront t: compression t (words in lanes 0..15)
        uint32_t v = li < 8 ? s_node[t * 8 + li] : s_node[(t + h) * 8 + (li - 8)];
        v = permute16_wave_shared(v, K);
        if (lane < 8) {
          s_node[t * 8 + lane] = v;
          next[(uint64_t)(w + t * G) * 8 + lane] = v;
        }
      }
    } else
#endif
      for (uint32_t t = slot; t < h; t += COOP_THREADS / 16) {
        uint32_t v = li < 8 ? s_node[t * 8 + li] : s_node[(t + h) * 8 + (li - 8)];
        permute16_coop(v, K);
        if (li < 8) {
          s_node[t * 8 + li] = v;
          next[(uint64_t)(w + t * G) * 8 + li] = v;
        }
      }
    __syncthreads();
    n >>= 1;
    next += (uint64_t)n * 8;
    kk = h;
  }
}

__global__ void __launch_bounds__(COOP_THREADS) merkle_coop_subtrees_kernel(JP<uint32_t> layer_, uint32_t n0, uint32_t G) {
  uint32_t* layer = layer_;
  __shared__ uint32_t s_node[COOP_MAX_NODES * 8];
  CoopP16 K;
  K.load();
  coop_subtrees_body(layer, n0, G, blockIdx.x, K, s_node);
}
}  // namespace kb
