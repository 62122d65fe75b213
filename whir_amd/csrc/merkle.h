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

namespace kb {

constexpr int MERKLE_THREADS = 256;

// One sponge per lane; ROW_WORDS is the flattened row length (base words).
template <int ROW_WORDS>
__global__ void __launch_bounds__(MERKLE_THREADS) merkle_leaves_kernel(const uint32_t* __restrict__ rows, uint64_t n_rows,
                                                                   uint32_t* __restrict__ leaves) {
  const uint64_t i = (uint64_t)blockIdx.x * MERKLE_THREADS + threadIdx.x;
  if (i >= n_rows) return;
  const uint32_t* row = rows + i * ROW_WORDS;
  uint32_t s[24];
#pragma unroll
  for (int j = 0; j < 24; ++j) s[j] = 0u;
  constexpr int CHUNKS = (ROW_WORDS + 15) / 16;
#pragma unroll
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
    permute<24>(s);
  }
  uint4* out = reinterpret_cast<uint4*>(leaves + i * 8);
  out[0] = make_uint4(s[0], s[1], s[2], s[3]);
  out[1] = make_uint4(s[4], s[5], s[6], s[7]);
}

// parent[j] = C(prev[j], prev[j + half]) for j < half.
__global__ void __launch_bounds__(MERKLE_THREADS) merkle_layer_kernel(const uint32_t* __restrict__ prev, uint64_t half,
                                                                  uint32_t* __restrict__ next) {
  const uint64_t j = (uint64_t)blockIdx.x * MERKLE_THREADS + threadIdx.x;
  if (j >= half) return;
  uint32_t l[8], r[8], o[8];
  const uint4* pl = reinterpret_cast<const uint4*>(prev + j * 8);
  const uint4* pr = reinterpret_cast<const uint4*>(prev + (j + half) * 8);
  uint4 a = pl[0], b = pl[1], c = pr[0], d = pr[1];
  l[0] = a.x; l[1] = a.y; l[2] = a.z; l[3] = a.w; l[4] = b.x; l[5] = b.y; l[6] = b.z; l[7] = b.w;
  r[0] = c.x; r[1] = c.y; r[2] = c.z; r[3] = c.w; r[4] = d.x; r[5] = d.y; r[6] = d.z; r[7] = d.w;
  compress(l, r, o);
  uint4* po = reinterpret_cast<uint4*>(next + j * 8);
  po[0] = make_uint4(o[0], o[1], o[2], o[3]);
  po[1] = make_uint4(o[4], o[5], o[6], o[7]);
}

// All layers from a layer of n <= 2*MERKLE_THREADS nodes up to the root, in one
// workgroup; writes each layer at its place in the concatenated layout.
__global__ void __launch_bounds__(MERKLE_THREADS) merkle_top_kernel(uint32_t* layer, uint64_t n) {
  while (n > 1) {
    const uint64_t half = n / 2;
    uint32_t* next = layer + n * 8;
    for (uint64_t j = threadIdx.x; j < half; j += MERKLE_THREADS) {
      uint32_t l[8], r[8], o[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        l[i] = layer[j * 8 + i];
        r[i] = layer[(j + half) * 8 + i];
      }
      compress(l, r, o);
#pragma unroll
      for (int i = 0; i < 8; ++i) next[j * 8 + i] = o[i];
    }
    __threadfence_block();
    __syncthreads();
    layer = next;
    n = half;
  }
}

}  // namespace kb
