// Device self-checks of a proof (whir_prover_set_checks): the reference's
// debug assertions on the prover path, evaluated on the device beside the
// proof's own kernels. A failed check adds 1 to the status word after the
// proof log (tail word 3), which the result calls report as WHIR_ECHECK.
//
//  * Merkle openings (MerkleTree::query's debug block, crates/merkle/src/
//    poseidon_packed.rs:158-230): every opening written to the proof hashes
//    (sponge of its row, then its siblings bottom-up) to the tree's root; its
//    first sibling is the hash of the sibling row; its row is the codeword's.
//  * Sumcheck rounds (debug_assert_eq!(*sum, self.prod()), crates/whir/src/
//    sumcheck/expr.rs:68): after a round's fold with challenge r, <f, w> of
//    the folded arrays equals the round polynomial at r, extrapolate_012(c0,
//    s - c0, c2, r), with s = <f, w> and (c0, c2) = coeffs(f, w) of the arrays
//    before the fold.
#pragma once
#include "protocol.h"

namespace kb {

constexpr uint32_t CHECK_THREADS = 1024;
constexpr uint32_t OPEN_CHECK_THREADS = 64;

__global__ void check_reset_kernel(JP<uint32_t> fail) {
  if (threadIdx.x == 0) *fail.get() = 0u;
}

// PaddingFreeSponge<P24, 24, 16, 8> of a row (the Merkle leaf hash).
__device__ __forceinline__ void sponge_row(const uint32_t* row, uint32_t row_words, uint32_t (&out)[8]) {
  uint32_t s[24];
#pragma unroll
  for (int j = 0; j < 24; ++j) s[j] = 0u;
#pragma unroll 1
  for (uint32_t c = 0; c < row_words; c += 16) {
    const uint32_t take = row_words - c < 16 ? row_words - c : 16u;
#pragma unroll
    for (uint32_t j = 0; j < 16; ++j)
      if (j < take) s[j] = row[c + j];
    permute<24>(s);
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) out[j] = s[j];
}

// One lane per query of the openings stir_query_kernel just wrote at the
// transcript's current log position (the next transcript kernel advances it).
__global__ void __launch_bounds__(OPEN_CHECK_THREADS) opening_check_kernel(
    JP<const DevTx> t_, JP<const uint32_t> log_, JP<const uint32_t> idx_, JP<const uint32_t> count_,
    JP<const uint32_t> cw_, uint32_t row_words, JP<const uint32_t> tree_, uint32_t depth, JP<uint32_t> fail_) {
  const DevTx* t = t_;
  const uint32_t* log = log_;
  const uint32_t* cw = cw_;
  const uint32_t q = blockIdx.x * OPEN_CHECK_THREADS + threadIdx.x;
  if (q >= *count_.get()) return;
  const uint32_t per = row_words + 8 * depth;
  const uint64_t base = (uint64_t)t->log_words + (uint64_t)q * per;
  if (base + per > t->log_cap) return;  // not written: the transcript reports the overflow
  const uint32_t index = idx_.get()[q];
  bool ok = true;
  for (uint32_t w = 0; w < row_words; ++w) ok &= log[base + w] == cw[(uint64_t)index * row_words + w];
  // the leaf from the row in the proof, and (depth >= 1) the sibling leaf from the codeword
  uint32_t node[8], sib_leaf[8];
#pragma unroll 1
  for (uint32_t h = 0; h < (depth ? 2u : 1u); ++h) {
    const uint32_t* row = h == 0 ? log + base : cw + (uint64_t)(index ^ (1u << (depth - 1))) * row_words;
    uint32_t d[8];
    sponge_row(row, row_words, d);
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      if (h == 0) node[w] = d[w];
      else sib_leaf[w] = d[w];
    }
  }
  uint32_t idx_asc = index;
#pragma unroll 1
  for (uint32_t i = 0; i < depth; ++i) {
    const uint32_t mid = 1u << (depth - i - 1);
    uint32_t sib[8], l[8], r[8], o[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      sib[w] = log[base + row_words + 8 * i + w];
      if (i == 0) ok &= sib[w] == sib_leaf[w];
      const bool right = (idx_asc & mid) != 0;  // split-half pairing: node idx_asc is a right child
      l[w] = right ? sib[w] : node[w];
      r[w] = right ? node[w] : sib[w];
    }
    compress(l, r, o);
#pragma unroll
    for (int w = 0; w < 8; ++w) node[w] = o[w];
    idx_asc &= mid - 1;
  }
  const uint32_t* root = tree_.get() + ((2ull << depth) - 2) * 8;
#pragma unroll
  for (int w = 0; w < 8; ++w) ok &= node[w] == root[w];
  if (!ok) atomicAdd(fail_.get(), 1u);
}

// Block sums of <f, w>, c0 and c2 over n (even) entries, then the check
// against the previous call's values when r is given; the values are kept in
// the job's scalars for the next call.
__global__ void __launch_bounds__(CHECK_THREADS) sumcheck_check_kernel(JP<const Ext> f_, JP<const Ext> w_, uint64_t n,
                                                                       JP<const Ext> r_, JP<JobScalars> js_,
                                                                       JP<uint32_t> fail_) {
  const Ext* f = f_;
  const Ext* w = w_;
  JobScalars* js = js_;
  __shared__ Ext sh[CHECK_THREADS];
  const uint32_t tid = threadIdx.x;
  Ext prod = ezero(), c0 = ezero(), c2 = ezero();
  for (uint64_t i = tid; i < n / 2; i += CHECK_THREADS) {
    const Ext f0 = ld_ext(f + 2 * i), f1 = ld_ext(f + 2 * i + 1);
    const Ext w0 = ld_ext(w + 2 * i), w1 = ld_ext(w + 2 * i + 1);
    const Ext e0 = emul(f0, w0);
    c0 = eadd(c0, e0);
    prod = eadd(prod, eadd(e0, emul(f1, w1)));
    c2 = eadd(c2, emul(esub(edbl(f1), f0), esub(edbl(w1), w0)));
  }
  Ext tot[3];
#pragma unroll 1
  for (int v = 0; v < 3; ++v) {
    sh[tid] = v == 0 ? prod : v == 1 ? c0 : c2;
    __syncthreads();
    for (uint32_t s = CHECK_THREADS / 2; s > 0; s >>= 1) {
      if (tid < s) sh[tid] = eadd(sh[tid], sh[tid + s]);
      __syncthreads();
    }
    tot[v] = sh[0];
    __syncthreads();
  }
  if (tid == 0) {
    if (r_.raw) {
      Ext l[3];
      lagrange012(ld_ext(r_.get()), l);
      const Ext want = eadd(eadd(emul(js->chk_c0, l[0]), emul(esub(js->chk_prod, js->chk_c0), l[1])),
                            emul(js->chk_c2, l[2]));
      if (!eeq(want, tot[0])) atomicAdd(fail_.get(), 1u);
    }
    js->chk_prod = tot[0];
    js->chk_c0 = tot[1];
    js->chk_c2 = tot[2];
  }
}

}  // namespace kb
