// A whole WHIR round for a small round polynomial (2^kc <= 2^10 extension
// elements) in one workgroup, and the final phase (final polynomial, last
// openings, checkpoint) in one more. The transcript, the arithmetic and the
// memory they leave behind are those of the kernel sequence prover_enqueue
// runs for a large round (whir.rs:205-312, 314-355):
//   transcript.write(root); n_ood x (draw z; write <f, eq(expand(kc, z))>);
//   n_stir draws -> sorted distinct indices -> openings of the previous tree
//   as hints; draw alpha; w += sum_o alpha^o eq(z_o) + sum_q alpha^(n_ood+q)
//   pow(omega^index_q) (combine.rs:37-64); f x (write c0, c2; draw r; fold)
//   (expr.rs:52-70, poly.rs:131-142).
// At this size a round is about twenty launches whose ~5 us floors are most of
// its time; here wave 0 runs the transcript and the whole workgroup does the
// small data-parallel steps between its draws. The round polynomial and its
// weights stay in global memory (L2-resident, a few tens of KiB); LDS holds
// the OOD eq tables, the STIR indexes and the pow tables.
#pragma once
#include "protocol.h"

namespace kb {

// Measured on MI355X (k = 24): the round of 2^8 elements takes 86 us here
// (91 us with 256 threads) against 133 us as a kernel sequence; at 2^12
// elements one workgroup is too small for the claim combination (34 us against
// 10 us on the grid), hence the 2^10 limit.
constexpr int SR_THREADS = 512;
constexpr int SR_WAVES = SR_THREADS / 64;
constexpr uint32_t SR_MAX_LOG_N = 10;
constexpr int SR_MAX_OOD = 8;
constexpr uint32_t SC_TAIL_MAX_LOG_N = 12;
constexpr uint32_t SR_POW_BYTES = 96 * 1024;  // pow tables: count * (4 * 2^m0 + 16 * 2^m1) bytes

// Host: whether a round of 2^kc elements with these claim counts fits the
// workgroup's tables (powR is placed at a 16-byte offset after powL).
inline bool small_round_fits(int kc, int n_ood, int n_stir) {
  if (kc < 1 || kc > (int)SR_MAX_LOG_N || n_ood > SR_MAX_OOD || n_stir < 1 || n_stir > MAX_STIR) return false;
  const uint64_t m0 = kc / 2, m1 = kc - kc / 2;
  const uint64_t l_bytes = (((uint64_t)n_stir * (4ull << m0)) + 15) & ~15ull;
  return l_bytes + (uint64_t)n_stir * (16ull << m1) <= SR_POW_BYTES;
}

struct SmallRoundArgs {
  JP<const uint32_t> root;      // this round's commitment root (8 words)
  JP<Ext> f_cur;                // round polynomial, 2^kc extension elements
  JP<Ext> w_cur;                // its sumcheck weights
  JP<Ext> f_alt;                // ping-pong buffers for the folds
  JP<Ext> w_alt;
  uint32_t kc, n_ood, n_stir, kfd, f, rs_idx, row_words;
  JP<const uint32_t> prev_cw;   // previous round's codeword rows (row_words words each)
  JP<const uint32_t> prev_tree; // its tree layers (depth kfd)
};

// Sums of two Ext per thread over the workgroup, into wave 0 (the transcript's
// wave). s_wave: 2 * (blockDim.x / 64) entries.
__device__ __forceinline__ void block_sum2_to_wave0(Ext& a, Ext& b, Ext* s_wave) {
  a = wave_sum_ext(a);
  b = wave_sum_ext(b);
  const uint32_t wave = threadIdx.x >> 6, nw = blockDim.x / 64;
  if ((threadIdx.x & 63) == 0) {
    s_wave[wave] = a;
    s_wave[nw + wave] = b;
  }
  __syncthreads();
  if (wave == 0)
    for (uint32_t i = 1; i < nw; ++i) {
      a = eadd(a, s_wave[i]);
      b = eadd(b, s_wave[nw + i]);
    }
}

// log[base + e] = word e of the openings of idx[0..count) (opening_word),
// e < count * per, by the whole workgroup with eight loads in flight per thread.
__device__ __forceinline__ void write_openings(uint32_t* log, uint32_t base, const uint32_t* idx, uint32_t count,
                                               const uint32_t* cw, uint32_t row_words, const uint32_t* tree,
                                               uint32_t depth) {
  const uint32_t per = row_words + 8 * depth, total = count * per, T = blockDim.x;
  for (uint32_t e0 = threadIdx.x; e0 < total; e0 += 8 * T) {
    uint32_t v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t e = e0 + (uint32_t)u * T;
      v[u] = e < total ? opening_word(cw, row_words, tree, depth, idx[e / per], e % per) : 0u;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t e = e0 + (uint32_t)u * T;
      if (e < total) log[base + e] = v[u];
    }
  }
}

// eq table of the m <= 6 coordinates z (LDS) into out[0..2^m) (LDS), by
// doubling: out[j + h] = out[j] z_i, out[j] -= that.
__device__ __forceinline__ void eq_small(const Ext* z, uint32_t m, Ext* out) {
  if (threadIdx.x == 0) out[0] = eone();
  __syncthreads();
  for (uint32_t i = 0; i < m; ++i) {
    const uint32_t h = 1u << i;
    if (threadIdx.x < h) {
      const Ext lo = out[threadIdx.x];
      const Ext hi = emul(lo, z[i]);
      out[threadIdx.x + h] = hi;
      out[threadIdx.x] = esub(lo, hi);
    }
    __syncthreads();
  }
}

// f sumcheck rounds of the whole workgroup on a polynomial of nn extension
// elements and its weights (expr.rs:52-70, poly.rs:131-142): per round
// c0 = sum w[2i] f[2i], c2 = sum (2w[2i+1]-w[2i])(2f[2i+1]-f[2i]), written by
// wave 0, the challenge r drawn into js->rs[rs_idx + round], then both arrays
// folded, a[i] = a[2i] + r (a[2i+1] - a[2i]), into the other buffer of the pair.
// After f rounds the folded arrays are in (F, Wt) if f is even, else in
// (Fo, Wo). Ends with a barrier.
// Launched with SR_THREADS threads.
__device__ __forceinline__ void sumcheck_block(WaveTx& tx, JobScalars* js, Ext* F, Ext* Wt, Ext* Fo, Ext* Wo,
                                               uint32_t nn, uint32_t f, uint32_t rs_idx, Ext* s_wave, Ext& s_x) {
  const uint32_t tid = threadIdx.x, T = SR_THREADS;
  for (uint32_t r = 0; r < f; ++r) {
    Ext c0 = ezero(), c2 = ezero();
    for (uint32_t i = tid; i < nn / 2; i += T)
      pair_coeffs(ld_ext(F + 2 * i), ld_ext(F + 2 * i + 1), ld_ext(Wt + 2 * i), ld_ext(Wt + 2 * i + 1), c0, c2);
    block_sum2_to_wave0(c0, c2, s_wave);
    if (is_wave0()) {
      tx.write_e(c0);
      tx.write_e(c2);
      const Ext ch = tx.sample_ext();
      if (tid == 0) {
        s_x = ch;
        js->rs[rs_idx + r] = ch;
      }
    }
    __syncthreads();
    const Ext rr = s_x;
    for (uint32_t i = tid; i < nn / 2; i += T) {
      const Ext f0 = ld_ext(F + 2 * i), f1 = ld_ext(F + 2 * i + 1);
      const Ext w0 = ld_ext(Wt + 2 * i), w1 = ld_ext(Wt + 2 * i + 1);
      st_ext(Fo + i, eadd(f0, emul(rr, esub(f1, f0))));
      st_ext(Wo + i, eadd(w0, emul(rr, esub(w1, w0))));
    }
    __syncthreads();
    Ext* tf = F;
    F = Fo;
    Fo = tf;
    Ext* tw = Wt;
    Wt = Wo;
    Wo = tw;
    nn /= 2;
  }
}

__global__ void __launch_bounds__(SR_THREADS) small_round_kernel(JP<DevTx> t_, JP<uint32_t> log_, JP<JobScalars> js_,
                                                                  SmallRoundArgs a) {
  DevTx* t = t_;
  uint32_t* log = log_;
  JobScalars* js = js_;
  __shared__ Ext s_eq[SR_MAX_OOD][2][64];   // split eq tables of the OOD points
  __shared__ Ext s_pt[SR_MAX_LOG_N];
  __shared__ Ext s_wave[2 * SR_WAVES];
  __shared__ Ext s_x;                        // broadcast of a reduction or a draw
  __shared__ Ext s_apow[SR_MAX_OOD];
  __shared__ uint32_t s_v[MAX_STIR];
  __shared__ uint32_t s_scan[MAX_STIR];     // the sort's scan (blockDim.x), then var_q (count)
  __shared__ uint32_t s_idx[MAX_STIR];
  __shared__ uint32_t s_u[2];
  __shared__ Ext s_scale[MAX_STIR];          // alpha^(n_ood + q)
  __shared__ __align__(16) unsigned char s_pow[SR_POW_BYTES];
  const uint32_t tid = threadIdx.x;
  const uint32_t n = 1u << a.kc;
  const uint32_t m0 = a.kc / 2, m1 = a.kc - a.kc / 2;  // split of the round polynomial's index
  WaveTx tx;
  if (is_wave0()) {
    tx_begin(tx, t, log, js);
    write_root_body(tx, a.root);
  }

  // ---- OOD claims: z -> expand(kc, z); value <f, eq(point)> (whir.rs:224-235) ----
  for (uint32_t o = 0; o < a.n_ood; ++o) {
    if (is_wave0()) {
      const Ext z = tx.sample_ext();
      expand_point_wave(z, a.kc, s_pt);
    }
    __syncthreads();
    eq_small(s_pt, m0, s_eq[o][0]);
    eq_small(s_pt + m0, m1, s_eq[o][1]);
    Ext acc = ezero(), unused = ezero();
    for (uint32_t x = tid; x < n; x += SR_THREADS)
      acc = eadd(acc, emul(ld_ext(a.f_cur + x), emul(s_eq[o][0][x & ((1u << m0) - 1)], s_eq[o][1][x >> m0])));
    block_sum2_to_wave0(acc, unused, s_wave);
    if (is_wave0()) tx.write_e(acc);
    __syncthreads();  // s_wave and s_pt are reused
  }

  // ---- STIR queries: draws, sort, openings of the previous tree as hints (whir.rs:237-253) ----
  if (is_wave0()) tx.sample_bits_many(a.n_stir, a.kfd, s_v);
  const uint32_t count = sort_dedup_block(s_v, a.n_stir, s_scan, s_idx);
  const uint32_t per = a.row_words + 8 * a.kfd;
  if (is_wave0()) {
    if (tid == 0) s_u[0] = tx.log_words;
    tx.log_words += count * per;
    if (tx.log_words > tx.log_cap) tx.overflow = 1;
  }
  __syncthreads();
  {
    const uint32_t base = s_u[0];
    if ((uint64_t)base + (uint64_t)count * per <= t->log_cap)
      write_openings(log, base, s_idx, count, a.prev_cw, a.row_words, a.prev_tree, a.kfd);
  }

  // ---- fold prologue: alpha; claim scales (prover.rs:377-384) ----
  if (is_wave0()) {
    const Ext al = tx.sample_ext();
    if (tid == 0) {
      s_x = al;
      js->alpha = al;
    }
  }
  __syncthreads();
  const Ext alpha = s_x;
  if (tid == 0) {
    Ext p = eone();
    for (uint32_t o = 0; o < a.n_ood; ++o) {
      s_apow[o] = p;
      p = emul(p, alpha);
    }
  }
  // pow tables: powL[q][e] = var_q^e (e < 2^m0), powR[q][e] = alpha^(n_ood+q) var_q^(e 2^m0)
  uint32_t* powL = reinterpret_cast<uint32_t*>(s_pow);
  Ext* powR = reinterpret_cast<Ext*>(s_pow + (((size_t)count * (4u << m0) + 15) & ~(size_t)15));
  const uint32_t n0 = 1u << m0, n1 = 1u << m1;
  {
    const uint32_t om = two_adic_generator(a.kfd);
    for (uint32_t q = tid; q < count; q += SR_THREADS) {
      s_scan[q] = pow(om, s_idx[q]);  // var_q (whir.rs:242-246)
      s_scale[q] = epow(alpha, (uint64_t)a.n_ood + q);
    }
    __syncthreads();
    for (uint32_t it = tid; it < count * (n0 + n1); it += SR_THREADS) {
      const uint32_t q = it / (n0 + n1), e = it % (n0 + n1);
      const uint32_t v = s_scan[q];
      if (e < n0) {
        powL[q * n0 + e] = pow(v, e);
      } else {
        powR[q * n1 + (e - n0)] = emul_base(s_scale[q], pow(v, (uint64_t)(e - n0) << m0));
      }
    }
    __syncthreads();
  }

  // ---- combine (combine.rs:37-64): lazy sums of the pow terms as combine_kernel ----
  for (uint32_t x = tid; x < n; x += SR_THREADS) {
    const uint32_t lo = x & (n0 - 1), hi = x >> m0;
    uint64_t acc[4] = {0, 0, 0, 0};
    uint32_t q = 0;
    for (; q + 1 < count; q += 2) {
      const uint32_t l0 = powL[q * n0 + lo], l1 = powL[(q + 1) * n0 + lo];
      const Ext r0 = powR[q * n1 + hi], r1 = powR[(q + 1) * n1 + hi];
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] += redc(mad64(l0, r0.c[c], (uint64_t)l1 * r1.c[c]));
    }
    if (q < count) {
      const uint32_t l0 = powL[q * n0 + lo];
      const Ext r0 = powR[q * n1 + hi];
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] += redc((uint64_t)l0 * r0.c[c]);
    }
    Ext part;
#pragma unroll
    for (int c = 0; c < 4; ++c) part.c[c] = canon2(mulr(redc(acc[c]), R2MOD));
    for (uint32_t o = 0; o < a.n_ood; ++o)
      part = eadd(part, emul(s_apow[o], emul(s_eq[o][0][lo], s_eq[o][1][hi])));
    st_ext(a.w_cur + x, eadd(ld_ext(a.w_cur + x), part));
  }
  __syncthreads();

  // ---- f sumcheck rounds, each followed by the fold (expr.rs:52-70) ----
  sumcheck_block(tx, js, a.f_cur, a.w_cur, a.f_alt, a.w_alt, n, a.f, a.rs_idx, s_wave, s_x);
  if (is_wave0()) tx_end(tx, t, js);
}

// The f sumcheck rounds of a later WHIR round whose combined polynomial has
// n <= 2^SC_TAIL_MAX_LOG_N extension elements, in one workgroup: the same
// transcript and folded arrays as coeffs_kernel, then f times round_finalize
// (fold_coeffs_kernel between them) and fold_kernel, in one launch instead of
// 2f + 1 whose ~5 us floors are most of their time at this size.
__global__ void __launch_bounds__(SR_THREADS) sumcheck_tail_kernel(JP<DevTx> t_, JP<uint32_t> log_, JP<JobScalars> js_, JP<Ext> f_,
                                                                   JP<Ext> w_, JP<Ext> fo_, JP<Ext> wo_, uint32_t n, uint32_t nf,
                                                                   uint32_t rs_idx) {
  DevTx* t = t_;
  uint32_t* log = log_;
  JobScalars* js = js_;
  Ext* f = f_;
  Ext* w = w_;
  Ext* fo = fo_;
  Ext* wo = wo_;
  __shared__ Ext s_wave[2 * SR_WAVES];
  __shared__ Ext s_x;
  WaveTx tx;
  if (is_wave0()) tx_begin(tx, t, log, js);
  sumcheck_block(tx, js, f, w, fo, wo, n, nf, rs_idx, s_wave, s_x);
  if (is_wave0()) tx_end(tx, t, js);
}

// The final phase (whir.rs:314-355 and test.rs:233): write the final
// polynomial (n extension elements), draw the last STIR indexes, write their
// openings of the last tree as hints, draw the checkpoint.
__global__ void __launch_bounds__(SERIAL_THREADS) final_phase_kernel(JP<DevTx> t_, JP<uint32_t> log_, JP<JobScalars> js_,
                                                                     JP<const Ext> poly_, uint32_t n, uint32_t n_stir,
                                                                     uint32_t kfd, JP<const uint32_t> cw_,
                                                                     uint32_t row_words, JP<const uint32_t> tree_) {
  DevTx* t = t_;
  uint32_t* log = log_;
  JobScalars* js = js_;
  const Ext* poly = poly_;
  const uint32_t* cw = cw_;
  const uint32_t* tree = tree_;
  __shared__ uint32_t s_v[MAX_STIR];
  __shared__ uint32_t s_scan[SERIAL_THREADS];
  __shared__ uint32_t s_idx[MAX_STIR];
  __shared__ uint32_t s_u;
  const uint32_t tid = threadIdx.x;
  WaveTx tx;
  if (is_wave0()) {
    tx_begin(tx, t, log, js);
    const uint32_t* words = reinterpret_cast<const uint32_t*>(poly);
    const uint32_t lane = tid & 63;
    for (uint32_t i0 = 0; i0 < n; i0 += 16) {
      const uint32_t v = i0 * 4 + lane < n * 4 ? words[i0 * 4 + lane] : 0u;
      const uint32_t m = n - i0 < 16 ? n - i0 : 16u;
      for (uint32_t j = 0; j < m; ++j) {
        Ext e;
#pragma unroll
        for (int c = 0; c < 4; ++c) e.c[c] = (uint32_t)__builtin_amdgcn_readlane((int)v, (int)(4 * j + c));
        tx.write_e(e);
      }
    }
    tx.sample_bits_many(n_stir, kfd, s_v);
  }
  const uint32_t count = sort_dedup_block(s_v, n_stir, s_scan, s_idx);
  const uint32_t per = row_words + 8 * kfd;
  if (is_wave0()) {
    if (tid == 0) s_u = tx.log_words;
    tx.log_words += count * per;
    if (tx.log_words > tx.log_cap) tx.overflow = 1;
    // the test's checkpoint draw, on a copy: the stored transcript is the one
    // Whir::open hands back (whir.rs:359), before the caller's next draw
    WaveTx after = tx;
    const uint32_t v = after.sample();
    if (tid == 0) {
      js->checkpoint = v;
      // the proof's status after the log's capacity, so one copy brings back
      // the proof and its length, overflow flag and checkpoint
      log[tx.log_cap] = tx.log_words;
      log[tx.log_cap + 1] = tx.overflow;
      log[tx.log_cap + 2] = v;
    }
    tx_end(tx, t, js);
  }
  __syncthreads();
  const uint32_t base = s_u;
  if ((uint64_t)base + (uint64_t)count * per <= t->log_cap) write_openings(log, base, s_idx, count, cw, row_words, tree, kfd);
}

}  // namespace kb
