// Data-parallel sumcheck / eq-polynomial kernels (the HBM passes of
// crates/poly/src/eq.rs and crates/whir/src/sumcheck/{expr,combine,prover}.rs).
// Conventions (crates/poly/src/point.rs:110-126): eq(z)[x] = prod_j (x_j ? z_j : 1 - z_j)
// with bit j of x bound to z_j; an eq table over m variables is stored split as
// eq0 (low m0 bits) and eq1 (high bits): eq[x] = eq0[x & (2^m0-1)] * eq1[x >> m0].
#pragma once
#include "field.h"
#include "poseidon2.h"
#include "tx.h"

namespace kb {

constexpr int SC_THREADS = 256;
constexpr int MAX_CLASSES = 32;  // distinct poly sizes of a stack

// Static description of one claim of the SVO phase (an evaluation claim on a
// stacked poly, or a virtual claim on the whole stacked vector), fixed by the
// job shape and uploaded once. Claims drawn at one point form a group; its
// first claim's tables are built at the draw.
struct ClaimDesc {
  uint64_t eq0_off;   // pool offset (Ext) of eq over the low m0 rest variables (the group's)
  uint64_t eq1_off;   // pool offset of this claim's eq over the high m1 rest variables
  uint64_t lead_eq1;  // pool offset of the group's first eq1 table (== eq1_off for the first claim)
  uint64_t poly_off;  // offset of the claim's poly in the stacked base vector
  uint32_t m0, m1;    // rest variables k_poly - f = m0 + m1
  uint32_t rank;      // alpha exponent: position in layout order (prover.rs:155-174)
  uint32_t pad;
};

// Size classes of a stack in the folded (k - f variable) vector, for the
// expression build: class c covers [start[c], start[c+1]) in blocks of
// 2^log[c]; block b of class c is segment seg0[c] + b, whose claims are
// seg_claims[seg_start[s] .. seg_start[s+1]).
struct StackBlocks {
  uint32_t n_class;
  uint32_t log[MAX_CLASSES];
  uint32_t seg0[MAX_CLASSES];
  uint64_t start[MAX_CLASSES + 1];
  JP<const uint32_t> seg_start;
  JP<const uint32_t> seg_claims;
};


__device__ __forceinline__ Ext ld_ext(const Ext* p) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  return Ext{{v.x, v.y, v.z, v.w}};
}
__device__ __forceinline__ void st_ext(Ext* p, const Ext& e) {
  *reinterpret_cast<uint4*>(p) = make_uint4(e.c[0], e.c[1], e.c[2], e.c[3]);
}
// Sum over the 64 lanes of a wavefront, returned to every lane, with DPP
// moves feeding the adds (no LDS crossbar traffic): rotations inside each
// 16-lane row give every lane its row's sum; row_bcast:15 adds row 0's sum
// into row 1 and row 2's into row 3, row_bcast:31 adds lane 31 (rows 0 + 1)
// into rows 2 and 3, so lane 63 holds the total, which readlane broadcasts.
// Rows outside a move's row mask read the move's "old" operand, 0. Every
// lane of the wavefront must be active.
constexpr int DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;
__device__ __forceinline__ uint32_t wave_sum_word(uint32_t x) {
  x = add(x, dpp<DPP_ROR(8)>(x));
  x = add(x, dpp<DPP_ROR(4)>(x));
  x = add(x, dpp<DPP_ROR(2)>(x));
  x = add(x, dpp<DPP_ROR(1)>(x));
  x = add(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_BCAST15, 0xA, 0xF, false));
  x = add(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, DPP_ROW_BCAST31, 0xC, 0xF, false));
  return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}
__device__ __forceinline__ Ext wave_sum_ext(Ext v) {
#pragma unroll
  for (int c = 0; c < 4; ++c) v.c[c] = wave_sum_word(v.c[c]);
  return v;
}

// Sum of one Ext per thread over a block of NT threads, returned to every
// thread: a butterfly inside each wavefront, then the NT/64 wave sums from LDS
// (sh: NT/64 entries; one barrier). Two calls in a row need a barrier between.
template <int NT>
__device__ __forceinline__ Ext block_reduce_ext(Ext v, Ext* sh) {
  v = wave_sum_ext(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  Ext s = sh[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) s = eadd(s, sh[i]);
  return s;
}

// Split eq tables of two points at once, into global memory, whole block
// (>= 128 threads) cooperating: table u (u < 2, m_u <= 2 EQ_LDS_LOG variables,
// times scale_u) is eq(z_u)[x] = prod_j (x_j ? z_j : 1 - z_j). It is the
// product of four sub-tables of <= 4 variables each (variables [0, a0),
// [a0, a), [a, a + b0), [a + b0, m) with a = min(m, 8), a0 = min(a, 4),
// b0 = min(m - a, 4)), each entry a product of its factors; pairs of
// sub-tables are multiplied into the low (2^a) and high (2^(m - a)) tables in
// LDS, and out[x] = lo[x mod 2^a] * hi[x >> a]. Two barriers; the scratch may
// be reused after a further barrier.
constexpr int EQ_LDS_LOG = 8;
struct EqPairScratch {
  Ext part[2][4][16];
  Ext lo[2][1 << EQ_LDS_LOG];
  Ext hi[2][1 << EQ_LDS_LOG];
};

__device__ void build_eq_pair(Ext* out0, const Ext* z0, int m0, Ext scale0, Ext* out1, const Ext* z1, int m1,
                              Ext scale1, EqPairScratch& sc) {
  const uint32_t tid = threadIdx.x;
  const int mu[2] = {m0, m1};
  auto split = [](int m, int& a, int& a0, int& b0) {
    a = m < EQ_LDS_LOG ? m : EQ_LDS_LOG;
    a0 = a < 4 ? a : 4;
    b0 = m - a < 4 ? m - a : 4;
  };
  if (tid < 128) {
    const int u = (int)(tid >> 6), p = (int)((tid >> 4) & 3);
    const uint32_t e = tid & 15;
    const int m = mu[u];
    const Ext* z = u ? z1 : z0;
    int a, a0, b0;
    split(m, a, a0, b0);
    const int off = p == 0 ? 0 : p == 1 ? a0 : p == 2 ? a : a + b0;
    const int nv = p == 0 ? a0 : p == 1 ? a - a0 : p == 2 ? b0 : m - a - b0;
    Ext f[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      f[j] = eone();
      if (j < nv) {
        const Ext zj = z[off + j];
        f[j] = ((e >> j) & 1) ? zj : esub(eone(), zj);
      }
    }
    const Ext sc0 = p == 0 ? (u ? scale1 : scale0) : eone();
    sc.part[u][p][e] = emul(emul(emul(f[0], f[1]), sc0), emul(f[2], f[3]));
  }
  __syncthreads();
  for (uint32_t it = tid; it < 4u << EQ_LDS_LOG; it += blockDim.x) {
    const int u = (int)(it >> (EQ_LDS_LOG + 1)), h = (int)((it >> EQ_LDS_LOG) & 1);
    const uint32_t x = it & ((1u << EQ_LDS_LOG) - 1);
    int a, a0, b0;
    split(mu[u], a, a0, b0);
    const int nb = h ? mu[u] - a : a, n_lo = h ? b0 : a0;
    if (x >= (1u << nb)) continue;
    const Ext v = emul(sc.part[u][2 * h][x & ((1u << n_lo) - 1)], sc.part[u][2 * h + 1][x >> n_lo]);
    if (h) sc.hi[u][x] = v;
    else sc.lo[u][x] = v;
  }
  __syncthreads();
  const uint32_t n0 = 1u << m0, n1 = 1u << m1;
  int a_0, a_1, t0, t1;
  split(m0, a_0, t0, t1);
  split(m1, a_1, t0, t1);
  for (uint32_t it = tid; it < n0 + n1; it += blockDim.x) {
    const int u = it >= n0;
    const uint32_t x = u ? it - n0 : it;
    const int a = u ? a_1 : a_0;
    st_ext((u ? out1 : out0) + x, emul(sc.lo[u][x & ((1u << a) - 1)], sc.hi[u][x >> a]));
  }
}

// ---- claim evaluation on the base polynomial (SplitEq::compress_hi, eq.rs:290-398) ----
// out[i] = sum_h eq(rest)[h] * poly[h * W + i] for i < W = 2^F_LOG; the block for
// h1 = blockIdx.x sums over h0 < 2^m0 and writes W partial values * eq1[h1].
// blockIdx.y selects one of the claims sharing the point: claim c0 + y reads
// its poly at desc[c0 + y].poly_off (desc == nullptr: offset 0) and writes its
// partials after the previous claims' gridDim.x * W.
template <int F_LOG>
__global__ void __launch_bounds__(SC_THREADS) compress_hi_kernel(JP<const uint32_t> poly_,
                                                             JP<const Ext> eq0_, JP<const Ext> eq1_,
                                                             uint32_t m0, JP<Ext> partials_,
                                                             JP<const ClaimDesc> desc_, uint32_t c0) {
  const uint32_t* __restrict__ poly = poly_;
  const Ext* __restrict__ eq0 = eq0_;
  const Ext* __restrict__ eq1 = eq1_;
  Ext* __restrict__ partials = partials_;
  const ClaimDesc* __restrict__ desc = desc_;
  if (desc) poly += desc[c0 + blockIdx.y].poly_off;
  partials += (uint64_t)blockIdx.y * gridDim.x * (1u << F_LOG);
  constexpr int W = 1 << F_LOG;
  constexpr int CPT = W < 4 ? W : 4;        // columns per thread
  constexpr int TPR = W / CPT;              // threads per row
  constexpr int RPP = SC_THREADS / TPR;     // rows per pass
  __shared__ Ext sh[SC_THREADS * CPT];
  const uint32_t t = threadIdx.x;
  const uint32_t cg = t % TPR, rl = t / TPR;
  const uint64_t h1 = blockIdx.x;
  const uint64_t n0 = 1ull << m0;
  // Lazy sums: two Montgomery products per REDC (2 p^2 < p 2^32), each REDC
  // output (< 2p) added unreduced in 64 bits; one modular fix per output.
  uint64_t acc[CPT][4];
#pragma unroll
  for (int c = 0; c < CPT; ++c)
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[c][k] = 0;
  auto load_row = [&](uint64_t h0, uint32_t (&v)[CPT]) {
    const uint32_t* row = poly + ((h0 + (h1 << m0)) << F_LOG) + cg * CPT;
    if constexpr (CPT == 4) {
      const uint4 q = *reinterpret_cast<const uint4*>(row);
      v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
#pragma unroll
      for (int c = 0; c < CPT; ++c) v[c] = row[c];
    }
  };
  uint64_t h0 = rl;
#pragma unroll 2
  for (; h0 + RPP < n0; h0 += 2 * RPP) {
    const Ext e0 = ld_ext(eq0 + h0), e1 = ld_ext(eq0 + h0 + RPP);
    uint32_t v0[CPT], v1[CPT];
    load_row(h0, v0);
    load_row(h0 + RPP, v1);
#pragma unroll
    for (int c = 0; c < CPT; ++c)
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[c][k] += redc(mad64(e0.c[k], v0[c], (uint64_t)e1.c[k] * v1[c]));
  }
  if (h0 < n0) {
    const Ext e0 = ld_ext(eq0 + h0);
    uint32_t v0[CPT];
    load_row(h0, v0);
#pragma unroll
    for (int c = 0; c < CPT; ++c)
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[c][k] += redc((uint64_t)e0.c[k] * v0[c]);
  }
#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    Ext r;
#pragma unroll
    for (int k = 0; k < 4; ++k) r.c[k] = canon2(mulr(redc(acc[c][k]), R2MOD));  // acc < 2^8 * 2p
    sh[t * CPT + c] = r;
  }
  __syncthreads();
  if (t < (uint32_t)W) {
    const uint32_t g = t / CPT, c = t % CPT;
    Ext s = ezero();
    for (uint32_t r = 0; r < RPP; ++r) s = eadd(s, sh[(r * TPR + g) * CPT + c]);
    st_ext(partials + h1 * W + t, emul(s, ld_ext(eq1 + h1)));
  }
}

// ---- evaluation of an extension polynomial (SplitEq::eval_ext, eq.rs:179-250) ----
__global__ void __launch_bounds__(SC_THREADS) eval_ext_kernel(JP<const Ext> poly_, JP<const Ext> eq0_,
                                                          JP<const Ext> eq1_, uint32_t m0,
                                                          JP<Ext> partials_) {
  const Ext* __restrict__ poly = poly_;
  const Ext* __restrict__ eq0 = eq0_;
  const Ext* __restrict__ eq1 = eq1_;
  Ext* __restrict__ partials = partials_;
  __shared__ Ext sh[SC_THREADS];
  const uint64_t h1 = blockIdx.x;
  const uint64_t n0 = 1ull << m0;
  Ext acc = ezero();
  for (uint64_t h0 = threadIdx.x; h0 < n0; h0 += SC_THREADS)
    acc = eadd(acc, emul(ld_ext(poly + h0 + (h1 << m0)), ld_ext(eq0 + h0)));
  const Ext s = block_reduce_ext<SC_THREADS>(acc, sh);
  if (threadIdx.x == 0) st_ext(partials + h1, emul(s, ld_ext(eq1 + h1)));
}

// ---- expression build after the SVO rounds (ProverLayout::expression, prover.rs:184-229) ----
// epoly[j] = sum_{i<W} eq_rs[i] * poly[j*W + i]            (compress_lo, eq.rs:491-533)
// ew[j]    = sum_c eq0_c[j & (2^m0-1)] * eq1s_c[j >> m0]    (combine_into, eq.rs:541-621;
//            eq1s_c already carries alpha^c * eq(svo_c, rs))
// Claims [full0, full0 + n_full) span the whole vector (their tables split at
// m0); the others are on one stacked poly's block (prover.rs:190-199), found
// through `sb` (n_class == 0: none).
template <int F_LOG>
__global__ void __launch_bounds__(SC_THREADS) expr_build_kernel(JP<const uint32_t> poly_,
                                                            JP<const Ext> eq_rs_, JP<const Ext> pool_,
                                                            JP<const ClaimDesc> desc_, uint32_t full0,
                                                            uint32_t n_full, uint32_t m0, uint64_t n,
                                                            JP<Ext> epoly_, JP<Ext> ew_,
                                                            StackBlocks sb, JP<const Ext> cscale_) {
  const uint32_t* __restrict__ poly = poly_;
  const Ext* __restrict__ eq_rs = eq_rs_;
  const Ext* __restrict__ pool = pool_;
  const ClaimDesc* __restrict__ desc = desc_;
  Ext* __restrict__ epoly = epoly_;
  Ext* __restrict__ ew = ew_;
  const Ext* __restrict__ cscale = cscale_;
  constexpr int W = 1 << F_LOG;
  __shared__ Ext s_eq[W];
  if (threadIdx.x < W) s_eq[threadIdx.x] = eq_rs[threadIdx.x];
  __syncthreads();
  const uint64_t j = (uint64_t)blockIdx.x * SC_THREADS + threadIdx.x;
  if (j >= n) return;
  const uint32_t* row = poly + (j << F_LOG);
  // lazy: two products per REDC, REDC outputs summed in 64 bits, one fix per coefficient
  uint64_t a4[4] = {0, 0, 0, 0};
  if constexpr (W >= 4) {
#pragma unroll
    for (int i = 0; i < W; i += 4) {
      const uint4 q = *reinterpret_cast<const uint4*>(row + i);
#pragma unroll
      for (int k = 0; k < 4; ++k)
        a4[k] += (uint64_t)redc(mad64(s_eq[i].c[k], q.x, (uint64_t)s_eq[i + 1].c[k] * q.y)) +
                 redc(mad64(s_eq[i + 2].c[k], q.z, (uint64_t)s_eq[i + 3].c[k] * q.w));
    }
  } else {
#pragma unroll
    for (int i = 0; i < W; ++i)
#pragma unroll
      for (int k = 0; k < 4; ++k) a4[k] += redc((uint64_t)s_eq[i].c[k] * row[i]);
  }
  Ext acc;
#pragma unroll
  for (int k = 0; k < 4; ++k) acc.c[k] = canon2(mulr(redc(a4[k]), R2MOD));
  st_ext(epoly + j, acc);
  const uint64_t lo = j & ((1ull << m0) - 1), hi = j >> m0;
  Ext w = ezero();
  if (m0 >= 6) {
    // a wavefront's 64 consecutive j share hi: eq1 entries as scalar loads
    const uint64_t hu = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)hi);
    for (uint32_t c = full0; c < full0 + n_full; ++c)
      w = eadd(w, emul(ld_ext(pool + desc[c].eq0_off + lo), emul(ld_ext(pool + desc[c].lead_eq1 + hu), cscale[c])));
  } else {
    for (uint32_t c = full0; c < full0 + n_full; ++c)
      w = eadd(w, emul(ld_ext(pool + desc[c].eq0_off + lo), emul(ld_ext(pool + desc[c].lead_eq1 + hi), cscale[c])));
  }
  if (sb.n_class) {
    uint32_t cl = 0;
    while (cl < sb.n_class && j >= sb.start[cl + 1]) ++cl;
    if (cl < sb.n_class) {
      const uint64_t rel = j - sb.start[cl];
      const uint32_t seg = sb.seg0[cl] + (uint32_t)(rel >> sb.log[cl]);
      const uint64_t x = rel & ((1ull << sb.log[cl]) - 1);
      for (uint32_t q = sb.seg_start[seg]; q < sb.seg_start[seg + 1]; ++q) {
        const uint32_t c = sb.seg_claims[q];
        const ClaimDesc& d = desc[c];
        w = eadd(w, emul(ld_ext(pool + d.eq0_off + (x & ((1ull << d.m0) - 1))),
                         emul(ld_ext(pool + d.lead_eq1 + (x >> d.m0)), cscale[c])));
      }
    }
  }
  st_ext(ew + j, w);
}

// ---- claim combination (combine_claims, combine.rs:37-64,178-203,407-468) ----
// w[x] += sum_c eq0_c[x_lo] * eq1s_c[x_hi]  +  sum_j powR[j][x_hi] * powL[j][x_lo]
// eq1s_c carries alpha^c; powL[j][e] = var_j^e (e < 2^m0), powR[j][e] = scale_j * var_j^(e*2^m0).
// A workgroup takes 256 / S consecutive outputs; its S wavefront groups split
// the terms (term j to group j mod S) and meet in LDS. Each wavefront covers 64
// consecutive outputs with one term per step, so for m0 >= 6 (UNI) the term's
// eq1 / powR entry is the same for the whole wavefront and is read as a scalar
// load; only the eq0 / powL row is a (coalesced) vector load. The pow terms are
// Montgomery products taken two per REDC (2 p^2 < p 2^32) and summed unreduced
// in 64 bits, so a term costs ~5 VALU instructions per coefficient.
template <int S, bool UNI>
__global__ void __launch_bounds__(SC_THREADS) combine_kernel(JP<Ext> w_, uint64_t n, uint32_t m0,
                                                         JP<const Ext> eq_pool_, uint32_t n_eq,
                                                         uint64_t eq_stride, JP<const uint32_t> powL_,
                                                         JP<const Ext> powR_, JP<const uint32_t> n_pow_ptr_,
                                                         uint32_t m1) {
  Ext* __restrict__ w = w_;
  const Ext* __restrict__ eq_pool = eq_pool_;
  const uint32_t* __restrict__ powL = powL_;
  const Ext* __restrict__ powR = powR_;
  const uint32_t* n_pow_ptr = n_pow_ptr_;
  constexpr uint32_t XB = SC_THREADS / S;  // outputs per workgroup, a multiple of 64
  __shared__ Ext s_part[S > 1 ? SC_THREADS : 1];
  const uint32_t sub = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / XB));
  const uint32_t xl = threadIdx.x % XB;
  const uint64_t x = (uint64_t)blockIdx.x * XB + xl;
  const bool valid = x < n;  // every lane takes part in the LDS reduction below
  const uint64_t xx = valid ? x : 0;
  const uint32_t n_pow = n_pow_ptr ? *n_pow_ptr : 0u;
  const uint64_t n0 = 1ull << m0, n1 = 1ull << m1;
  const uint64_t lo = xx & (n0 - 1);
  uint64_t hi = xx >> m0;
  if constexpr (UNI) hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)hi);
  uint64_t a[4] = {0, 0, 0, 0};
  uint32_t j = sub;
#pragma unroll 2
  for (; j + S < n_pow; j += 2 * S) {
    const uint32_t l0 = powL[(uint64_t)j * n0 + lo], l1 = powL[(uint64_t)(j + S) * n0 + lo];
    const Ext r0 = ld_ext(powR + (uint64_t)j * n1 + hi), r1 = ld_ext(powR + (uint64_t)(j + S) * n1 + hi);
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] += redc(mad64(l0, r0.c[c], (uint64_t)l1 * r1.c[c]));
  }
  if (j < n_pow) {
    const uint32_t l0 = powL[(uint64_t)j * n0 + lo];
    const Ext r0 = ld_ext(powR + (uint64_t)j * n1 + hi);
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] += redc((uint64_t)l0 * r0.c[c]);
  }
  Ext part;
#pragma unroll
  for (int c = 0; c < 4; ++c) part.c[c] = canon2(mulr(redc(a[c]), R2MOD));  // a mod p
  for (uint32_t c = sub; c < n_eq; c += S) {
    const Ext* base = eq_pool + c * eq_stride;
    part = eadd(part, emul(ld_ext(base + lo), ld_ext(base + n0 + hi)));
  }
  if constexpr (S > 1) {
    s_part[threadIdx.x] = part;
    __syncthreads();
    if (sub != 0) return;
#pragma unroll
    for (int k = 1; k < S; ++k) part = eadd(part, s_part[xl + k * XB]);
  }
  if (valid) st_ext(w + x, eadd(ld_ext(w + x), part));
}

// Host: launch combine_kernel for 2^(m0 + m1) = n outputs; S = 4 when there
// are pow terms (n_pow_ptr set) or more than four eq terms, one lane per output
// otherwise.
inline void launch_combine(hipStream_t s, uint32_t jobs, Ext* w, uint64_t n, uint32_t m0, const Ext* eq_pool, uint32_t n_eq,
                           uint64_t eq_stride, const uint32_t* powL, const Ext* powR, const uint32_t* n_pow_ptr,
                           uint32_t m1) {
  const bool uni = m0 >= 6;
  const unsigned S = n_pow_ptr || n_eq > 4 ? 4 : 1;
  const unsigned grid = (unsigned)((n + SC_THREADS / S - 1) / (SC_THREADS / S));
  if (S == 4) {
    if (uni)
      hipLaunchKernelGGL((combine_kernel<4, true>), dim3(grid, 1, jobs), dim3(SC_THREADS), 0, s, w, n, m0, eq_pool, n_eq,
                         eq_stride, powL, powR, n_pow_ptr, m1);
    else
      hipLaunchKernelGGL((combine_kernel<4, false>), dim3(grid, 1, jobs), dim3(SC_THREADS), 0, s, w, n, m0, eq_pool, n_eq,
                         eq_stride, powL, powR, n_pow_ptr, m1);
  } else {
    if (uni)
      hipLaunchKernelGGL((combine_kernel<1, true>), dim3(grid, 1, jobs), dim3(SC_THREADS), 0, s, w, n, m0, eq_pool, n_eq,
                         eq_stride, powL, powR, n_pow_ptr, m1);
    else
      hipLaunchKernelGGL((combine_kernel<1, false>), dim3(grid, 1, jobs), dim3(SC_THREADS), 0, s, w, n, m0, eq_pool, n_eq,
                         eq_stride, powL, powR, n_pow_ptr, m1);
  }
}

// powL[j][e] = var_j^e for e < 2^m0; powR[j][e] = scale_j * var_j^(e * 2^m0) for e < 2^m1.
__global__ void __launch_bounds__(SC_THREADS) pow_tables_kernel(JP<const uint32_t> vars_,
                                                            JP<const Ext> scales_, JP<const uint32_t> count_ptr_,
                                                            uint32_t m0, uint32_t m1, JP<uint32_t> powL_,
                                                            JP<Ext> powR_) {
  const uint32_t* __restrict__ vars = vars_;
  const Ext* __restrict__ scales = scales_;
  const uint32_t* count_ptr = count_ptr_;
  uint32_t* __restrict__ powL = powL_;
  Ext* __restrict__ powR = powR_;
  const uint32_t count = *count_ptr;
  const uint64_t n0 = 1ull << m0, n1 = 1ull << m1;
  const uint64_t per = n0 + n1;
  const uint64_t id = (uint64_t)blockIdx.x * SC_THREADS + threadIdx.x;
  const uint64_t j = id / per, e = id % per;
  if (j >= count) return;
  const uint32_t v = vars[j];
  if (e < n0) {
    powL[j * n0 + e] = pow(v, e);
  } else {
    const uint64_t e1 = e - n0;
    st_ext(powR + j * n1 + e1, emul_base(scales[j], pow(v, e1 << m0)));
  }
}

// ---- sumcheck round coefficients (expr.rs:14-27) and fold (poly.rs:131-142) ----
// c0 = sum w[2i] f[2i];  c2 = sum (2w[2i+1]-w[2i]) (2f[2i+1]-f[2i]); one (c0,c2)
// partial per block.
__device__ __forceinline__ void pair_coeffs(const Ext& f0, const Ext& f1, const Ext& w0, const Ext& w1, Ext& c0,
                                            Ext& c2) {
  c0 = eadd(c0, emul(w0, f0));
  c2 = eadd(c2, emul(esub(edbl(w1), w0), esub(edbl(f1), f0)));
}

// Block blk of nblk (blockDim.x = SC_THREADS); sh: SC_THREADS entries.
__device__ __forceinline__ void coeffs_body(const Ext* __restrict__ f, const Ext* __restrict__ w, uint64_t n,
                                            Ext* __restrict__ partials, uint32_t blk, uint32_t nblk, Ext* sh) {
  Ext c0 = ezero(), c2 = ezero();
  const uint64_t npairs = n / 2;
  for (uint64_t i = (uint64_t)blk * SC_THREADS + threadIdx.x; i < npairs; i += (uint64_t)nblk * SC_THREADS)
    pair_coeffs(ld_ext(f + 2 * i), ld_ext(f + 2 * i + 1), ld_ext(w + 2 * i), ld_ext(w + 2 * i + 1), c0, c2);
  const Ext s0 = block_reduce_ext<SC_THREADS>(c0, sh);
  __syncthreads();
  const Ext s2 = block_reduce_ext<SC_THREADS>(c2, sh);
  if (threadIdx.x == 0) {
    st_ext(partials + 2 * blk, s0);
    st_ext(partials + 2 * blk + 1, s2);
  }
}

__global__ void __launch_bounds__(SC_THREADS) coeffs_kernel(JP<const Ext> f_, JP<const Ext> w_,
                                                        uint64_t n, JP<Ext> partials_) {
  const Ext* __restrict__ f = f_;
  const Ext* __restrict__ w = w_;
  Ext* __restrict__ partials = partials_;
  __shared__ Ext sh[SC_THREADS];
  coeffs_body(f, w, n, partials, blockIdx.x, gridDim.x, sh);
}

// Fold both arrays by r (a[i] = a[2i] + r (a[2i+1] - a[2i])) and emit the next
// round's (c0, c2) partials computed on the folded values.
// Block blk of nblk (blockDim.x = SC_THREADS); sh: SC_THREADS entries.
__device__ __forceinline__ void fold_coeffs_body(const Ext* __restrict__ f, const Ext* __restrict__ w, uint64_t n,
                                                 const Ext r, Ext* __restrict__ fo, Ext* __restrict__ wo,
                                                 Ext* __restrict__ partials, uint32_t blk, uint32_t nblk, Ext* sh) {
  Ext c0 = ezero(), c2 = ezero();
  const uint64_t nq = n / 4;
  for (uint64_t i = (uint64_t)blk * SC_THREADS + threadIdx.x; i < nq; i += (uint64_t)nblk * SC_THREADS) {
    const Ext a0 = ld_ext(f + 4 * i), a1 = ld_ext(f + 4 * i + 1), a2 = ld_ext(f + 4 * i + 2), a3 = ld_ext(f + 4 * i + 3);
    const Ext b0 = ld_ext(w + 4 * i), b1 = ld_ext(w + 4 * i + 1), b2 = ld_ext(w + 4 * i + 2), b3 = ld_ext(w + 4 * i + 3);
    const Ext F0 = eadd(a0, emul(r, esub(a1, a0)));
    const Ext F1 = eadd(a2, emul(r, esub(a3, a2)));
    const Ext W0 = eadd(b0, emul(r, esub(b1, b0)));
    const Ext W1 = eadd(b2, emul(r, esub(b3, b2)));
    st_ext(fo + 2 * i, F0);
    st_ext(fo + 2 * i + 1, F1);
    st_ext(wo + 2 * i, W0);
    st_ext(wo + 2 * i + 1, W1);
    pair_coeffs(F0, F1, W0, W1, c0, c2);
  }
  const Ext s0 = block_reduce_ext<SC_THREADS>(c0, sh);
  __syncthreads();
  const Ext s2 = block_reduce_ext<SC_THREADS>(c2, sh);
  if (threadIdx.x == 0) {
    st_ext(partials + 2 * blk, s0);
    st_ext(partials + 2 * blk + 1, s2);
  }
}

__global__ void __launch_bounds__(SC_THREADS) fold_coeffs_kernel(JP<const Ext> f_, JP<const Ext> w_,
                                                             uint64_t n, JP<const Ext> r_ptr_,
                                                             JP<Ext> fo_, JP<Ext> wo_,
                                                             JP<Ext> partials_) {
  const Ext* __restrict__ f = f_;
  const Ext* __restrict__ w = w_;
  const Ext* __restrict__ r_ptr = r_ptr_;
  Ext* __restrict__ fo = fo_;
  Ext* __restrict__ wo = wo_;
  Ext* __restrict__ partials = partials_;
  __shared__ Ext sh[SC_THREADS];
  fold_coeffs_body(f, w, n, ld_ext(r_ptr), fo, wo, partials, blockIdx.x, gridDim.x, sh);
}

__device__ __forceinline__ void fold_body(const Ext* __restrict__ f, const Ext* __restrict__ w, uint64_t n,
                                          const Ext r, Ext* __restrict__ fo, Ext* __restrict__ wo, uint32_t blk) {
  const uint64_t i = (uint64_t)blk * SC_THREADS + threadIdx.x;
  if (i >= n / 2) return;
  const Ext a0 = ld_ext(f + 2 * i), a1 = ld_ext(f + 2 * i + 1);
  const Ext b0 = ld_ext(w + 2 * i), b1 = ld_ext(w + 2 * i + 1);
  st_ext(fo + i, eadd(a0, emul(r, esub(a1, a0))));
  st_ext(wo + i, eadd(b0, emul(r, esub(b1, b0))));
}

__global__ void __launch_bounds__(SC_THREADS) fold_kernel(JP<const Ext> f_, JP<const Ext> w_,
                                                      uint64_t n, JP<const Ext> r_ptr_, JP<Ext> fo_,
                                                      JP<Ext> wo_) {
  const Ext* __restrict__ f = f_;
  const Ext* __restrict__ w = w_;
  const Ext* __restrict__ r_ptr = r_ptr_;
  Ext* __restrict__ fo = fo_;
  Ext* __restrict__ wo = wo_;
  fold_body(f, w, n, ld_ext(r_ptr), fo, wo, blockIdx.x);
}

}  // namespace kb
