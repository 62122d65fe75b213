// Column-batched DFT over KoalaBear (Reed–Solomon encoding), replacing Plonky3
// Radix2DFTSmallBatch::dft_batch / dft_algebra_batch as called at
// crates/whir/src/whir.rs:34,59:
//   out[i][c] = sum_{j < n_in} in[j][c] * w_N^(i*j),  w_N = two_adic_generator(log2 N)
// for the N-row, C-column row-major result of zero-padding the n_in-row input
// (the rate padding, whir.rs:32,57). An extension matrix is the same transform
// on its 4*C base columns.
//
// Coset form. With n = 2^ceil(log2 n_in) and E = N / n, output row i = t E + e
// (t < n, e < E) is
//   out[t E + e][c] = sum_{j < n} (in[j][c] * w_N^(e j)) * w_n^(t j),
// i.e. a size-n DFT of the "twisted" n x (E C) matrix X[j][(e, c)] =
// in[j][c] w_N^(e j), and the n x (E C) result is exactly the N x C output in
// natural row-major order. The zero rows are never touched: a rate-1/2^r LDE
// costs log2(n) butterfly stages, not log2(N), and the late WHIR rounds (a
// 16-row input spread over 2^15 rows) become one short pass.
//
// The size-n DFT is a four-step decomposition n = L_1 ... L_P (L_p <= 256,
// P <= 3). Pass p works on blocks of M_{p-1} = L_p M_p rows: for block b and
// every word offset w = m B + col (m < M_p, B = E C columns) it transforms the
// L_p elements at rows j M_p + m (j < L_p) and multiplies output t by
// w_{M_{p-1}}^(m t) (not in the final pass). Outputs end up in digit-reversed
// block order; the final pass of a multi-pass transform stores row
// t_1 + L_1 t_2 (+ L_1 L_2 t_3) directly.
//
// Per workgroup: QV "lines" (V consecutive columns of one (b, m)) x L rows.
// L = R1 * R2: stage A, each thread loads R1 rows of its line straight from
// HBM (consecutive lanes -> consecutive 16 B), runs an R1-point DFT in
// registers and the inner twiddles w_L^(j_b t_a); one LDS exchange; stage B
// runs R2-point DFTs in registers and stores straight to HBM. All index math is
// 32-bit (a transform holds < 2^32 words).
#pragma once
#include "field.h"

namespace kb {

// ---- compile-time Montgomery constants for the register DFTs ----
constexpr uint32_t cx_redc(uint64_t x) {
  const uint32_t m = (uint32_t)x * NPINV;
  const uint32_t t = (uint32_t)((x + (uint64_t)m * P) >> 32);
  return t >= P ? t - P : t;
}
constexpr uint32_t cx_mul(uint32_t a, uint32_t b) { return cx_redc((uint64_t)a * b); }
constexpr uint32_t cx_pow(uint32_t a, uint64_t e) {
  uint32_t r = ONE;
  while (e) {
    if (e & 1) r = cx_mul(r, a);
    a = cx_mul(a, a);
    e >>= 1;
  }
  return r;
}
struct Omega32 {
  uint32_t w[32];
};
constexpr Omega32 make_omega32() {
  Omega32 t{};
  const uint32_t g = (uint32_t)(((uint64_t)GENERATOR << 32) % P);  // the generator in Montgomery form
  const uint32_t w32 = cx_pow(g, (uint64_t)(P - 1) >> 5);          // two_adic_generator(5)
  uint32_t x = ONE;
  for (int i = 0; i < 32; ++i) {
    t.w[i] = x;
    x = cx_mul(x, w32);
  }
  return t;
}
constexpr Omega32 kOmega32 = make_omega32();

template <int V>
struct Vec {
  uint32_t v[V];
};

// The transform's arrays are device memory. Their pointers reach the kernels
// through JP (job-tagged addresses), which hides the address space from the
// compiler; on the device the loads and stores name the global space, so they
// are global_load / global_store (in-order vmcnt waits) rather than flat ones
// (which also count on lgkmcnt and are waited for as a whole).
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gptr(const uint32_t* p) {
  return (const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gptr(uint32_t* p) {
  return (__attribute__((address_space(1))) T*)p;
}
#else
template <class T>
KB_HD const T* gptr(const uint32_t* p) {
  return reinterpret_cast<const T*>(p);
}
template <class T>
KB_HD T* gptr(uint32_t* p) {
  return reinterpret_cast<T*>(p);
}
#endif
template <int V>
KB_HD Vec<V> vload(const uint32_t* p) {
  Vec<V> r;
  if constexpr (V == 4) {
    const uint4 t = *gptr<uint4>(p);
    r.v[0] = t.x;
    r.v[1] = t.y;
    r.v[2] = t.z;
    r.v[3] = t.w;
  } else if constexpr (V == 2) {
    const uint2 t = *gptr<uint2>(p);
    r.v[0] = t.x;
    r.v[1] = t.y;
  } else {
    r.v[0] = *gptr<uint32_t>(p);
  }
  return r;
}
template <int V>
KB_HD void vstore(uint32_t* p, const Vec<V>& x) {
  if constexpr (V == 4) *gptr<uint4>(p) = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
  else if constexpr (V == 2) *gptr<uint2>(p) = make_uint2(x.v[0], x.v[1]);
  else *gptr<uint32_t>(p) = x.v[0];
}
template <int V>
KB_HD void vscale(Vec<V>& x, uint32_t w) {
#pragma unroll
  for (int c = 0; c < V; ++c) x.v[c] = mul(x.v[c], w);
}

constexpr int bitrev(int i, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b) r |= ((i >> b) & 1) << (bits - 1 - b);
  return r;
}

// Register slot i after stage s of dft_regs holds a value that stage s + 1
// only multiplies by a twiddle (the upper half of a butterfly with j != 0);
// after the last stage, bit i of the caller's LAZY_OUT says so.
constexpr bool dft_lazy(int LR, int s, int i, uint32_t lazy_out) {
  if (s + 1 == LR) return ((lazy_out >> i) & 1u) != 0;
  const int half = 1 << (s + 1);
  return (i % (2 * half)) >= half && (i % half) != 0;
}

// x <- DFT_R(x) (natural order in and out, R = 2^LR <= 32): radix-2 DIT on the
// bit-reversed registers; twiddles w_{2^(s+1)}^j are compile-time constants and
// j = 0 multiplies are skipped (17 multiplies for R = 16, 49 for 32). Inputs are in
// [0, p). A sum or difference that is next only multiplied by a twiddle is
// left unreduced in [0, 2p) (u + t, or u - t + p): the Montgomery product of
// a factor < 2p and a twiddle < p is below p 2^32 and comes out in [0, p), so
// such values skip the reduction of the add or sub (2 or 1 instructions).
// LAZY_OUT marks the outputs the caller multiplies next; the others are in
// [0, p).
template <int LR, int V, uint32_t LAZY_OUT = 0>
KB_HD void dft_regs(Vec<V> (&x)[1 << LR]) {
  constexpr int R = 1 << LR;
  static_assert(LR >= 0 && LR <= 5, "register DFT of up to 32 points");
  Vec<V> y[R];
#pragma unroll
  for (int i = 0; i < R; ++i) y[i] = x[bitrev(i, LR)];
#pragma unroll
  for (int s = 0; s < LR; ++s) {
    const int half = 1 << s;
#pragma unroll
    for (int g = 0; g < R; g += 2 * half) {
#pragma unroll
      for (int j = 0; j < half; ++j) {
        const int i0 = g + j, i1 = g + j + half;
        const uint32_t w = kOmega32.w[j << (4 - s)];  // w_{2^(s+1)}^j
        const bool lazy0 = dft_lazy(LR, s, i0, LAZY_OUT), lazy1 = dft_lazy(LR, s, i1, LAZY_OUT);
#pragma unroll
        for (int c = 0; c < V; ++c) {
          const uint32_t u = y[i0].v[c];
          const uint32_t t = j == 0 ? y[i1].v[c] : mul(y[i1].v[c], w);
          y[i0].v[c] = lazy0 ? u + t : add(u, t);
          y[i1].v[c] = lazy1 ? u - t + P : sub(u, t);
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < R; ++i) x[i] = y[i];
}
template <int V>
KB_HD void vcanon(Vec<V>& x) {  // [0, 2p) -> [0, p)
#pragma unroll
  for (int c = 0; c < V; ++c) {
    const uint32_t u = x.v[c] - P;
    x.v[c] = x.v[c] < u ? x.v[c] : u;
  }
}
// all outputs but the first (index 0 is never multiplied by a twiddle)
template <int LR>
constexpr uint32_t lazy_but_first() {
  return (uint32_t)((1ull << (1 << LR)) - 2u);
}

struct NttArgs {
  JP<const uint32_t> src;    // first pass: the n_in x C input; later: the n x B work array
  JP<uint32_t> dst;          // the n x B work array, or the output (final pass)
  JP<const uint32_t> tw_lo;  // w24^i, i < 4096
  JP<const uint32_t> tw_hi;  // w24^(4096 i), i < 4096
  uint32_t n_in;          // first pass: input rows (rows >= n_in read as zero)
  uint32_t log_c;         // first pass: input columns C (>= V)
  uint32_t log_nn;        // first pass: log2 N (order of the coset twist)
  uint32_t log_b;         // B = E C columns of the work array
  uint32_t log_n;         // n rows of the work array
  uint32_t log_m;         // M_p (this pass transforms L rows spaced M_p apart)
  uint32_t log_qv;        // lines per workgroup
  uint32_t first;         // read the input matrix and apply the coset twist
  uint32_t twiddle_out;   // multiply output t by w_{M_{p-1}}^(m t) (not the final pass)
  uint32_t natural;       // final pass of P >= 2: store rows in natural order
  uint32_t log_l1, log_l2;  // L_1, L_2 (natural-order store)
  uint32_t n_pass;
};

KB_HD uint32_t tw24(const uint32_t* lo, const uint32_t* hi, uint32_t e) {
  return mul(*gptr<uint32_t>(lo + (e & 4095u)), *gptr<uint32_t>(hi + ((e >> 12) & 4095u)));
}

// The first pass's R input rows r = r0 + j 2^log_step (j < R) of its line, in
// coset e and input column c (col = e C + c), times the coset twist w_N^(e r);
// rows >= n_in read as zero. Every row's load is issued before any is used:
// a load under a per-row branch, followed by the row's twist multiply, waited
// for each row in turn (16 round trips to HBM per thread in the base LDE's
// radix-16 first pass). Rows past the input load row 0 and are zeroed after.
template <int R, int V>
KB_HD void load_input_rows(const NttArgs& a, uint32_t r0, uint32_t log_step, uint32_t col,
                           Vec<V> (&x)[R]) {
  const uint32_t e = col >> a.log_c, c = col & ((1u << a.log_c) - 1);
  if (a.n_in == 0) {
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
      for (int q = 0; q < V; ++q) x[j].v[q] = 0u;
    return;
  }
  // the twist's table entries first, so that they are in flight with the rows
  const uint32_t sh = 24 - a.log_nn, i0 = (e * r0) << sh, i1 = (e << log_step) << sh;
  uint32_t t0[2] = {ONE, ONE}, t1[2] = {ONE, ONE};
  if (e) {
    const uint32_t* lo = a.tw_lo;
    const uint32_t* hi = a.tw_hi;
    t0[0] = *gptr<uint32_t>(lo + (i0 & 4095u));
    t0[1] = *gptr<uint32_t>(hi + ((i0 >> 12) & 4095u));
    t1[0] = *gptr<uint32_t>(lo + (i1 & 4095u));
    t1[1] = *gptr<uint32_t>(hi + ((i1 >> 12) & 4095u));
  }
  const uint32_t* in = a.src + c;
  if (r0 + ((uint32_t)(R - 1) << log_step) < a.n_in) {  // every row inside the input
#pragma unroll
    for (int j = 0; j < R; ++j) x[j] = vload<V>(in + ((r0 + ((uint32_t)j << log_step)) << a.log_c));
  } else {
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const uint32_t r = r0 + ((uint32_t)j << log_step);
      x[j] = vload<V>(in + ((r < a.n_in ? r : 0u) << a.log_c));
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const bool past = r0 + ((uint32_t)j << log_step) >= a.n_in;
#pragma unroll
      for (int q = 0; q < V; ++q) x[j].v[q] = past ? 0u : x[j].v[q];
    }
  }
  if (e) {
    uint32_t tw = mul(t0[0], t0[1]);  // tw24 of i0 and i1
    const uint32_t st = mul(t1[0], t1[1]);
#pragma unroll
    for (int j = 0; j < R; ++j) {
      vscale<V>(x[j], tw);
      tw = mul(tw, st);
    }
  }
}

// Workgroup `blk` of one pass (ntt_kernel runs it as blockIdx.x; the
// work-queue kernels of jobtable.h as the block of a dequeued item). `lds`
// holds the pass's tile; the pass has one workgroup barrier (none for a
// register-only pass).
template <int LR1, int LR2, int V>
__device__ __forceinline__ void ntt_block(const NttArgs& a, uint32_t blk, uint32_t* lds) {
  constexpr int R1 = 1 << LR1, R2 = 1 << LR2, LL = LR1 + LR2;
  constexpr int LV = V == 4 ? 2 : V == 2 ? 1 : 0;
  const uint32_t QV = 1u << a.log_qv;
  const uint32_t tid = threadIdx.x;
  const uint32_t cq = tid & (QV - 1);
  const uint32_t jb = tid >> a.log_qv;  // stage A: j_b; stage B: first t_a
  const uint32_t log_lpb = a.log_m + a.log_b - LV;  // lines per block
  const uint32_t line = (blk << a.log_qv) + cq;
  const uint32_t b = line >> log_lpb;
  const uint32_t w = (line & ((1u << log_lpb) - 1)) << LV;  // m B + col
  const uint32_t m = w >> a.log_b;
  const uint32_t col = w & ((1u << a.log_b) - 1);
  const uint32_t log_mprev = a.log_m + LL;
  const uint32_t jstride = 1u << (a.log_m + a.log_b);
  const uint32_t base = (b << (log_mprev + a.log_b)) + w;

  // ---- stage A: rows j = j_a R2 + j_b ----
  Vec<V> x[R1];
  if (!a.first) {
#pragma unroll
    for (int ja = 0; ja < R1; ++ja) x[ja] = vload<V>(a.src + base + ((uint32_t)ja * R2 + jb) * jstride);
  } else {
    // input row r = (j_a R2 + j_b) M_1 + m (one block)
    load_input_rows<R1, V>(a, (jb << a.log_m) + m, LR2 + a.log_m, col, x);
  }
  if constexpr (LR2 == 0) {
    // register-only pass (an early pass of a multi-pass transform): output t
    // of this line is x[t]; no exchange, no LDS. Its time is the strided
    // loads, not arithmetic: reduced outputs throughout (the lazy variant
    // measured 63.0 -> 68.5 us for the base LDE's first pass).
    dft_regs<LR1, V>(x);
    if (a.twiddle_out && m) {
      const uint32_t st = tw24(a.tw_lo, a.tw_hi, m << (24 - log_mprev));  // w_{M_{p-1}}^m
      uint32_t cur = st;
#pragma unroll
      for (int ta = 1; ta < R1; ++ta) {
        vscale<V>(x[ta], cur);
        cur = mul(cur, st);
      }
    }
#pragma unroll
    for (int ta = 0; ta < R1; ++ta) vstore<V>(a.dst + base + (uint32_t)ta * jstride, x[ta]);
    return;
  }
  // outputs 1.. come out in [0, 2p): the twiddle multiply that follows
  // reduces them; a line that has none reduces them explicitly
  dft_regs<LR1, V, lazy_but_first<LR1>()>(x);
  if (R2 > 1 && jb) {  // inner twiddles w_L^(j_b t_a) = tw_hi[(j_b t_a mod L) 2^(12 - LL)] (L <= 4096)
#pragma unroll
    for (int ta = 1; ta < R1; ++ta) vscale<V>(x[ta], *gptr<uint32_t>(a.tw_hi + (((jb * (uint32_t)ta) & ((1u << LL) - 1)) << (12 - LL))));
  } else {
#pragma unroll
    for (int ta = 1; ta < R1; ++ta) vcanon<V>(x[ta]);
  }
  Vec<V>* l4 = reinterpret_cast<Vec<V>*>(lds);
#pragma unroll
  for (int ta = 0; ta < R1; ++ta) l4[((uint32_t)ta * R2 + jb) * QV + cq] = x[ta];
  __syncthreads();

  // ---- stage B: tasks t_a = j_b + k R2, outputs t = t_a + R1 t_b ----
  uint32_t nat0 = 0;
  if (a.natural) nat0 = a.n_pass == 2 ? b : (b >> a.log_l2) + ((b & ((1u << a.log_l2) - 1)) << a.log_l1);
  const uint32_t log_nat_stride = a.n_pass == 2 ? a.log_l1 : a.log_l1 + a.log_l2;
#pragma unroll
  for (int k = 0; k < R1 / R2; ++k) {
    const uint32_t ta = jb + (uint32_t)k * R2;
    Vec<V> z[R2];
#pragma unroll
    for (int j = 0; j < R2; ++j) z[j] = l4[(ta * R2 + j) * QV + cq];
    if (a.twiddle_out) {  // every output is multiplied next (reduced instead when m = 0)
      dft_regs<LR2, V, (uint32_t)((1ull << R2) - 1)>(z);
      if (m) {
        const uint32_t sh = 24 - log_mprev;
        uint32_t cur = tw24(a.tw_lo, a.tw_hi, (m * ta) << sh);
        const uint32_t st = tw24(a.tw_lo, a.tw_hi, (m << LR1) << sh);
#pragma unroll
        for (int tb = 0; tb < R2; ++tb) {
          vscale<V>(z[tb], cur);
          cur = mul(cur, st);
        }
      } else {
#pragma unroll
        for (int tb = 0; tb < R2; ++tb) vcanon<V>(z[tb]);
      }
    } else {
      dft_regs<LR2, V>(z);
    }
#pragma unroll
    for (int tb = 0; tb < R2; ++tb) {
      const uint32_t t = ta + (uint32_t)tb * R1;
      uint32_t* p;
      if (a.natural) p = a.dst + (((nat0 + (t << log_nat_stride))) << a.log_b) + col;
      else p = a.dst + base + t * jstride;
      vstore<V>(p, z[tb]);
    }
  }
}

template <int LR1, int LR2, int V>
__global__ void __launch_bounds__(256) ntt_kernel(NttArgs a) {
  extern __shared__ uint32_t lds[];
  ntt_block<LR1, LR2, V>(a, blockIdx.x, lds);
}

// Workgroup i of a grid of n (n a multiple of 16) takes tile
// 16 (i / 16) + 2 (i mod 8) + (i / 8 mod 2): tiles 2t and 2t + 1 go to launch
// slots 8 apart. Workgroups are dealt to the 8 XCDs in turn, so the two run on
// one XCD at about the same time and the second finds in that XCD's L2 the
// 128-byte lines the first brought in. Other grids keep the identity.
__host__ __device__ __forceinline__ uint32_t xcd_pair_tile(uint32_t i, uint32_t n) {
  if (n & 15u) return i;
  return (i & ~15u) | ((i & 7u) << 1) | ((i >> 3) & 1u);
}

// Passes of 512 or 1024 rows in two register stages of up to 32 points
// (ntt_block with 2^LR1 x 2^LR2, LR1 = 5), 256 threads, two words per lane and
// line: a workgroup's tile is 1024 rows x 16 columns (64 KiB of LDS, two
// workgroups per CU) or 512 rows x 32 columns. A 1024-row tile covers half of
// each 128-byte row of a 32-column work array; the tile of the other half runs
// beside it on the same XCD (xcd_pair_tile), so each line is fetched from HBM
// once (FETCH_SIZE of a pass = its 128 MiB). The base LDE of a 2^24-coefficient
// commitment (2^20 x 16 -> 2^21 rows) in two such passes measured 201 us
// against 189 us for the three-pass plan (50 launches, host-timed; one word
// per lane: 235 us, 74.6 VALU instructions per element per pass against 48.3
// for an 8-bit pass): the passes save HBM traffic (2.3x the algorithmic bytes
// instead of 3.7x) but not VALU work, which bounds the transform, and their
// tiles leave two waves per SIMD. It stays a tuner candidate.
template <int LR1, int LR2, int V>
__global__ void __launch_bounds__(256) ntt_wide_kernel(NttArgs a) {
  extern __shared__ uint32_t lds[];
  ntt_block<LR1, LR2, V>(a, xcd_pair_tile(blockIdx.x, gridDim.x), lds);
}

// ---- passes of L = R1 R2 R3 rows: three register stages, two LDS exchanges ----
// A transform of 2^17..2^20 rows in two such passes (L up to 1024) reads and
// writes the work array once less than in three passes of <= 256 rows. Row
// index of a line's element: j = j1 T + j2 R3 + j3 (T = R2 R3 threads per
// line, thread jb = j2 R3 + j3); output t = t1 + R1 t2 + R1 R2 t3:
//   stage A (thread jb):   DFT_R1 over j1 -> t1, times w_L^(jb t1);
//   stage B (task t1, j3): DFT_R2 over j2 -> t2, times w_T^(j3 t2);
//   stage C (task t1, t2): DFT_R3 over j3 -> t3.
// The stages are host/device functions of (workgroup, thread) with the
// workgroup barriers between them, so the same code runs in a host emulation.
template <int LR1, int LR2, int LR3, int V>
struct Ntt3 {
  static constexpr int R1 = 1 << LR1, R2 = 1 << LR2, R3 = 1 << LR3;
  static constexpr int LT = LR2 + LR3, T = 1 << LT;  // threads per line
  static constexpr int LL = LR1 + LT;                // log2 L
  static constexpr int KB = (R1 * R3) >> LT;         // stage-B tasks per thread
  static constexpr int KC = (R1 * R2) >> LT;         // stage-C tasks per thread
  static constexpr int LV = V == 4 ? 2 : V == 2 ? 1 : 0;
  static_assert(LR1 >= LR2 && LR1 >= LR3, "stage A is the widest");

  struct Lane {  // the thread's place: line (b, m, col) and its offsets
    uint32_t cq, jb, b, w, m, col, base, jstride;
  };
  static KB_HD Lane lane(const NttArgs& a, uint32_t bid, uint32_t tid) {
    Lane l;
    l.cq = tid & ((1u << a.log_qv) - 1);
    l.jb = tid >> a.log_qv;
    const uint32_t log_lpb = a.log_m + a.log_b - LV;  // lines per block
    const uint32_t line = (bid << a.log_qv) + l.cq;
    l.b = line >> log_lpb;
    l.w = (line & ((1u << log_lpb) - 1)) << LV;  // m B + col
    l.m = l.w >> a.log_b;
    l.col = l.w & ((1u << a.log_b) - 1);
    l.jstride = 1u << (a.log_m + a.log_b);
    l.base = (l.b << (a.log_m + LL + a.log_b)) + l.w;
    return l;
  }
  // LDS slot of (row position i, line cq) in Vec<V> units
  static KB_HD uint32_t slot(const NttArgs& a, uint32_t i, uint32_t cq) { return (i << a.log_qv) + cq; }

  static KB_HD void stage_a(const NttArgs& a, const Lane& l, Vec<V>* lds) {
    Vec<V> x[R1];
    if (!a.first) {
      #pragma unroll
      for (int j1 = 0; j1 < R1; ++j1) x[j1] = vload<V>(a.src + l.base + (((uint32_t)j1 << LT) + l.jb) * l.jstride);
    } else {
      // input row r = (j1 T + jb) M + m
      load_input_rows<R1, V>(a, (l.jb << a.log_m) + l.m, LT + a.log_m, l.col, x);
    }
    dft_regs<LR1, V, lazy_but_first<LR1>()>(x);
    if (l.jb) {  // w_L^(jb t1) = tw_hi[(jb t1 mod L) 2^(12 - LL)] (L <= 4096)
      #pragma unroll
      for (int t1 = 1; t1 < R1; ++t1) vscale<V>(x[t1], *gptr<uint32_t>(a.tw_hi + (((l.jb * (uint32_t)t1) & ((1u << LL) - 1)) << (12 - LL))));
    } else {
      #pragma unroll
      for (int t1 = 1; t1 < R1; ++t1) vcanon<V>(x[t1]);
    }
    #pragma unroll
    for (int t1 = 0; t1 < R1; ++t1) lds[slot(a, ((uint32_t)t1 << LT) + l.jb, l.cq)] = x[t1];
  }

  // stage B, reading half: task q = jb + k T -> (t1, j3); y[k] its R2 outputs
  static KB_HD void stage_b_read(const NttArgs& a, const Lane& l, const Vec<V>* lds, Vec<V> (&y)[KB][R2]) {
    #pragma unroll
    for (int k = 0; k < KB; ++k) {
      const uint32_t q = l.jb + ((uint32_t)k << LT);
      const uint32_t t1 = q >> LR3, j3 = q & (R3 - 1);
      #pragma unroll
      for (int j2 = 0; j2 < R2; ++j2) y[k][j2] = lds[slot(a, (t1 << LT) + ((uint32_t)j2 << LR3) + j3, l.cq)];
      dft_regs<LR2, V, lazy_but_first<LR2>()>(y[k]);
      if (j3) {  // w_T^(j3 t2) = tw_hi[(j3 t2 mod T) 2^(12 - LT)]
        #pragma unroll
        for (int t2 = 1; t2 < R2; ++t2) vscale<V>(y[k][t2], *gptr<uint32_t>(a.tw_hi + (((j3 * (uint32_t)t2) & ((1u << LT) - 1)) << (12 - LT))));
      } else {
        #pragma unroll
        for (int t2 = 1; t2 < R2; ++t2) vcanon<V>(y[k][t2]);
      }
    }
  }
  static KB_HD void stage_b_write(const NttArgs& a, const Lane& l, Vec<V>* lds, const Vec<V> (&y)[KB][R2]) {
    #pragma unroll
    for (int k = 0; k < KB; ++k) {
      const uint32_t q = l.jb + ((uint32_t)k << LT);
      const uint32_t t1 = q >> LR3, j3 = q & (R3 - 1);
      #pragma unroll
      for (int t2 = 0; t2 < R2; ++t2)
        lds[slot(a, ((((uint32_t)t2 << LR1) + t1) << LR3) + j3, l.cq)] = y[k][t2];
    }
  }

  // stage C: task u = jb + k T = t1 + R1 t2; outputs t = u + R1 R2 t3
  static KB_HD void stage_c(const NttArgs& a, const Lane& l, const Vec<V>* lds) {
    uint32_t nat0 = 0;
    if (a.natural) nat0 = a.n_pass == 2 ? l.b : (l.b >> a.log_l2) + ((l.b & ((1u << a.log_l2) - 1)) << a.log_l1);
    const uint32_t log_nat_stride = a.n_pass == 2 ? a.log_l1 : a.log_l1 + a.log_l2;
    const uint32_t log_mprev = a.log_m + LL;
    #pragma unroll
    for (int k = 0; k < KC; ++k) {
      const uint32_t u = l.jb + ((uint32_t)k << LT);
      Vec<V> z[R3];
      #pragma unroll
      for (int j3 = 0; j3 < R3; ++j3) z[j3] = lds[slot(a, (u << LR3) + (uint32_t)j3, l.cq)];
      if (a.twiddle_out) {  // w_{M_{p-1}}^(m t); every output is multiplied next (reduced when m = 0)
        dft_regs<LR3, V, (uint32_t)((1ull << R3) - 1)>(z);
        if (l.m) {
          const uint32_t sh = 24 - log_mprev;
          uint32_t cur = tw24(a.tw_lo, a.tw_hi, (l.m * u) << sh);
          const uint32_t st = tw24(a.tw_lo, a.tw_hi, (l.m << (LR1 + LR2)) << sh);
          #pragma unroll
          for (int t3 = 0; t3 < R3; ++t3) {
            vscale<V>(z[t3], cur);
            cur = mul(cur, st);
          }
        } else {
          #pragma unroll
          for (int t3 = 0; t3 < R3; ++t3) vcanon<V>(z[t3]);
        }
      } else {
        dft_regs<LR3, V>(z);
      }
      #pragma unroll
      for (int t3 = 0; t3 < R3; ++t3) {
        const uint32_t t = u + ((uint32_t)t3 << (LR1 + LR2));
        uint32_t* p;
        if (a.natural) p = a.dst + (((nat0 + (t << log_nat_stride))) << a.log_b) + l.col;
        else p = a.dst + l.base + t * l.jstride;
        vstore<V>(p, z[t3]);
      }
    }
  }
};

template <int LR1, int LR2, int LR3, int V>
__global__ void __launch_bounds__(512) ntt3_kernel(NttArgs a) {
  using K = Ntt3<LR1, LR2, LR3, V>;
  extern __shared__ uint32_t lds_words[];
  Vec<V>* lds = reinterpret_cast<Vec<V>*>(lds_words);
  const typename K::Lane l = K::lane(a, blockIdx.x, threadIdx.x);
  K::stage_a(a, l, lds);
  __syncthreads();
  Vec<V> y[K::KB][K::R2];
  K::stage_b_read(a, l, lds, y);
  __syncthreads();
  K::stage_b_write(a, l, lds, y);
  __syncthreads();
  K::stage_c(a, l, lds);
}

}  // namespace kb
