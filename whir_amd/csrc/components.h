// Component entry points: the concrete sumcheck / eq / Merkle-opening
// functions of the reference as individual C-ABI calls on device buffers, for
// a caller that swaps one function body at its call site rather than the whole
// prover (crates/poly/src/{point,eq,poly}.rs, crates/whir/src/sumcheck/
// {expr,combine}.rs, crates/merkle/src/poseidon_packed.rs:145-233).
// Scalars (scale, r, alpha) are passed by value as 4 Montgomery words; arrays
// are device pointers; work is enqueued on `stream` (scratch comes from the
// stream-ordered allocator and is released on the same stream).
// Included at the end of whir_amd.hip (it uses the kernels and helpers there).
#pragma once

namespace kb {

struct ExtArg {
  uint32_t c[4];
};
__device__ __forceinline__ Ext ext_of(const ExtArg& a) { return Ext{{a.c[0], a.c[1], a.c[2], a.c[3]}}; }

// Split eq tables of one point per workgroup: block j builds eq0 (the low m0
// variables, unscaled) and eq1 (the high m1, times scale * alpha^j) of
// points[j] (k = m0 + m1 extension coordinates) into pool + j * stride.
__global__ void __launch_bounds__(SERIAL_THREADS) eq_split_kernel(const Ext* __restrict__ points, uint32_t k,
                                                                  ExtArg scale, ExtArg alpha, uint32_t m0,
                                                                  uint32_t m1, Ext* pool, uint64_t stride) {
  __shared__ Ext s_z[MAX_K];
  __shared__ Ext s_scale;
  __shared__ EqPairScratch sc;
  const uint32_t j = blockIdx.x;
  for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) s_z[i] = ld_ext(points + (uint64_t)j * k + i);
  if (threadIdx.x == 0) s_scale = emul(ext_of(scale), epow(ext_of(alpha), j));
  __syncthreads();
  Ext* eq0 = pool + j * stride;
  build_eq_pair(eq0, s_z, (int)m0, eone(), eq0 + (1ull << m0), s_z + m0, (int)m1, s_scale, sc);
}

// out[x] = eq0[x mod 2^m0] * eq1[x >> m0]
__global__ void __launch_bounds__(SC_THREADS) eq_expand_kernel(const Ext* __restrict__ eq0, const Ext* __restrict__ eq1,
                                                           uint32_t m0, uint64_t n, Ext* __restrict__ out) {
  const uint64_t x = (uint64_t)blockIdx.x * SC_THREADS + threadIdx.x;
  if (x >= n) return;
  st_ext(out + x, emul(ld_ext(eq0 + (x & ((1ull << m0) - 1))), ld_ext(eq1 + (x >> m0))));
}

// out[i] = sum_p partials[p * w + i], i < w (one workgroup).
__global__ void __launch_bounds__(SERIAL_THREADS) reduce_partials_kernel(const Ext* __restrict__ partials,
                                                                         uint32_t n_part, uint32_t w,
                                                                         Ext* __restrict__ out) {
  __shared__ Ext sh[SERIAL_THREADS];
  const uint32_t tid = threadIdx.x;
  const uint32_t lanes_per = SERIAL_THREADS / w;
  const uint32_t i = tid % w, p0 = tid / w;
  Ext acc = ezero();
  if (p0 < lanes_per)
    for (uint32_t p = p0; p < n_part; p += lanes_per) acc = eadd(acc, ld_ext(partials + (uint64_t)p * w + i));
  sh[tid] = acc;
  __syncthreads();
  if (tid < w) {
    Ext s = ezero();
    for (uint32_t r = 0; r < lanes_per; ++r) s = eadd(s, sh[r * w + tid]);
    st_ext(out + tid, s);
  }
}

// SplitEq::compress_lo (eq.rs:31-49, 400-539): out[j] = sum_{i < 2^m} eq(z)[i] poly[j 2^m + i], m <= MAX_F.
__global__ void __launch_bounds__(SC_THREADS) compress_lo_kernel(const uint32_t* __restrict__ poly,
                                                             const Ext* __restrict__ z, uint32_t m, uint64_t n_out,
                                                             Ext* __restrict__ out) {
  __shared__ Ext s_eq[1 << MAX_F];
  if (threadIdx.x == 0) {
    s_eq[0] = eone();
    for (uint32_t i = 0; i < m; ++i) {
      const uint32_t h = 1u << i;
      const Ext zi = ld_ext(z + i);
      for (uint32_t j = 0; j < h; ++j) {
        const Ext hi = emul(s_eq[j], zi);
        s_eq[j + h] = hi;
        s_eq[j] = esub(s_eq[j], hi);
      }
    }
  }
  __syncthreads();
  const uint64_t j = (uint64_t)blockIdx.x * SC_THREADS + threadIdx.x;
  if (j >= n_out) return;
  const uint32_t W = 1u << m;
  const uint32_t* row = poly + j * W;
  Ext acc = ezero();
  for (uint32_t i = 0; i < W; ++i) acc = eadd(acc, emul_base(s_eq[i], row[i]));
  st_ext(out + j, acc);
}

// Poly::fix_lo_var_mut (poly.rs:131-142): out[i] = a[2i] + r (a[2i+1] - a[2i]).
__global__ void __launch_bounds__(SC_THREADS) fix_lo_var_kernel(const Ext* __restrict__ a, uint64_t n_out, ExtArg r,
                                                            Ext* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * SC_THREADS + threadIdx.x;
  if (i >= n_out) return;
  const Ext a0 = ld_ext(a + 2 * i), a1 = ld_ext(a + 2 * i + 1);
  st_ext(out + i, eadd(a0, emul(ext_of(r), esub(a1, a0))));
}

// scales[j] = alpha^(shift + j), *count = n (inputs of pow_tables_kernel).
__global__ void pow_scales_kernel(ExtArg alpha, uint64_t shift, uint32_t n, Ext* scales, uint32_t* count) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0) *count = n;
  if (j < n && scales) st_ext(scales + j, epow(ext_of(alpha), shift + j));  // no scales: count only
}

// MerkleTree::query (poseidon_packed.rs:145-233): per index, the row
// (row_words base words) then the log_h sibling digests bottom-up.
__global__ void __launch_bounds__(SC_THREADS) merkle_open_kernel(const uint32_t* __restrict__ layers, uint32_t log_h,
                                                             const uint32_t* __restrict__ rows, uint32_t row_words,
                                                             const uint32_t* __restrict__ idx,
                                                             uint32_t* __restrict__ out) {
  const uint32_t q = blockIdx.x;
  const uint32_t index = idx[q];
  if (log_h < 32 && (index >> log_h) != 0) return;  // not a leaf index: output left untouched
  const uint32_t per = row_words + 8 * log_h;
  uint32_t* o = out + (uint64_t)q * per;
  for (uint32_t w = threadIdx.x; w < row_words; w += SC_THREADS) o[w] = rows[(uint64_t)index * row_words + w];
  for (uint32_t e = threadIdx.x; e < 8 * log_h; e += SC_THREADS) {
    const uint32_t i = e >> 3, w = e & 7;
    const uint32_t mid = 1u << (log_h - i - 1);
    const uint32_t sib = (index & ((mid << 1) - 1)) ^ mid;
    const uint64_t layer_off = (2ull << log_h) - (2ull << (log_h - i));  // sum_{j<i} 2^(log_h-j)
    o[row_words + e] = layers[(layer_off + sib) * 8 + w];
  }
}

// The sorted().dedup() of stir_indexes (crates/whir/src/whir.rs:63-77) for n
// given values: one workgroup of SERIAL_THREADS threads running
// sort_dedup_block, as stir_begin_kernel does on the drawn indices.
__global__ void __launch_bounds__(SERIAL_THREADS) sort_dedup_kernel(const uint32_t* __restrict__ in, uint32_t n,
                                                                   uint32_t* __restrict__ out, uint32_t* count) {
  __shared__ uint32_t s_v[MAX_STIR];
  __shared__ uint32_t s_scan[SERIAL_THREADS];
  __shared__ uint32_t s_out[MAX_STIR];
  for (uint32_t i = threadIdx.x; i < n; i += SERIAL_THREADS) s_v[i] = in[i];
  const uint32_t c = sort_dedup_block(s_v, n, s_scan, s_out);
  for (uint32_t i = threadIdx.x; i < c; i += SERIAL_THREADS) out[i] = s_out[i];
  if (threadIdx.x == 0) *count = c;
}

}  // namespace kb

// ============================ component C ABI ============================

static ExtArg ext_arg(const uint32_t* v) {
  ExtArg a;
  for (int i = 0; i < 4; ++i) a.c[i] = v ? v[i] : 0u;
  return a;
}

// Split eq tables of n points into a fresh pool; returns the pool (caller frees
// it on the stream) and the split.
static int eq_split_enqueue(const uint32_t* d_points, uint32_t n, uint32_t k, const uint32_t* scale,
                            const uint32_t* alpha, hipStream_t s, Ext** pool, uint32_t& m0, uint32_t& m1,
                            uint64_t& stride) {
  m0 = k / 2;
  m1 = k - m0;
  stride = (1ull << m0) + (1ull << m1);
  HIP_TRY(hipMallocAsync((void**)pool, stride * n * 16, s));
  const uint32_t one[4] = {ONE, 0, 0, 0};
  hipLaunchKernelGGL(eq_split_kernel, dim3(n), dim3(SERIAL_THREADS), 0, s, (const Ext*)d_points, k,
                     ext_arg(scale ? scale : one), ext_arg(alpha ? alpha : one), m0, m1, *pool, stride);
  return WHIR_OK;
}

extern "C" int whir_eq_table(whir_ctx* c, const uint32_t* d_z, uint32_t k, const uint32_t* scale, uint32_t* d_out,
                             void* stream) {
  if (!c || !d_z || !d_out || !scale || k < 1 || k > MAX_K) return WHIR_EINVAL;
  hipStream_t s = pick_stream(c, stream);
  Ext* pool;
  uint32_t m0, m1;
  uint64_t stride;
  int rc = eq_split_enqueue(d_z, 1, k, scale, nullptr, s, &pool, m0, m1, stride);
  if (rc) return rc;
  hipLaunchKernelGGL(eq_expand_kernel, dim3(blocks_for(1ull << k)), dim3(SC_THREADS), 0, s, pool, pool + (1ull << m0),
                     m0, 1ull << k, (Ext*)d_out);
  HIP_TRY(hipFreeAsync(pool, s));
  HIP_TRY(hipGetLastError());
  return WHIR_OK;
}

extern "C" int whir_eval_ext(whir_ctx* c, const uint32_t* d_poly, uint32_t k, const uint32_t* d_z, uint32_t* d_out,
                             void* stream) {
  if (!c || !d_poly || !d_z || !d_out || k < 1 || k > MAX_K) return WHIR_EINVAL;
  hipStream_t s = pick_stream(c, stream);
  Ext* pool;
  uint32_t m0, m1;
  uint64_t stride;
  int rc = eq_split_enqueue(d_z, 1, k, nullptr, nullptr, s, &pool, m0, m1, stride);
  if (rc) return rc;
  Ext* partials;
  HIP_TRY(hipMallocAsync((void**)&partials, (1ull << m1) * 16, s));
  hipLaunchKernelGGL(eval_ext_kernel, dim3(1u << m1), dim3(SC_THREADS), 0, s, (const Ext*)d_poly, pool,
                     pool + (1ull << m0), m0, partials);
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(SERIAL_THREADS), 0, s, partials, 1u << m1, 1u,
                     (Ext*)d_out);
  HIP_TRY(hipFreeAsync(partials, s));
  HIP_TRY(hipFreeAsync(pool, s));
  HIP_TRY(hipGetLastError());
  return WHIR_OK;
}

// compress_hi over the top m variables of a 2^k base polynomial, k - m <= 5;
// eval_base is the case m = k. Scratch: the point's split eq tables (`pool`,
// compress_hi_pool_words(m) Ext) and the per-block partial sums (`partials`,
// compress_hi_part_words(k, m) Ext); both are written in full before they are
// read, so neither needs clearing.
static bool compress_hi_args_ok(const uint32_t* d_poly, uint32_t k, const uint32_t* d_z, uint32_t m,
                                const uint32_t* d_out) {
  return d_poly && d_z && d_out && m >= 1 && m <= k && k - m <= (uint32_t)MAX_F && m <= (uint32_t)MAX_K;
}
static uint64_t compress_hi_pool_ext(uint32_t m) { return (1ull << (m / 2)) + (1ull << (m - m / 2)); }
static uint64_t compress_hi_part_ext(uint32_t k, uint32_t m) { return (1ull << (m - m / 2)) << (k - m); }

static int compress_hi_enqueue(const uint32_t* d_poly, uint32_t k, const uint32_t* d_z, uint32_t m, uint32_t* d_out,
                               hipStream_t s, Ext* pool, Ext* partials) {
  const uint32_t m0 = m / 2, m1 = m - m0;
  const uint32_t one[4] = {ONE, 0, 0, 0};
  hipLaunchKernelGGL(eq_split_kernel, dim3(1), dim3(SERIAL_THREADS), 0, s, (const Ext*)d_z, m, ext_arg(one),
                     ext_arg(one), m0, m1, pool, compress_hi_pool_ext(m));
  const uint32_t W = 1u << (k - m);
  const unsigned grid = 1u << m1;
  const Ext* eq1 = pool + (1ull << m0);
  switch (k - m) {
    case 0: hipLaunchKernelGGL(compress_hi_kernel<0>, dim3(grid), dim3(SC_THREADS), 0, s, d_poly, pool, eq1, m0, partials, nullptr, 0u); break;
    case 1: hipLaunchKernelGGL(compress_hi_kernel<1>, dim3(grid), dim3(SC_THREADS), 0, s, d_poly, pool, eq1, m0, partials, nullptr, 0u); break;
    case 2: hipLaunchKernelGGL(compress_hi_kernel<2>, dim3(grid), dim3(SC_THREADS), 0, s, d_poly, pool, eq1, m0, partials, nullptr, 0u); break;
    case 3: hipLaunchKernelGGL(compress_hi_kernel<3>, dim3(grid), dim3(SC_THREADS), 0, s, d_poly, pool, eq1, m0, partials, nullptr, 0u); break;
    case 4: hipLaunchKernelGGL(compress_hi_kernel<4>, dim3(grid), dim3(SC_THREADS), 0, s, d_poly, pool, eq1, m0, partials, nullptr, 0u); break;
    default: hipLaunchKernelGGL(compress_hi_kernel<5>, dim3(grid), dim3(SC_THREADS), 0, s, d_poly, pool, eq1, m0, partials, nullptr, 0u); break;
  }
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(SERIAL_THREADS), 0, s, partials, grid, W, (Ext*)d_out);
  HIP_TRY(hipGetLastError());
  return WHIR_OK;
}

extern "C" int whir_compress_hi(whir_ctx* c, const uint32_t* d_poly, uint32_t k, const uint32_t* d_z, uint32_t m,
                                uint32_t* d_out, void* stream) {
  if (!c || !compress_hi_args_ok(d_poly, k, d_z, m, d_out)) return WHIR_EINVAL;
  hipStream_t s = pick_stream(c, stream);
  Ext *pool = nullptr, *partials = nullptr;
  HIP_TRY(hipMallocAsync((void**)&pool, compress_hi_pool_ext(m) * 16, s));
  if (hipMallocAsync((void**)&partials, compress_hi_part_ext(k, m) * 16, s) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFreeAsync(pool, s);
    return WHIR_EOOM;
  }
  const int rc = compress_hi_enqueue(d_poly, k, d_z, m, d_out, s, pool, partials);
  HIP_TRY(hipFreeAsync(partials, s));
  HIP_TRY(hipFreeAsync(pool, s));
  return rc;
}

// whir_compress_hi with its scratch in the library's static device array
// (whir_ctx::scratch, the address of `scratch` on the context's device):
// no allocation or release of any kind.
extern "C" int whir_compress_hi_noalloc(whir_ctx* c, const uint32_t* d_poly, uint32_t k, const uint32_t* d_z,
                                        uint32_t m, uint32_t* d_out, void* stream) {
  if (!c || !c->scratch || !compress_hi_args_ok(d_poly, k, d_z, m, d_out)) return WHIR_EINVAL;
  const uint64_t pool_bytes = (compress_hi_pool_ext(m) * 16 + 255) & ~255ull;
  const uint64_t need = pool_bytes + compress_hi_part_ext(k, m) * 16;
  if (need > WHIR_NOALLOC_SCRATCH_BYTES) {
    fprintf(stderr, "whir_compress_hi_noalloc: k=%u m=%u needs %llu bytes of scratch, above the static %u\n", k, m,
            (unsigned long long)need, (unsigned)WHIR_NOALLOC_SCRATCH_BYTES);
    return WHIR_EINVAL;
  }
  hipStream_t s = pick_stream(c, stream);
  Ext* pool = (Ext*)c->scratch;
  Ext* partials = (Ext*)((char*)c->scratch + pool_bytes);
  return compress_hi_enqueue(d_poly, k, d_z, m, d_out, s, pool, partials);
}

extern "C" int whir_compress_lo(whir_ctx* c, const uint32_t* d_poly, uint32_t k, const uint32_t* d_z, uint32_t m,
                                uint32_t* d_out, void* stream) {
  if (!c || !d_poly || !d_z || !d_out || m > k || m > (uint32_t)MAX_F) return WHIR_EINVAL;
  hipStream_t s = pick_stream(c, stream);
  const uint64_t n_out = 1ull << (k - m);
  hipLaunchKernelGGL(compress_lo_kernel, dim3(blocks_for(n_out)), dim3(SC_THREADS), 0, s, d_poly, (const Ext*)d_z, m,
                     n_out, (Ext*)d_out);
  HIP_TRY(hipGetLastError());
  return WHIR_OK;
}

extern "C" int whir_sumcheck_coeffs(whir_ctx* c, const uint32_t* d_f, const uint32_t* d_w, uint32_t k,
                                    uint32_t* d_c0c2, void* stream) {
  if (!c || !d_f || !d_w || !d_c0c2 || k < 1 || k > 30) return WHIR_EINVAL;
  hipStream_t s = pick_stream(c, stream);
  const uint64_t n = 1ull << k;
  const unsigned g = std::min<unsigned>(1024, blocks_for(n / 2));
  Ext* partials;
  HIP_TRY(hipMallocAsync((void**)&partials, (uint64_t)g * 2 * 16, s));
  hipLaunchKernelGGL(coeffs_kernel, dim3(g), dim3(SC_THREADS), 0, s, (const Ext*)d_f, (const Ext*)d_w, n, partials);
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(SERIAL_THREADS), 0, s, partials, g, 2u, (Ext*)d_c0c2);
  HIP_TRY(hipFreeAsync(partials, s));
  HIP_TRY(hipGetLastError());
  return WHIR_OK;
}

extern "C" int whir_fix_lo_var(whir_ctx* c, const uint32_t* d_poly, uint32_t k, const uint32_t* r, uint32_t* d_out,
                               void* stream) {
  if (!c || !d_poly || !r || !d_out || k < 1 || k > 30) return WHIR_EINVAL;
  hipStream_t s = pick_stream(c, stream);
  const uint64_t n_out = 1ull << (k - 1);
  hipLaunchKernelGGL(fix_lo_var_kernel, dim3(blocks_for(n_out)), dim3(SC_THREADS), 0, s, (const Ext*)d_poly, n_out,
                     ext_arg(r), (Ext*)d_out);
  HIP_TRY(hipGetLastError());
  return WHIR_OK;
}

extern "C" int whir_combine_eqs(whir_ctx* c, uint32_t* d_w, uint32_t k, const uint32_t* d_points, uint32_t n,
                                const uint32_t* alpha, void* stream) {
  if (!c || !d_w || (!d_points && n) || !alpha || k < 1 || k > MAX_K || n > 4096) return WHIR_EINVAL;
  if (n == 0) return WHIR_OK;
  hipStream_t s = pick_stream(c, stream);
  Ext* pool;
  uint32_t m0, m1;
  uint64_t stride;
  int rc = eq_split_enqueue(d_points, n, k, nullptr, alpha, s, &pool, m0, m1, stride);
  if (rc) return rc;
  const uint64_t N = 1ull << k;
  launch_combine(s, 1u, (Ext*)d_w, N, m0, pool, n, stride, nullptr, nullptr, nullptr, m1);
  HIP_TRY(hipFreeAsync(pool, s));
  HIP_TRY(hipGetLastError());
  return WHIR_OK;
}

// d_w[i] += sum_j scale_j vars_j^i: split pow tables (powL = vars^e for the low
// m0 bits of i, powR = scale vars^(e 2^m0) for the high m1 bits) and the
// combine kernel. With `alpha` set the scales are alpha^(shift + j)
// (combine_pows), else the caller's n device scales (combine_pows_scaled).
static int combine_pows_enqueue(hipStream_t s, uint32_t* d_w, uint32_t k, const uint32_t* d_vars, uint32_t n,
                                const uint32_t* alpha, uint64_t shift, const Ext* d_scales) {
  const uint32_t m0 = k / 2, m1 = k - m0;
  Ext* scales = nullptr;
  uint32_t* count;
  uint32_t* powL;
  Ext* powR;
  HIP_TRY(hipMallocAsync((void**)&count, 4, s));
  if (alpha) HIP_TRY(hipMallocAsync((void**)&scales, (uint64_t)n * 16, s));
  HIP_TRY(hipMallocAsync((void**)&powL, ((uint64_t)n << m0) * 4, s));
  HIP_TRY(hipMallocAsync((void**)&powR, ((uint64_t)n << m1) * 16, s));
  if (alpha)
    hipLaunchKernelGGL(pow_scales_kernel, dim3((n + 255) / 256), dim3(256), 0, s, ext_arg(alpha), shift, n, scales,
                       count);
  else
    hipLaunchKernelGGL(pow_scales_kernel, dim3(1), dim3(1), 0, s, ext_arg(nullptr), 0ull, n, (Ext*)nullptr, count);
  hipLaunchKernelGGL(pow_tables_kernel, dim3(blocks_for((uint64_t)n * ((1ull << m0) + (1ull << m1)))),
                     dim3(SC_THREADS), 0, s, d_vars, alpha ? (const Ext*)scales : d_scales, (const uint32_t*)count,
                     m0, m1, powL, powR);
  const uint64_t N = 1ull << k;
  launch_combine(s, 1u, (Ext*)d_w, N, m0, nullptr, 0u, 0, powL, powR, count, m1);
  HIP_TRY(hipFreeAsync(powR, s));
  HIP_TRY(hipFreeAsync(powL, s));
  if (scales) HIP_TRY(hipFreeAsync(scales, s));
  HIP_TRY(hipFreeAsync(count, s));
  HIP_TRY(hipGetLastError());
  return WHIR_OK;
}

extern "C" int whir_combine_pows(whir_ctx* c, uint32_t* d_w, uint32_t k, const uint32_t* d_vars, uint32_t n,
                                 const uint32_t* alpha, uint64_t shift, void* stream) {
  if (!c || !d_w || (!d_vars && n) || !alpha || k < 1 || k > 30 || n > 4096) return WHIR_EINVAL;
  if (n == 0) return WHIR_OK;
  return combine_pows_enqueue(pick_stream(c, stream), d_w, k, d_vars, n, alpha, shift, nullptr);
}

extern "C" int whir_combine_pows_scaled(whir_ctx* c, uint32_t* d_w, uint32_t k, const uint32_t* d_vars, uint32_t n,
                                        const uint32_t* d_scales, void* stream) {
  if (!c || !d_w || ((!d_vars || !d_scales) && n) || k < 1 || k > 30 || n > 4096) return WHIR_EINVAL;
  if (n == 0) return WHIR_OK;
  return combine_pows_enqueue(pick_stream(c, stream), d_w, k, d_vars, n, nullptr, 0, (const Ext*)d_scales);
}

extern "C" int whir_merkle_open(whir_ctx* c, const uint32_t* d_layers, uint32_t log_h, const uint32_t* d_rows,
                                uint32_t row_words, const uint32_t* d_idx, uint32_t n, uint32_t* d_out,
                                void* stream) {
  if (!c || !d_layers || !d_rows || (!d_idx && n) || (!d_out && n) || row_words == 0 || log_h > 28) return WHIR_EINVAL;
  if (n == 0) return WHIR_OK;
  hipStream_t s = pick_stream(c, stream);
  hipLaunchKernelGGL(merkle_open_kernel, dim3(n), dim3(SC_THREADS), 0, s, d_layers, log_h, d_rows, row_words, d_idx,
                     d_out);
  HIP_TRY(hipGetLastError());
  return WHIR_OK;
}

extern "C" int whir_sort_dedup(whir_ctx* c, const uint32_t* d_in, uint32_t n, uint32_t* d_out, uint32_t* d_count,
                               void* stream) {
  if (!c || (!d_in && n) || (!d_out && n) || !d_count || n > (uint32_t)MAX_STIR) return WHIR_EINVAL;
  hipStream_t s = pick_stream(c, stream);
  hipLaunchKernelGGL(sort_dedup_kernel, dim3(1), dim3(SERIAL_THREADS), 0, s, d_in, n, d_out, d_count);
  HIP_TRY(hipGetLastError());
  return WHIR_OK;
}
