// MI355X (gfx950) WHIR prover: one translation unit holding every kernel (they
// share the __constant__ Poseidon2 tables) plus the host orchestration and the
// C ABI declared in include/whir_gpu.h.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "field.h"
#include "poseidon2.h"
#include "tx.h"
#include "ntt.h"
#include "ntt_plan.h"
#include "merkle.h"
#include "sumcheck.h"
#include "protocol.h"
#include "round_small.h"
#include "checks.h"
#include "host_params.h"
#include "host_rng.h"
#include "stack.h"
#include "../../include/whir_gpu.h"

using namespace kb;
using whir_host::Plan;
using whir_host::RoundPlan;

#define HIP_TRY(x)                                                                     \
  do {                                                                                 \
    hipError_t _e = (x);                                                               \
    if (_e != hipSuccess) {                                                            \
      fprintf(stderr, "whir_amd: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(_e), \
              __FILE__, __LINE__);                                                     \
      return WHIR_EDEVICE;                                                             \
    }                                                                                  \
  } while (0)

// ============================ context ============================

// Static device scratch of the allocation-free entry points (one array per
// GPU, alive until the process ends). Host code uses the per-device address
// from hipGetSymbolAddress (whir_ctx::scratch), never this name.
static __device__ __align__(256) unsigned char scratch[WHIR_NOALLOC_SCRATCH_BYTES];

// How one NTT runs (results do not depend on it): largest pass 2^max_logl rows,
// 2^log_threads threads per workgroup; the first pass of a transform of >= 3
// passes as an LDS-free radix-2^first_log register pass (0: all passes even);
// 2^max_lv words per lane and line; transforms of 2^17..2^20 rows in two passes
// of up to 1024 rows with three register stages (ntt3_kernel) when two_pass is
// 1, transforms of 2^18..2^20 rows in two passes of two register stages of up
// to 32 points (ntt_wide_kernel) when it is 2.
struct NttConfig {
  int max_logl = 8, log_threads = 8, first_log = 4, max_lv = 1, two_pass = 0;
};
// How one Merkle tree runs: two lanes per permutation (p2pair.h) for leaf
// layers and compression layers of at most pair_max rows / compressions
// (measured on MI355X: below ~2^16 one lane's instruction stream is exposed,
// 2^15 rows of 64 words 55.7 -> 38.2 us, a 2^15-compression layer 15.3 -> 12.0
// us; from 2^17 up one lane per permutation wins); from coop_from nodes down,
// cooperative LDS-resident subtrees of up to COOP_MAX_NODES nodes per launch
// (the latency-bound top of the tree).
struct MerkleConfig {
#if defined(WHIR_FIELD_BABYBEAR)
  uint64_t pair_max = 0;  // no two-lane kernels in this build
#else
  uint64_t pair_max = 1ull << 15;
#endif
  uint64_t coop_from = 1ull << 15;
};

struct whir_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* scratch = nullptr;  // this device's address of `scratch`
  NttConfig ntt;            // defaults (whir_ctx_set_ntt_*)
  MerkleConfig merkle;      // default (whir_ctx_set_merkle_tuning)
  // autotuner: per transform (n_in, log N, log C) and per tree (log H, row
  // words), the configuration measured fastest on this device
  bool autotune = false;
  std::mutex tune_mu;
  // (keys end with the job-count class of the launches: lockstep batches
  // are tuned with their own job count)
  std::map<std::tuple<uint64_t, int, int, uint32_t>, NttConfig> ntt_tuned;
  std::map<std::tuple<int, uint32_t, uint32_t>, MerkleConfig> merkle_tuned;
  std::vector<std::string> tune_log;  // one JSON object per tuned transform / tree
  uint32_t* tw = nullptr;  // tw_lo[4096] || tw_hi[4096]
  const uint32_t* tw_lo() const { return tw; }
  const uint32_t* tw_hi() const { return tw + 4096; }
};

// Job-count class of a launch over `jobs` lockstep proofs: 1, 2, 4, 8 or 16+.
static uint32_t jobs_class(uint32_t jobs) {
  uint32_t c = 1;
  while (c < jobs && c < 16) c <<= 1;
  return c;
}
// The tuned configuration for this size and job count, else the one tuned for
// single proofs, else the context's default.
static NttConfig ntt_config(whir_ctx* c, uint64_t n_in, int log_nn, int log_c, uint32_t jobs = 1) {
  std::lock_guard<std::mutex> g(c->tune_mu);
  auto it = c->ntt_tuned.find(std::make_tuple(n_in, log_nn, log_c, jobs_class(jobs)));
  if (it == c->ntt_tuned.end()) it = c->ntt_tuned.find(std::make_tuple(n_in, log_nn, log_c, 1u));
  return it == c->ntt_tuned.end() ? c->ntt : it->second;
}
static MerkleConfig merkle_config(whir_ctx* c, int log_h, uint32_t row_words, uint32_t jobs = 1) {
  std::lock_guard<std::mutex> g(c->tune_mu);
  auto it = c->merkle_tuned.find(std::make_tuple(log_h, row_words, jobs_class(jobs)));
  if (it == c->merkle_tuned.end()) it = c->merkle_tuned.find(std::make_tuple(log_h, row_words, 1u));
  return it == c->merkle_tuned.end() ? c->merkle : it->second;
}

// The constant tables are per-device globals built in static buffers:
// contexts created from several host threads build and upload them one at a
// time.
static std::mutex g_constants_mu;

#if defined(WHIR_FIELD_BABYBEAR)
// Poseidon2BabyBear<16|24>::new_from_rng_128(StdRng::seed_from_u64(seed)): per
// width a fresh generator, 4 initial external rounds of W constants, 4
// terminal, R_P internal (the same draw order as KoalaBear's).
template <int W>
static void plain_consts_draw(uint64_t seed, Poseidon2Plain<W>& C) {
  whir_host::ChaChaRng r(seed);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < W; ++j) C.ext_init[i][j] = r.next_field();
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < W; ++j) C.ext_term[i][j] = r.next_field();
  for (int i = 0; i < Poseidon2Plain<W>::RP; ++i) C.internal[i] = r.next_field();
  plain_diag_build<W>(C);
}

static int upload_constants(uint64_t seed) {
  std::lock_guard<std::mutex> g(g_constants_mu);
  static Poseidon2PlainConsts pc;
  plain_consts_draw<16>(seed, pc.w16);
  plain_consts_draw<24>(seed, pc.w24);
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_p2b), &pc, sizeof(pc)));
  return WHIR_OK;
}
#else
static int upload_constants(uint64_t seed) {
  std::lock_guard<std::mutex> g(g_constants_mu);
  Poseidon2Consts c;
  {
    whir_host::ChaChaRng r(seed);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 16; ++j) c.ext16_init[i][j] = r.next_field();
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 16; ++j) c.ext16_term[i][j] = r.next_field();
    for (int i = 0; i < 20; ++i) c.int16[i] = r.next_field();
  }
  {
    whir_host::ChaChaRng r(seed);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 24; ++j) c.ext24_init[i][j] = r.next_field();
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 24; ++j) c.ext24_term[i][j] = r.next_field();
    for (int i = 0; i < 23; ++i) c.int24[i] = r.next_field();
  }
  {
    static Poseidon2LazyConsts lz;
    lazy_consts_build<16>(c.ext16_init, c.ext16_term, c.int16, lz.w16);
    lazy_consts_build<24>(c.ext24_init, c.ext24_term, c.int24, lz.w24);
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_p2l), &lz, sizeof(lz)));
    static Poseidon2Wave16 wz;
    wave16_consts_build(c.ext16_term, c.int16, wz);
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_p2w), &wz, sizeof(wz)));
  }
  return WHIR_OK;
}
#endif

extern "C" int whir_ctx_create(int device, uint64_t seed, whir_ctx** out) {
  if (!out) return WHIR_EINVAL;
  HIP_TRY(hipSetDevice(device));
  whir_ctx* c = new whir_ctx();
  c->device = device;
  if (hipGetSymbolAddress(&c->scratch, HIP_SYMBOL(scratch)) != hipSuccess || !c->scratch) {
    (void)hipGetLastError();
    delete c;
    return WHIR_EDEVICE;
  }
  int rc = upload_constants(seed);
  if (rc) {
    delete c;
    return rc;
  }
  std::vector<uint32_t> tw(8192);
  const uint32_t g = kb::two_adic_generator(24);
  uint32_t x = kb::ONE;
  for (int i = 0; i < 4096; ++i) {
    tw[i] = x;
    x = kb::mul(x, g);
  }
  const uint32_t g4096 = kb::pow(g, 4096);
  x = kb::ONE;
  for (int i = 0; i < 4096; ++i) {
    tw[4096 + i] = x;
    x = kb::mul(x, g4096);
  }
  HIP_TRY(hipMalloc(&c->tw, 8192 * 4));
  HIP_TRY(hipMemcpy(c->tw, tw.data(), 8192 * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipStreamCreate(&c->stream));
  *out = c;
  return WHIR_OK;
}

// Release calls in a destroy path: every one is made; the first failure is
// reported (WHIR_EDEVICE) after the rest.
static void release_step(int& rc, hipError_t e) {
  if (e != hipSuccess && rc == WHIR_OK) rc = WHIR_EDEVICE;
}

extern "C" int whir_ctx_destroy(whir_ctx* c) {
  if (!c) return WHIR_OK;
  int rc = WHIR_OK;
  release_step(rc, hipFree(c->tw));
  if (c->stream) release_step(rc, hipStreamDestroy(c->stream));
  delete c;
  return rc;
}

extern "C" int whir_ctx_set_ntt_shape(whir_ctx* c, int first_pass_log, int log_vec) {
  if (!c || !(first_pass_log == 0 || (first_pass_log >= 2 && first_pass_log <= 4)) || log_vec < 0 || log_vec > 2)
    return WHIR_EINVAL;
  c->ntt.first_log = first_pass_log;
  c->ntt.max_lv = log_vec;
  return WHIR_OK;
}

extern "C" int whir_ctx_set_ntt_tuning(whir_ctx* c, int max_log_pass, int log_threads) {
  if (!c || max_log_pass < 1 || max_log_pass > 8 || log_threads < 6 || log_threads > 9) return WHIR_EINVAL;
  c->ntt.max_logl = max_log_pass;
  c->ntt.log_threads = log_threads;
  return WHIR_OK;
}

extern "C" int whir_ctx_set_ntt_passes(whir_ctx* c, int two_pass) {
  if (!c || two_pass < 0 || two_pass > 2) return WHIR_EINVAL;
  c->ntt.two_pass = two_pass;
  return WHIR_OK;
}

extern "C" int whir_ctx_set_merkle_tuning(whir_ctx* c, int pair_max_log, int coop_from_log) {
  if (!c || pair_max_log < 0 || pair_max_log > 24 || coop_from_log < 5 || coop_from_log > 24) return WHIR_EINVAL;
  c->merkle.pair_max = 1ull << pair_max_log;
  c->merkle.coop_from = 1ull << coop_from_log;
  return WHIR_OK;
}

extern "C" int whir_ctx_set_autotune(whir_ctx* c, int enable) {
  if (!c) return WHIR_EINVAL;
  c->autotune = enable != 0;
  return WHIR_OK;
}

static hipStream_t pick_stream(whir_ctx* c, void* s) { return s ? (hipStream_t)s : c->stream; }

// ============================ DFT ============================

typedef void (*NttKernelFn)(NttArgs);

template <int V>
static NttKernelFn ntt_reg_kernel_for(int logl) {
  switch (logl) {
    case 2: return ntt_kernel<2, 0, V>;
    case 3: return ntt_kernel<3, 0, V>;
    case 4: return ntt_kernel<4, 0, V>;
    default: return nullptr;
  }
}

template <int V>
static NttKernelFn ntt3_kernel_for(int logl) {
  switch (logl) {
    case 9: return ntt3_kernel<3, 3, 3, V>;
    case 10: return ntt3_kernel<4, 3, 3, V>;
    default: return nullptr;
  }
}

template <int V>
static NttKernelFn ntt_wide_kernel_for(int logl) {
  switch (logl) {
    case 9: return ntt_wide_kernel<NTT_WIDE_LR1, 4, V>;
    case 10: return ntt_wide_kernel<NTT_WIDE_LR1, 5, V>;
    default: return nullptr;
  }
}

template <int V>
static NttKernelFn ntt_kernel_for(int logl) {
  switch (logl) {
    case 1: return ntt_kernel<1, 0, V>;
    case 2: return ntt_kernel<1, 1, V>;
    case 3: return ntt_kernel<2, 1, V>;
    case 4: return ntt_kernel<2, 2, V>;
    case 5: return ntt_kernel<3, 2, V>;
    case 6: return ntt_kernel<3, 3, V>;
    case 7: return ntt_kernel<4, 3, V>;
    case 8: return ntt_kernel<4, 4, V>;
    default: return nullptr;
  }
}

// out (N x C, natural rows) = DFT of the n_in-row matrix `in` zero-padded to
// N = 2^log_nn rows; `tmp` has room for N * C words (used when P >= 2).
static bool ntt_plan_for(const NttConfig& cfg, uint64_t n_in, int log_nn, int log_c, NttPlan& pl) {
  return ntt_plan(n_in, log_nn, log_c, pl, cfg.max_logl, cfg.first_log, cfg.max_lv, cfg.two_pass);
}

// `force`: run with this configuration instead of the tuned or default one.
static int ntt_enqueue(whir_ctx* c, const uint32_t* in, uint64_t n_in, int log_nn, int log_c, uint32_t* out,
                       uint32_t* tmp, hipStream_t s, const NttConfig* force = nullptr, uint32_t jobs = 1) {
  const NttConfig cfg = force ? *force : ntt_config(c, n_in, log_nn, log_c, jobs);
  NttPlan pl;
  if (!ntt_plan_for(cfg, n_in, log_nn, log_c, pl)) return WHIR_EINVAL;
  for (int p = 0; p < pl.P; ++p) {
    NttPass ps;
    ntt_pass_setup(pl, p, n_in, cfg.log_threads, in, out, tmp, c->tw_lo(), c->tw_hi(), ps);
    const int l = pl.logl[p];
    NttKernelFn fn = ps.kind == NTT_PASS_REG
                         ? (pl.lv == 2 ? ntt_reg_kernel_for<4>(l) : pl.lv == 1 ? ntt_reg_kernel_for<2>(l)
                                                                               : ntt_reg_kernel_for<1>(l))
                     : ps.kind == NTT_PASS_WIDE ? (pl.lv == 1 ? ntt_wide_kernel_for<2>(l) : nullptr)
                     : ps.kind == NTT_PASS_TRI
                         ? (pl.lv == 2 ? ntt3_kernel_for<4>(l) : pl.lv == 1 ? ntt3_kernel_for<2>(l)
                                                                            : ntt3_kernel_for<1>(l))
                         : (pl.lv == 2 ? ntt_kernel_for<4>(l) : pl.lv == 1 ? ntt_kernel_for<2>(l)
                                                                           : ntt_kernel_for<1>(l));
    if (!fn || ps.lds > 65536) return WHIR_EINVAL;
    hipLaunchKernelGGL(fn, dim3(ps.grid, 1, jobs), dim3(ps.threads), ps.lds, s, ps.a);
  }
  return WHIR_OK;
}

extern "C" int whir_dft_batch(whir_ctx* c, const uint32_t* d_in, uint64_t n_in_rows, uint32_t log_h, uint32_t width,
                              uint32_t* d_out, void* stream) {
  if (!c || !d_in || !d_out || width == 0 || (width & (width - 1)) || n_in_rows > (1ull << log_h))
    return WHIR_EINVAL;
  int log_c = 0;
  while ((1u << log_c) < width) ++log_c;
  hipStream_t s = pick_stream(c, stream);
  uint32_t* tmp = nullptr;
  NttPlan pl;
  if (!ntt_plan_for(ntt_config(c, n_in_rows, (int)log_h, log_c), n_in_rows, (int)log_h, log_c, pl))
    return WHIR_EINVAL;
  if (pl.P > 1) HIP_TRY(hipMallocAsync((void**)&tmp, ((size_t)width << log_h) * 4, s));
  int rc = ntt_enqueue(c, d_in, n_in_rows, (int)log_h, log_c, d_out, tmp, s);
  if (tmp) HIP_TRY(hipFreeAsync(tmp, s));
  HIP_TRY(hipGetLastError());
  return rc;
}

// ============================ Merkle ============================

static int leaves_enqueue(const uint32_t* rows, uint64_t n_rows, uint32_t row_words, uint32_t* leaves,
                          hipStream_t s, const MerkleConfig& mc, uint32_t jobs) {
  const dim3 grid((unsigned)((n_rows + MERKLE_THREADS - 1) / MERKLE_THREADS), 1, jobs), block(MERKLE_THREADS);
#if !defined(WHIR_FIELD_BABYBEAR)
  if (n_rows <= mc.pair_max && row_words % 16 == 0) {
    const dim3 grid2((unsigned)((2 * n_rows + MERKLE_THREADS - 1) / MERKLE_THREADS), 1, jobs);
    switch (row_words) {
#define LEAF_PAIR_CASE(W) \
  case W: hipLaunchKernelGGL(merkle_leaves_pair_kernel<W>, grid2, block, 0, s, rows, n_rows, leaves); return WHIR_OK;
      LEAF_PAIR_CASE(16) LEAF_PAIR_CASE(32) LEAF_PAIR_CASE(64) LEAF_PAIR_CASE(128)
#undef LEAF_PAIR_CASE
      default: break;
    }
  }
#endif
  switch (row_words) {
#define LEAF_CASE(W) \
  case W: hipLaunchKernelGGL(merkle_leaves_kernel<W>, grid, block, 0, s, rows, n_rows, leaves); break;
    LEAF_CASE(1) LEAF_CASE(2) LEAF_CASE(4) LEAF_CASE(8) LEAF_CASE(16) LEAF_CASE(32) LEAF_CASE(64) LEAF_CASE(128)
#undef LEAF_CASE
    default: return WHIR_EINVAL;
  }
  return WHIR_OK;
}

// Leaves + every layer into `layers` (2*2^log_h - 1 digests). Root at the end.
// Wide layers use one compression per lane (throughput-bound); from
// mc.coop_from nodes down, cooperative subtrees (latency-bound part).
// With `top` set, the last cooperative launch (the top <= COOP_MAX_NODES nodes
// down to the root) is left to the caller: *top / *top_n = that layer and its
// node count (tree_top_claim_kernel runs it with the transcript step).
// With `done_levels` >= 0 the leaves and that many wide layers are already in
// `layers` (the commit work queue of a set built them), and the launches start
// at the next layer.
static int merkle_enqueue(const uint32_t* rows, int log_h, uint32_t row_words, uint32_t* layers, hipStream_t s,
                          const MerkleConfig& mc, hipEvent_t* leaf_ev = nullptr, uint32_t** top = nullptr,
                          uint32_t* top_n = nullptr, uint32_t jobs = 1, int done_levels = -1) {
  const uint64_t H = 1ull << log_h;
  uint32_t* cur = layers;
  uint64_t n = H;
  if (done_levels < 0) {
    if (leaf_ev) HIP_TRY(hipEventRecord(leaf_ev[0], s));
    int rc = leaves_enqueue(rows, H, row_words, layers, s, mc, jobs);
    if (rc) return rc;
    if (leaf_ev) HIP_TRY(hipEventRecord(leaf_ev[1], s));
  } else {
    for (int l = 0; l < done_levels; ++l) {
      cur += n * 8;
      n /= 2;
    }
  }
  while (n > std::max<uint64_t>(mc.coop_from, COOP_MAX_NODES)) {
    const uint64_t half = n / 2;
#if !defined(WHIR_FIELD_BABYBEAR)
    if (half <= mc.pair_max)
      hipLaunchKernelGGL(merkle_layer_pair_kernel, dim3((unsigned)((2 * half + MERKLE_THREADS - 1) / MERKLE_THREADS), 1, jobs),
                         dim3(MERKLE_THREADS), 0, s, cur, half, cur + n * 8);
    else
#endif
      hipLaunchKernelGGL(merkle_layer_kernel, dim3((unsigned)((half + MERKLE_THREADS - 1) / MERKLE_THREADS), 1, jobs),
                         dim3(MERKLE_THREADS), 0, s, cur, half, cur + n * 8);
    cur += n * 8;
    n = half;
  }
  // cooperative stages: G residue classes of at most COOP_MAX_NODES nodes,
  // each reduced to its node of the G-node layer
  while (n > 1) {
    const uint64_t kk = std::min<uint64_t>(n, COOP_MAX_NODES);
    const uint32_t G = (uint32_t)(n / kk);
    if (G == 1 && top) break;
    hipLaunchKernelGGL(merkle_coop_subtrees_kernel, dim3(G, 1, jobs), dim3(COOP_THREADS), 0, s, cur, (uint32_t)n, G);
    while (n > G) {  // advance to the G-node layer
      cur += n * 8;
      n /= 2;
    }
  }
  if (top) {
    *top = cur;
    *top_n = (uint32_t)n;
  }
  return WHIR_OK;
}

// Layers merkle_enqueue computes with one or two lanes per compression before
// its cooperative stages (the ones a commit work queue may take over).
static int merkle_wide_levels(int log_h, const MerkleConfig& mc) {
  int l = 0;
  for (uint64_t n = 1ull << log_h; n > std::max<uint64_t>(mc.coop_from, COOP_MAX_NODES); n /= 2) ++l;
  return l;
}

extern "C" int whir_merkle_commit(whir_ctx* c, const uint32_t* d_rows, uint32_t log_h, uint32_t row_words,
                                  uint32_t* d_layers, uint32_t* root_host, void* stream) {
  if (!c || !d_rows || !d_layers || log_h > 28) return WHIR_EINVAL;
  hipStream_t s = pick_stream(c, stream);
  int rc = merkle_enqueue(d_rows, (int)log_h, row_words, d_layers, s, merkle_config(c, (int)log_h, row_words));
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  if (root_host) {
    const uint64_t H = 1ull << log_h;
    HIP_TRY(hipMemcpyAsync(root_host, d_layers + (2 * H - 2) * 8, 32, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  return WHIR_OK;
}

template <int W>
__global__ void permute_kernel(uint32_t* st, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t s[W];
#pragma unroll
  for (int j = 0; j < W; ++j) s[j] = st[i * W + j];
  permute<W>(s);
#pragma unroll
  for (int j = 0; j < W; ++j) st[i * W + j] = s[j];
}

// Cooperative P16, one wavefront per state: WAVE = 0, the 16-lane row form of
// the Merkle tree tops (every row a copy); WAVE = 1, the device transcript's
// one-wavefront form (KoalaBear; BabyBear's transcript uses the row form).
template <int WAVE>
__global__ void permute16_coop_kernel(uint32_t* st, uint64_t n) {
  const uint64_t i = blockIdx.x;
  if (i >= n) return;
  uint32_t s = st[i * 16 + (threadIdx.x & 15)];
  CoopP16 K;
  K.load();
#if !defined(WHIR_FIELD_BABYBEAR)
  if constexpr (WAVE) permute16_wave(s, K);
  else
#endif
    permute16_coop(s, K);
  if (threadIdx.x < 16) st[i * 16 + threadIdx.x] = s;
}

extern "C" int whir_poseidon2_permute(whir_ctx* c, int width, uint32_t* d_states, uint64_t n, void* stream) {
  if (!c || !d_states) return WHIR_EINVAL;
  hipStream_t s = pick_stream(c, stream);
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  if (width == 16) hipLaunchKernelGGL(permute_kernel<16>, grid, block, 0, s, d_states, n);
  else if (width == 24) hipLaunchKernelGGL(permute_kernel<24>, grid, block, 0, s, d_states, n);
  else if (width == -16) hipLaunchKernelGGL(permute16_coop_kernel<0>, dim3((unsigned)n), dim3(64), 0, s, d_states, n);
  else if (width == -64) hipLaunchKernelGGL(permute16_coop_kernel<1>, dim3((unsigned)n), dim3(64), 0, s, d_states, n);
  else return WHIR_EINVAL;
  HIP_TRY(hipGetLastError());
  return WHIR_OK;
}

// ============================ prover ============================

// Status words after a proof log's capacity: words used, overflow flag,
// checkpoint draw (written by final_phase_kernel), failed self-checks
// (whir_prover_set_checks; zeroed when the checks are on).
constexpr uint32_t LOG_TAIL = 4;

struct whir_prover {
  whir_ctx* ctx = nullptr;
  Plan plan;
  uint32_t jobs = 1;  // proofs run in lockstep (gridDim.z) on workspace copies at a fixed stride
  void* arena = nullptr;
  size_t arena_bytes = 0;
  // carved buffers
  DevTx* tx;
  JobScalars* js;
  uint32_t* log;
  uint32_t* input;
  uint32_t* cw0;
  uint32_t* tree0;
  uint32_t* cw[2];
  uint32_t* tree[2];
  uint32_t* ntt_tmp;
  Ext* ep[2];
  Ext* ew[2];
  ClaimSlot* slots;
  Ext* svo_pool;
  Ext* ood_pool;
  ClaimDesc* desc;         // per SVO-phase claim (device copy of hdesc)
  uint32_t* group_start;   // first claim of each point group, n_groups + 1
  Ext* claim_scale;        // per claim: alpha^rank eq(svo_c, rs), applied to its group's eq1 table
  uint32_t* seg_start;     // stacked polys in offset order: their claims (CSR)
  uint32_t* seg_claims;
  Ext* pe_buf;             // svo_kernel scratch
  Ext* gpe_buf;
  Ext* partials;
  uint32_t* stir_idx;
  uint32_t* stir_count;
  QueryArrays qa;
  uint32_t* powL;
  Ext* powR;
  uint64_t svo_stride, ood_stride;
  int m0_svo, m1_svo;
  int n_svo_claims;        // evaluation claims + virtual claims
  // stack layout and the SVO-phase claim table (one poly: a stack of one)
  whir_host::StackPlan stack;
  struct Group {
    uint32_t c0, count, k;
  };
  std::vector<Group> groups;  // evaluation-claim groups, then one per virtual claim
  int n_user_groups = 0;
  std::vector<ClaimDesc> hdesc;
  std::vector<uint32_t> h_group_start, h_seg_start, h_seg_claims;
  uint32_t full0 = 0, n_full = 0;  // claims spanning the whole stacked vector
  StackBlocks sb{};
  uint32_t* host_log = nullptr;  // pinned: the log, LOG_TAIL status words, the transcript (tx_off)
  uint32_t* host_log_dev = nullptr;  // host_log's device-mapped address (nullptr: copy with the DMA engine)
  size_t tx_off = 0;             // byte offset of the DevTx in the log allocation (and in host_log)
  size_t host_log_bytes = 0;     // needed by this shape
  size_t host_log_cap = 0;       // allocated
  bool result_ok = false;        // the last enqueue went through; whir_prover_result may read it
  // commit / eval / open session: 0 idle, 1 committed (next_claim claims done), 2 opened
  int session = 0;
  int next_claim = 0;
  const uint32_t* session_poly = nullptr;
  // optional HIP-event timers around kernel classes (whir_prover_set_timers)
  bool timers = false;
  hipEvent_t ev[WHIR_TIMER_COUNT][2] = {};
  double timer_ms[WHIR_TIMER_COUNT] = {};
  uint64_t timer_launches[WHIR_TIMER_COUNT] = {};
  bool timer_pending = false;
  // stream of the last enqueue; optional hipGraph of the whole proof
  hipStream_t last_stream = nullptr;
  bool graph_mode = false;
  bool checks = false;  // device self-checks (checks.h)
  // >= 0: the base LDE, the leaves and this many wide layers of the base tree
  // come from a set's commit work queue (jobtable.h), not from this sequence
  int ext_commit_levels = -1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipStream_t graph_stream = nullptr;
  hipEvent_t done = nullptr;
};

static inline size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

struct ArenaItem {
  void** ptr;
  size_t bytes;
};

// Fills P's plan-derived fields and the list of device buffers of one proof.
// Claim groups, pool offsets and expression blocks of the SVO phase: the
// evaluation claims (transcript order) then n_ood0 virtual claims on the whole
// stacked vector (write_virtual_eval, prover.rs:231-277: its per-poly sum is
// the stacked vector's claim at the same point).
static int claim_table(whir_prover* P, int n_ood0, uint64_t& pool_ext, uint64_t& part_ext) {
  const whir_host::StackPlan& S = P->stack;
  const int f = P->plan.f, k = S.k;
  const int n_user = (int)S.claim_poly.size();
  const int nc = n_user + n_ood0;
  P->groups.clear();
  for (int c = 0; c < n_user; ++c) {
    const uint32_t kc = (uint32_t)S.k_polys[S.claim_poly[c]];
    if (S.claim_fresh[c]) P->groups.push_back({(uint32_t)c, 1u, kc});
    else P->groups.back().count++;
  }
  P->n_user_groups = (int)P->groups.size();
  for (int v = 0; v < n_ood0; ++v) P->groups.push_back({(uint32_t)(n_user + v), 1u, (uint32_t)k});
  P->hdesc.assign((size_t)nc, ClaimDesc{});
  P->h_group_start.clear();
  pool_ext = 0;
  part_ext = 0;
  for (const auto& G : P->groups) {
    P->h_group_start.push_back(G.c0);
    const uint32_t m = G.k - (uint32_t)f, m0 = m / 2, m1 = m - m / 2;
    part_ext = std::max<uint64_t>(part_ext, (uint64_t)G.count << (m1 + f));
    for (uint32_t j = 0; j < G.count; ++j) {
      const uint32_t c = G.c0 + j;
      ClaimDesc& d = P->hdesc[c];
      d.m0 = m0;
      d.m1 = m1;
      if (j == 0) {
        d.eq0_off = pool_ext;
        d.eq1_off = pool_ext + (1ull << m0);
        pool_ext += (1ull << m0) + (1ull << m1);
      } else {
        d.eq0_off = P->hdesc[G.c0].eq0_off;
        d.eq1_off = pool_ext;
        pool_ext += 1ull << m1;
      }
      d.lead_eq1 = P->hdesc[G.c0].eq1_off;
      d.poly_off = (int)c < n_user ? S.offset[S.claim_poly[c]] : 0;
      d.rank = (int)c < n_user ? S.rank[c] : c;
    }
  }
  P->h_group_start.push_back((uint32_t)nc);
  if (pool_ext >> 32) return WHIR_EINVAL;  // svo_kernel keeps pool offsets in 32 bits
  // whole-vector claims: the virtual ones, and every claim of a one-poly stack
  const bool one = S.k_polys.size() == 1;
  P->full0 = one ? 0u : (uint32_t)n_user;
  P->n_full = (uint32_t)nc - P->full0;
  // blocks: polys in offset order, grouped into runs of equal size
  P->sb = StackBlocks{};
  P->h_seg_start.assign(1, 0);
  P->h_seg_claims.clear();
  if (!one) {
    std::vector<int> by_off(S.k_polys.size());
    for (size_t i = 0; i < by_off.size(); ++i) by_off[i] = (int)i;
    std::sort(by_off.begin(), by_off.end(), [&](int a, int b) { return S.offset[a] < S.offset[b]; });
    for (size_t sgi = 0; sgi < by_off.size(); ++sgi) {
      const int id = by_off[sgi];
      const uint32_t lg = (uint32_t)(S.k_polys[id] - f);
      if (P->sb.n_class == 0 || P->sb.log[P->sb.n_class - 1] != lg) {
        if (P->sb.n_class == (uint32_t)MAX_CLASSES) return WHIR_EINVAL;
        P->sb.log[P->sb.n_class] = lg;
        P->sb.seg0[P->sb.n_class] = (uint32_t)sgi;
        P->sb.start[P->sb.n_class] = S.offset[id] >> f;
        P->sb.n_class++;
      }
      P->sb.start[P->sb.n_class] = (S.offset[id] >> f) + (1ull << lg);
      for (int c = 0; c < n_user; ++c)
        if (S.claim_poly[c] == id) P->h_seg_claims.push_back((uint32_t)c);
      P->h_seg_start.push_back((uint32_t)P->h_seg_claims.size());
    }
  }
  if (P->h_seg_claims.empty()) P->h_seg_claims.push_back(0);
  return WHIR_OK;
}

static int prover_layout(whir_prover* P, const whir_params* prm, const whir_stack* st,
                         std::vector<ArenaItem>& items) {
  // one poly unless a stack is given: n_points claims, each at a fresh point (test.rs:223-227)
  whir_host::StackPlan S;
  if (st) {
    if (!whir_host::make_stack_plan(st->k_polys, st->n_polys, st->claim_poly, st->claim_fresh, st->n_claims,
                                    prm->folding, S))
      return WHIR_EINVAL;
  } else {
    if (prm->n_points < 0 || prm->n_points > MAX_CLAIMS) return WHIR_EINVAL;
    const int32_t kk = prm->k;
    std::vector<int32_t> cp((size_t)prm->n_points, 0);
    std::vector<uint8_t> cf((size_t)prm->n_points, 1);
    if (!whir_host::make_stack_plan(&kk, 1, cp.data(), cf.data(), prm->n_points, prm->folding, S))
      return WHIR_EINVAL;
  }
  if (S.k != prm->k) return WHIR_EINVAL;
  Plan pl = whir_host::make_plan(prm->k, prm->folding, prm->rate, prm->initial_reduction, prm->soundness,
                                 prm->security_level, (int)S.claim_poly.size());
  if (!pl.ok) return WHIR_EINVAL;
  const int k = pl.k, f = pl.f;
  if (k > 28 || f > MAX_F || pl.k + pl.rate - f > 24) return WHIR_EINVAL;
  const int n_svo = pl.n_points + pl.n_ood0;
  if (n_svo < 1 || n_svo > MAX_CLAIMS) return WHIR_EINVAL;
  int max_ood = 0, max_stir = pl.final_n_stir;
  uint64_t max_cw_words = 0;
  for (const RoundPlan& r : pl.rounds) {
    max_ood = std::max(max_ood, r.n_ood);
    max_stir = std::max(max_stir, r.n_stir);
    if (r.log_rows > 24) return WHIR_EINVAL;
    max_cw_words = std::max<uint64_t>(max_cw_words, (4ull << f) << r.log_rows);
  }
  if (max_stir > MAX_STIR || max_ood > 16) return WHIR_EINVAL;
  P->plan = pl;
  P->n_svo_claims = n_svo;
  P->stack = std::move(S);
  uint64_t pool_ext = 0, group_part_ext = 0;
  if (int rc = claim_table(P, pl.n_ood0, pool_ext, group_part_ext)) return rc;
  const int m = k - f;
  P->m0_svo = m / 2;
  P->m1_svo = m - m / 2;
  P->svo_stride = (1ull << P->m0_svo) + (1ull << P->m1_svo);
  P->ood_stride = P->svo_stride;  // round polynomials have at most k - f variables
  const uint64_t H0 = 1ull << (k + pl.rate - f);
  const uint64_t cw0_words = H0 << f;
  uint64_t max_rows = 1;
  for (const RoundPlan& r : pl.rounds) max_rows = std::max<uint64_t>(max_rows, 1ull << r.log_rows);
  const uint64_t n_ext = 1ull << m;
  const int m_pow = std::max(0, k - 2 * f);
  const uint64_t pow_l = 1ull << (m_pow / 2), pow_r = 1ull << (m_pow - m_pow / 2);
  const uint64_t part_ext = std::max<uint64_t>({group_part_ext, 1ull << P->m1_svo, 4096});
  constexpr uint64_t PE_PER = (2 << MAX_F) - 2;
  // the transcript state sits right after the log and its status words, so
  // one copy brings back the proof and the transcript that open hands back
  P->tx_off = ((pl.log_words + LOG_TAIL) * 4 + 255) & ~(size_t)255;
  P->host_log_bytes = P->tx_off + sizeof(DevTx);
  items = {
      {(void**)&P->js, sizeof(JobScalars)},
      {(void**)&P->log, P->host_log_bytes},
      {(void**)&P->input, (1ull << k) * 4},
      {(void**)&P->cw0, cw0_words * 4},
      {(void**)&P->tree0, (2 * H0 - 1) * 32},
      {(void**)&P->cw[0], max_cw_words * 4},
      {(void**)&P->cw[1], max_cw_words * 4},
      {(void**)&P->tree[0], (2 * max_rows - 1) * 32},
      {(void**)&P->tree[1], (2 * max_rows - 1) * 32},
      {(void**)&P->ntt_tmp, std::max(cw0_words, max_cw_words) * 4},
      {(void**)&P->ep[0], n_ext * 16},
      {(void**)&P->ep[1], n_ext * 16},
      {(void**)&P->ew[0], n_ext * 16},
      {(void**)&P->ew[1], n_ext * 16},
      {(void**)&P->slots, sizeof(ClaimSlot) * (n_svo + 16)},
      {(void**)&P->svo_pool, pool_ext * 16},
      {(void**)&P->desc, sizeof(ClaimDesc) * n_svo},
      {(void**)&P->group_start, P->h_group_start.size() * 4},
      {(void**)&P->claim_scale, sizeof(Ext) * n_svo},
      {(void**)&P->seg_start, P->h_seg_start.size() * 4},
      {(void**)&P->seg_claims, P->h_seg_claims.size() * 4},
      {(void**)&P->pe_buf, PE_PER * n_svo * 16},
      {(void**)&P->gpe_buf, PE_PER * P->groups.size() * 16},
      {(void**)&P->ood_pool, P->ood_stride * std::max(1, max_ood) * 16},
      {(void**)&P->partials, part_ext * 16},
      {(void**)&P->stir_idx, MAX_STIR * 4},
      {(void**)&P->stir_count, 4},
      {(void**)&P->qa.index, MAX_STIR * 4},
      {(void**)&P->qa.ws, MAX_STIR * 4},
      {(void**)&P->qa.ts, MAX_STIR * 4},
      {(void**)&P->qa.var, MAX_STIR * 4},
      {(void**)&P->qa.scale, MAX_STIR * 16},
      {(void**)&P->qa.G, (size_t)MAX_STIR * (1u << MAX_F) * 16},
      {(void**)&P->powL, MAX_STIR * pow_l * 4},
      {(void**)&P->powR, MAX_STIR * pow_r * 16},
  };
  return WHIR_OK;
}

static size_t arena_size(const std::vector<ArenaItem>& items) {
  size_t total = 0;
  for (auto& it : items) total += align_up(it.bytes);
  return total;
}

// Device bytes one proof of this shape needs (0 if the shape is invalid).
extern "C" size_t whir_prover_bytes_for(const whir_params* prm) {
  whir_prover tmp;
  std::vector<ArenaItem> items;
  if (!prm || prover_layout(&tmp, prm, nullptr, items)) return 0;
  return arena_size(items);
}

// Carves P's buffers (laid out in `items`) out of P->arena and uploads the
// shape's static claim tables.
static int prover_bind(whir_prover* P, const std::vector<ArenaItem>& items) {
  char* p = (char*)P->arena;
  for (auto& it : items) {
    *it.ptr = p;
    p += align_up(it.bytes);
  }
  P->tx = (DevTx*)((char*)P->log + P->tx_off);
  void* host_dev = nullptr;
  P->host_log_dev = P->host_log && hipHostGetDevicePointer(&host_dev, P->host_log, 0) == hipSuccess
                        ? (uint32_t*)host_dev : nullptr;
  (void)hipGetLastError();
  const struct {
    void* dst;
    const void* src;
    size_t bytes;
  } up[] = {
      {P->desc, P->hdesc.data(), P->hdesc.size() * sizeof(ClaimDesc)},
      {P->group_start, P->h_group_start.data(), P->h_group_start.size() * 4},
      {P->seg_start, P->h_seg_start.data(), P->h_seg_start.size() * 4},
      {P->seg_claims, P->h_seg_claims.data(), P->h_seg_claims.size() * 4},
  };
  for (const auto& u : up)
    if (hipMemcpy(u.dst, u.src, u.bytes, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipGetLastError();
      return WHIR_EDEVICE;
    }
  P->sb.seg_start = P->seg_start;
  P->sb.seg_claims = P->seg_claims;
  return WHIR_OK;
}

static int prover_create(whir_ctx* c, const whir_params* prm, const whir_stack* st, whir_prover** out) {
  if (!c || !prm || !out) return WHIR_EINVAL;
  whir_prover* P = new whir_prover();
  P->ctx = c;
  std::vector<ArenaItem> items;
  int rc = prover_layout(P, prm, st, items);
  if (rc) {
    delete P;
    return rc;
  }
  const size_t total = arena_size(items);
  if (hipMalloc(&P->arena, total) != hipSuccess) {
    (void)hipGetLastError();
    delete P;
    return WHIR_EOOM;
  }
  P->arena_bytes = total;
  if (hipHostMalloc((void**)&P->host_log, P->host_log_bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    whir_prover_destroy(P);
    return WHIR_EOOM;
  }
  P->host_log_cap = P->host_log_bytes;
  if ((rc = prover_bind(P, items))) {
    whir_prover_destroy(P);
    return rc;
  }
  *out = P;
  return WHIR_OK;
}

// Re-lays out the idle one-poly workspace P for the proof shape `prm` inside
// P's own arena: no device memory is allocated or released (a release waits
// for the whole device). P's pinned host log is kept when it is large enough;
// otherwise a new one is allocated and the old one goes to `host_to_free`.
// Returns the new workspace, P having been consumed, or nullptr with P
// unchanged when the shape does not fit.
static whir_prover* prover_relayout(whir_prover* P, const whir_params* prm, std::vector<void*>& host_to_free) {
  whir_prover* Q = new whir_prover();
  Q->ctx = P->ctx;
  std::vector<ArenaItem> items;
  if (prover_layout(Q, prm, nullptr, items) || arena_size(items) > P->arena_bytes) {
    delete Q;
    return nullptr;
  }
  Q->arena = P->arena;
  Q->arena_bytes = P->arena_bytes;
  if (P->host_log_cap >= Q->host_log_bytes) {
    Q->host_log = P->host_log;
    Q->host_log_cap = P->host_log_cap;
  } else {
    if (hipHostMalloc((void**)&Q->host_log, Q->host_log_bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      delete Q;
      return nullptr;
    }
    Q->host_log_cap = Q->host_log_bytes;
  }
  if (prover_bind(Q, items)) {
    if (Q->host_log != P->host_log) host_to_free.push_back(Q->host_log);
    delete Q;
    return nullptr;
  }
  if (Q->host_log != P->host_log) host_to_free.push_back(P->host_log);
  // P's remaining resources: its graph (captured for the old shape), events
  if (P->gexec) (void)hipGraphExecDestroy(P->gexec);
  if (P->graph) (void)hipGraphDestroy(P->graph);
  for (int i = 0; i < WHIR_TIMER_COUNT; ++i)
    for (int j = 0; j < 2; ++j)
      if (P->ev[i][j]) (void)hipEventDestroy(P->ev[i][j]);
  Q->done = P->done;
  delete P;
  return Q;
}

// ============================ autotuner ============================
// At workspace creation (whir_ctx_set_autotune), every transform and tree of
// the proof shape that the context has not tuned yet is run in each candidate
// configuration on the workspace's own buffers; the fastest (minimum of three
// timed runs, HIP events on the context's stream) is used from then on by every
// launch of that size. Candidates differ only in how the work is laid out, so
// every one gives the same words (tests/test_gpu_autotune.py).

static double time_min_ms(hipStream_t s, const std::function<int()>& run, int reps = 3) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    (void)hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    return -1.0;
  }
  double best = -1.0;
  if (run() == WHIR_OK) {  // warm-up
    for (int r = 0; r < reps; ++r) {
      float ms = 0.f;
      if (hipEventRecord(e0, s) != hipSuccess || run() != WHIR_OK || hipEventRecord(e1, s) != hipSuccess ||
          hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        (void)hipGetLastError();
        best = -1.0;
        break;
      }
      if (best < 0 || ms < best) best = ms;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

// One candidate per distinct launch sequence (plan, kernels, workgroup shape).
static std::vector<NttConfig> ntt_candidates(whir_ctx* c, uint64_t n_in, int log_nn, int log_c) {
  std::vector<NttConfig> out;
  std::vector<std::string> seen;
  for (int two : {0, 1, 2})
    for (int fl : {4, 3, 0})
      for (int lt : {8, 7, 9})  // 9: 512-thread three-stage passes (others are capped at 256 threads)
        for (int lv : {2, 1}) {
          NttConfig cfg;
          cfg.two_pass = two;
          cfg.first_log = fl;
          cfg.log_threads = lt;
          cfg.max_lv = lv;
          NttPlan pl;
          if (!ntt_plan_for(cfg, n_in, log_nn, log_c, pl)) continue;
          std::string sig;
          bool ok = true;
          for (int p = 0; p < pl.P; ++p) {
            NttPass ps;
            ntt_pass_setup(pl, p, n_in, cfg.log_threads, nullptr, nullptr, nullptr, c->tw_lo(), c->tw_hi(), ps);
            ok &= ps.lds <= 65536;
            sig += std::to_string(ps.kind) + "," + std::to_string(pl.logl[p]) + "," + std::to_string(pl.lv) + "," +
                   std::to_string(ps.threads) + ";";
          }
          if (!ok || std::find(seen.begin(), seen.end(), sig) != seen.end()) continue;
          seen.push_back(sig);
          out.push_back(cfg);
        }
  return out;
}

static std::string merkle_signature(int log_h, uint32_t row_words, const MerkleConfig& mc) {
  const uint64_t H = 1ull << log_h;
  std::string sig = (H <= mc.pair_max && row_words % 16 == 0) ? "P" : "L";
  uint64_t n = H;
  while (n > std::max<uint64_t>(mc.coop_from, COOP_MAX_NODES)) {
    sig += (n / 2 <= mc.pair_max) ? 'p' : 'l';
    n /= 2;
  }
  return sig + "|" + std::to_string(n);
}

static std::vector<MerkleConfig> merkle_candidates(int log_h, uint32_t row_words, uint32_t jobs = 1) {
  std::vector<MerkleConfig> out;
  std::vector<std::string> seen;
#if defined(WHIR_FIELD_BABYBEAR)
  std::vector<int> pair_logs = {-1};  // no two-lane kernels in this build
#else
  std::vector<int> pair_logs = {15, 14, 16, 13};
#endif
  std::vector<int> coop_logs = {15, 13, 14, 16, 17};
  if (jobs > 1) {  // a batch has jobs x the lanes per level: cheaper wide levels further up
    coop_logs.insert(coop_logs.end(), {12, 11, 10});
#if !defined(WHIR_FIELD_BABYBEAR)
    pair_logs.insert(pair_logs.end(), {12, 10});
#endif
  }
  for (int pm : pair_logs)
    for (int cf : coop_logs) {
      MerkleConfig mc;
      mc.pair_max = pm < 0 ? 0 : 1ull << pm;
      mc.coop_from = 1ull << cf;
      const std::string sig = merkle_signature(log_h, row_words, mc);
      if (std::find(seen.begin(), seen.end(), sig) != seen.end()) continue;
      seen.push_back(sig);
      out.push_back(mc);
    }
  return out;
}

// Transforms and trees below these sizes are launch-bound: not tuned.
constexpr uint64_t TUNE_MIN_NTT_WORDS = 1ull << 16;
constexpr int TUNE_MIN_TREE_LOG = 12;

static void tune_ntt(whir_ctx* c, const uint32_t* in, uint64_t n_in, int log_nn, int log_c, uint32_t* out,
                     uint32_t* tmp, uint32_t jobs) {
  const auto key = std::make_tuple(n_in, log_nn, log_c, jobs_class(jobs));
  {
    std::lock_guard<std::mutex> g(c->tune_mu);
    if (c->ntt_tuned.count(key)) return;
  }
  if (((uint64_t)1 << (log_nn + log_c)) < TUNE_MIN_NTT_WORDS) return;
  const std::vector<NttConfig> cand = ntt_candidates(c, n_in, log_nn, log_c);
  if (cand.size() < 2) return;
  NttConfig best = c->ntt;
  double best_ms = -1.0, default_ms = -1.0;
  for (const NttConfig& cfg : cand) {
    const double ms =
        time_min_ms(c->stream, [&] { return ntt_enqueue(c, in, n_in, log_nn, log_c, out, tmp, c->stream, &cfg, jobs); });
    if (ms < 0) continue;
    if (cfg.two_pass == c->ntt.two_pass && cfg.first_log == c->ntt.first_log && cfg.log_threads == c->ntt.log_threads &&
        cfg.max_lv == c->ntt.max_lv)
      default_ms = ms;
    if (best_ms < 0 || ms < best_ms) {
      best_ms = ms;
      best = cfg;
    }
  }
  if (best_ms < 0) return;
  char buf[320];
  snprintf(buf, sizeof(buf),
           "{\"kind\": \"ntt\", \"n_in\": %llu, \"log_rows\": %d, \"log_cols\": %d, \"jobs\": %u, "
           "\"candidates\": %zu, \"two_pass\": %d, \"first_pass_log\": %d, \"log_threads\": %d, \"log_vec\": %d, "
           "\"best_us\": %.1f, \"default_us\": %.1f}",
           (unsigned long long)n_in, log_nn, log_c, jobs_class(jobs), cand.size(), best.two_pass, best.first_log,
           best.log_threads, best.max_lv, best_ms * 1e3, default_ms * 1e3);
  std::lock_guard<std::mutex> g(c->tune_mu);
  c->ntt_tuned[key] = best;
  c->tune_log.push_back(buf);
}

static void tune_merkle(whir_ctx* c, const uint32_t* rows, int log_h, uint32_t row_words, uint32_t* layers,
                        uint32_t jobs) {
  const auto key = std::make_tuple(log_h, row_words, jobs_class(jobs));
  {
    std::lock_guard<std::mutex> g(c->tune_mu);
    if (c->merkle_tuned.count(key)) return;
  }
  if (log_h < TUNE_MIN_TREE_LOG) return;
  const std::vector<MerkleConfig> cand = merkle_candidates(log_h, row_words, jobs);
  if (cand.size() < 2) return;
  MerkleConfig best = c->merkle;
  double best_ms = -1.0, default_ms = -1.0;
  for (const MerkleConfig& mc : cand) {
    const double ms = time_min_ms(c->stream, [&] {
      return merkle_enqueue(rows, log_h, row_words, layers, c->stream, mc, nullptr, nullptr, nullptr, jobs);
    });
    if (ms < 0) continue;
    if (mc.pair_max == c->merkle.pair_max && mc.coop_from == c->merkle.coop_from) default_ms = ms;
    if (best_ms < 0 || ms < best_ms) {
      best_ms = ms;
      best = mc;
    }
  }
  if (best_ms < 0) return;
  int pm = 0, cf = 0;
  while ((1ull << pm) < best.pair_max) ++pm;
  while ((1ull << cf) < best.coop_from) ++cf;
  char buf[320];
  snprintf(buf, sizeof(buf),
           "{\"kind\": \"merkle\", \"log_rows\": %d, \"row_words\": %u, \"jobs\": %u, \"candidates\": %zu, "
           "\"pair_max_log\": %d, \"coop_from_log\": %d, \"best_us\": %.1f, \"default_us\": %.1f}",
           log_h, row_words, jobs_class(jobs), cand.size(), pm, cf, best_ms * 1e3, default_ms * 1e3);
  std::lock_guard<std::mutex> g(c->tune_mu);
  c->merkle_tuned[key] = best;
  c->tune_log.push_back(buf);
}

// Every transform and tree of P's proof shape, timed on P's buffers (their
// contents are scratch until a proof runs).
static void prover_autotune(whir_prover* P) {
  whir_ctx* c = P->ctx;
  const Plan& pl = P->plan;
  const int k = pl.k, f = pl.f;
  const uint32_t W = 1u << f;
  const int log_h0 = k + pl.rate - f;
  tune_ntt(c, P->input, 1ull << (k - f), log_h0, f, P->cw0, P->ntt_tmp, P->jobs);
  tune_merkle(c, P->cw0, log_h0, W, P->tree0, P->jobs);
  for (const RoundPlan& rp : pl.rounds) {
    tune_ntt(c, (const uint32_t*)P->ep[0], 1ull << (rp.kc - f), rp.log_rows, f + 2, P->cw[0], P->ntt_tmp, P->jobs);
    tune_merkle(c, P->cw[0], rp.log_rows, 4 * W, P->tree[0], P->jobs);
  }
  (void)hipGetLastError();
}

extern "C" int whir_ctx_tuning(whir_ctx* c, char* buf, size_t cap, size_t* len) {
  if (!c) return WHIR_EINVAL;
  std::string js = "[";
  {
    std::lock_guard<std::mutex> g(c->tune_mu);
    for (size_t i = 0; i < c->tune_log.size(); ++i) js += (i ? ", " : "") + c->tune_log[i];
  }
  js += "]";
  if (len) *len = js.size();
  if (!buf || cap < js.size() + 1) return buf ? WHIR_ETRANSCRIPT : WHIR_OK;
  memcpy(buf, js.c_str(), js.size() + 1);
  return WHIR_OK;
}

extern "C" int whir_prover_create(whir_ctx* c, const whir_params* prm, whir_prover** out) {
  int rc = prover_create(c, prm, nullptr, out);
  if (rc == WHIR_OK && c->autotune) prover_autotune(*out);
  return rc;
}

extern "C" int whir_prover_create_stacked(whir_ctx* c, const whir_params* prm, const whir_stack* st,
                                          whir_prover** out) {
  if (!st) return WHIR_EINVAL;
  int rc = prover_create(c, prm, st, out);
  if (rc == WHIR_OK && c->autotune) prover_autotune(*out);
  return rc;
}

extern "C" int64_t whir_prover_stack_offset(const whir_prover* P, int32_t id) {
  if (!P || id < 0 || id >= (int32_t)P->stack.k_polys.size()) return -1;
  return (int64_t)P->stack.offset[id];
}

extern "C" int whir_prover_destroy(whir_prover* P) {
  if (!P) return WHIR_OK;
  int rc = WHIR_OK;
  if (P->arena) release_step(rc, hipFree(P->arena));
  if (P->host_log) release_step(rc, hipHostFree(P->host_log));
  for (int i = 0; i < WHIR_TIMER_COUNT; ++i)
    for (int j = 0; j < 2; ++j)
      if (P->ev[i][j]) release_step(rc, hipEventDestroy(P->ev[i][j]));
  if (P->gexec) release_step(rc, hipGraphExecDestroy(P->gexec));
  if (P->graph) release_step(rc, hipGraphDestroy(P->graph));
  if (P->done) release_step(rc, hipEventDestroy(P->done));
  delete P;
  return rc;
}

extern "C" int whir_prover_set_timers(whir_prover* P, int enable) {
  if (!P) return WHIR_EINVAL;
  if (enable && !P->ev[0][0])
    for (int i = 0; i < WHIR_TIMER_COUNT; ++i)
      for (int j = 0; j < 2; ++j) HIP_TRY(hipEventCreate(&P->ev[i][j]));
  P->timers = enable != 0;
  for (int i = 0; i < WHIR_TIMER_COUNT; ++i) {
    P->timer_ms[i] = 0.0;
    P->timer_launches[i] = 0;
  }
  return WHIR_OK;
}

extern "C" int whir_prover_timer(const whir_prover* P, int which, double* total_ms, uint64_t* launches,
                                 uint64_t* work_units) {
  if (!P || which < 0 || which >= WHIR_TIMER_COUNT) return WHIR_EINVAL;
  if (total_ms) *total_ms = P->timer_ms[which];
  if (launches) *launches = P->timer_launches[which];
  if (work_units) {
    const Plan& pl = P->plan;
    const uint64_t rows = 1ull << (pl.k + pl.rate - pl.f);
    // leaf hash: Poseidon2-24 permutations per launch; DFT: input + output words
    *work_units = which == WHIR_TIMER_LEAVES_BASE ? rows * (((1ull << pl.f) + 15) / 16)
                                                  : (1ull << pl.k) + (rows << pl.f);
  }
  return WHIR_OK;
}

extern "C" size_t whir_prover_proof_capacity(const whir_prover* P) { return P ? P->plan.log_words * 4 : 0; }
extern "C" uint32_t* whir_prover_input(whir_prover* P) { return P ? P->input : nullptr; }

static unsigned blocks_for(uint64_t n, unsigned per = SC_THREADS) { return (unsigned)((n + per - 1) / per); }
// A launch grid over the prover's jobs (gridDim.z).
static dim3 zg(dim3 g, uint32_t jobs) {
  g.z = jobs;
  return g;
}

template <int FL>
static void launch_compress_hi(dim3 grid, hipStream_t s, const uint32_t* poly, const Ext* eq0, const Ext* eq1,
                               uint32_t m0, Ext* partials, const ClaimDesc* desc, uint32_t c0, uint32_t jobs) {
  hipLaunchKernelGGL(compress_hi_kernel<FL>, zg(grid, jobs), dim3(SC_THREADS), 0, s, poly, eq0, eq1, m0, partials, desc, c0);
}
template <int FL>
static void launch_expr(unsigned grid, hipStream_t s, const uint32_t* poly, const Ext* eq_rs, const whir_prover* P,
                        uint32_t m0, uint64_t n, Ext* ep, Ext* ew) {
  hipLaunchKernelGGL(expr_build_kernel<FL>, zg(dim3(grid), P->jobs), dim3(SC_THREADS), 0, s, poly, eq_rs, (const Ext*)P->svo_pool,
                     (const ClaimDesc*)P->desc, P->full0, P->n_full, m0, n, ep, ew, P->sb,
                     (const Ext*)P->claim_scale);
}

#define FSWITCH(f, CALL)          \
  switch (f) {                    \
    case 1: { constexpr int FL = 1; CALL; } break; \
    case 2: { constexpr int FL = 2; CALL; } break; \
    case 3: { constexpr int FL = 3; CALL; } break; \
    case 4: { constexpr int FL = 4; CALL; } break; \
    case 5: { constexpr int FL = 5; CALL; } break; \
    default: return WHIR_EINVAL;  \
  }

// Claim group g's begin (its point and eq tables); an empty begin past the
// last group.
static ClaimBegin group_begin(const whir_prover* P, int g) {
  ClaimBegin cb{};
  if (g >= 0 && g < (int)P->groups.size()) {
    const whir_prover::Group& G = P->groups[g];
    const ClaimDesc& d = P->hdesc[G.c0];
    cb.slot = P->slots + G.c0;
    cb.count = G.count;
    cb.eq0 = P->svo_pool + d.eq0_off;
    cb.eq1 = P->svo_pool + d.eq1_off;
    cb.k = G.k;
    cb.fsv = (uint32_t)P->plan.f;
    cb.m0 = d.m0;
    cb.m1 = d.m1;
  }
  return cb;
}

// ---- commit: RS encode + Merkle + transcript.write(root) (whir.rs:19-36,
// 127-138). `first` is the claim begin that runs in the root's kernel (the
// fused proof draws the first claim point there); `import` continues a
// caller's transcript instead of starting a fresh one.
static int enqueue_commit(whir_prover* P, const uint32_t* poly, hipStream_t s, const ClaimBegin& first,
                          const DevTx* import) {
  whir_ctx* c = P->ctx;
  const Plan& pl = P->plan;
  const int k = pl.k, f = pl.f;
  const uint32_t W = 1u << f;
  int rc;
  if (import) hipLaunchKernelGGL(tx_import_kernel, zg(dim3(1), P->jobs), dim3(64), 0, s, P->tx, P->js, *import);
  const int log_h0 = k + pl.rate - f;
  const bool ext = P->ext_commit_levels >= 0;
  const bool timers = P->timers && !ext;
  if (timers) HIP_TRY(hipEventRecord(P->ev[WHIR_TIMER_DFT_BASE][0], s));
  if (!ext && (rc = ntt_enqueue(c, poly, 1ull << (k - f), log_h0, f, P->cw0, P->ntt_tmp, s, nullptr, P->jobs)))
    return rc;
  if (timers) HIP_TRY(hipEventRecord(P->ev[WHIR_TIMER_DFT_BASE][1], s));
  uint32_t* top = nullptr;
  uint32_t top_n = 0;
  if ((rc = merkle_enqueue(P->cw0, log_h0, W, P->tree0, s, merkle_config(c, log_h0, W, P->jobs),
                           timers ? P->ev[WHIR_TIMER_LEAVES_BASE] : nullptr, &top, &top_n, P->jobs,
                           P->ext_commit_levels)))
    return rc;
  P->timer_pending = timers;
  // the tree's top levels, transcript.write(root) and the first group's draw
  hipLaunchKernelGGL(tree_top_claim_kernel, zg(dim3(1), P->jobs), dim3(SERIAL_THREADS), 0, s, top, top_n, P->tx, P->log, P->js,
                     first, import ? 0u : (uint32_t)pl.log_words);
  return WHIR_OK;
}

// ---- evaluation claims + OOD virtual claims (prover.rs:64-90, 231-277) ----
// Claims at one point form a group (one compress launch over its polys).
// Groups g0..g1-1, each begun already; each group's end runs in the same
// kernel as the next group's begin (groups up to g1 - 1).
static int enqueue_groups(whir_prover* P, const uint32_t* poly, hipStream_t s, int g0, int g1) {
  const int f = P->plan.f;
  for (int g = g0; g < g1; ++g) {
    const ClaimBegin cb = group_begin(P, g);
    const whir_prover::Group& G = P->groups[g];
    FSWITCH(f, launch_compress_hi<FL>(dim3(1u << cb.m1, G.count), s, poly, cb.eq0, cb.eq1, cb.m0, P->partials,
                                      P->desc, G.c0, P->jobs));
    const ClaimBegin next = g + 1 < g1 ? group_begin(P, g + 1) : ClaimBegin{};
    hipLaunchKernelGGL(claim_end_kernel, zg(dim3(1), P->jobs), dim3(SERIAL_THREADS), 0, s, P->tx, P->log, P->js, cb.slot, G.count,
                       P->partials, 1u << cb.m1, (uint32_t)f, next);
  }
  return WHIR_OK;
}

static int enqueue_open(whir_prover* P, const uint32_t* poly, hipStream_t s);

// The whole proof as one kernel sequence on stream s (no host synchronisation):
// the test driver's commit, n claims at drawn points and open
// (crates/whir/src/test.rs:218-235).
static int prover_enqueue(whir_prover* P, const uint32_t* poly, hipStream_t s) {
  int rc;
  if ((rc = enqueue_commit(P, poly, s, group_begin(P, 0), nullptr))) return rc;
  if ((rc = enqueue_groups(P, poly, s, 0, (int)P->groups.size()))) return rc;
  return enqueue_open(P, poly, s);
}

// Whir::open once every claim group has been evaluated (whir.rs:168-359): SVO
// rounds, expression, WHIR rounds, final phase.
static int enqueue_open(whir_prover* P, const uint32_t* poly, hipStream_t s) {
  whir_ctx* c = P->ctx;
  const Plan& pl = P->plan;
  const int k = pl.k, f = pl.f;
  const uint32_t W = 1u << f;
  int rc;
  const int n_groups = (int)P->groups.size();
  const uint32_t m0 = P->m0_svo;
  uint32_t* fail = P->log + P->plan.log_words + 3;  // failed self-checks (tail word 3)
  if (P->checks) hipLaunchKernelGGL(check_reset_kernel, zg(dim3(1), P->jobs), dim3(64), 0, s, fail);
  // the sumcheck self-check of the rounds on the grid after round 0 (round 0's
  // weights have query terms that live in the per-query tables, not in ew)
  auto sumcheck_check = [&](int cur_, uint64_t n_, const Ext* r_) {
    hipLaunchKernelGGL(sumcheck_check_kernel, zg(dim3(1), P->jobs), dim3(CHECK_THREADS), 0, s, P->ep[cur_],
                       P->ew[cur_], n_, r_, P->js, fail);
  };
  // ---- SVO rounds + expression (prover.rs:279-353, 184-229) ----
  hipLaunchKernelGGL(svo_kernel, zg(dim3(1), P->jobs), dim3(SERIAL_THREADS), 0, s, P->tx, P->log, P->js, P->slots, P->desc,
                     (uint32_t)P->n_svo_claims, P->group_start, (uint32_t)n_groups, (uint32_t)f, P->pe_buf,
                     P->gpe_buf, P->claim_scale);
  uint64_t n = 1ull << (k - f);
  FSWITCH(f, launch_expr<FL>(blocks_for(n), s, poly, P->js->eq_rs, P, m0, n, P->ep[0], P->ew[0]));
  int cur = 0;
  uint32_t rs_idx = (uint32_t)f;
  uint32_t* top = nullptr;
  uint32_t top_n = 0;

  // ---- WHIR rounds (whir.rs:205-312) ----
  for (int r = 0; r < pl.n_rounds; ++r) {
    const RoundPlan& rp = pl.rounds[r];
    uint32_t* cw = P->cw[r & 1];
    uint32_t* tree = P->tree[r & 1];
    // commit_ext (whir.rs:38-61): 2^(kc-f) rows of 2^f extension elements
    if ((rc = ntt_enqueue(c, (const uint32_t*)P->ep[cur], 1ull << (rp.kc - f), rp.log_rows, f + 2, cw, P->ntt_tmp, s,
                          nullptr, P->jobs)))
      return rc;
    const bool small = r > 0 && small_round_fits(rp.kc, rp.n_ood, rp.n_stir);
    if ((rc = merkle_enqueue(cw, rp.log_rows, 4 * W, tree, s, merkle_config(c, rp.log_rows, 4 * W, P->jobs), nullptr,
                             small ? nullptr : &top, &top_n, P->jobs)))
      return rc;
    const uint64_t Hr = 1ull << rp.log_rows;
    const uint32_t* root = tree + (2 * Hr - 2) * 8;
    if (small) {
      // the whole round in one workgroup (round_small.h)
      SmallRoundArgs sa{root, P->ep[cur], P->ew[cur], P->ep[cur ^ 1], P->ew[cur ^ 1], (uint32_t)rp.kc,
                        (uint32_t)rp.n_ood, (uint32_t)rp.n_stir, (uint32_t)rp.kfd, (uint32_t)f, rs_idx, 4 * W,
                        P->cw[(r - 1) & 1], P->tree[(r - 1) & 1]};
      hipLaunchKernelGGL(small_round_kernel, zg(dim3(1), P->jobs), dim3(SR_THREADS), 0, s, P->tx, P->log, P->js, sa);
      n >>= f;
      cur ^= f & 1;
      rs_idx += f;
      continue;
    }
    // OOD claims (whir.rs:224-235), the first one drawn in the root's kernel
    const uint32_t om0 = rp.kc / 2, om1 = rp.kc - rp.kc / 2;
    auto ood_claim = [&](int o) {
      ClaimBegin cb{};
      if (o < rp.n_ood) {
        cb.slot = P->slots + P->n_svo_claims + o;
        cb.eq0 = P->ood_pool + o * P->ood_stride;
        cb.eq1 = cb.eq0 + (1ull << om0);
        cb.k = (uint32_t)rp.kc;
        cb.fsv = 0;
        cb.m0 = om0;
        cb.m1 = om1;
        cb.count = 1;
      }
      return cb;
    };
    // the tree's top levels, transcript.write(root) and the first OOD draw
    hipLaunchKernelGGL(tree_top_claim_kernel, zg(dim3(1), P->jobs), dim3(SERIAL_THREADS), 0, s, top, top_n, P->tx, P->log, P->js,
                       rp.n_ood > 0 ? ood_claim(0) : ClaimBegin{}, 0u);
    for (int o = 0; o < rp.n_ood; ++o) {
      const ClaimBegin cb = ood_claim(o);
      hipLaunchKernelGGL(eval_ext_kernel, zg(dim3(1u << om1), P->jobs), dim3(SC_THREADS), 0, s, P->ep[cur], cb.eq0, cb.eq1, om0,
                         P->partials);
      if (o + 1 < rp.n_ood)
        hipLaunchKernelGGL(claim_end_kernel, zg(dim3(1), P->jobs), dim3(SERIAL_THREADS), 0, s, P->tx, P->log, P->js, cb.slot, 1u,
                           P->partials, 1u << om1, 0u, ood_claim(o + 1));
    }
    // STIR draws (whir.rs:237-241) and the fold's alpha (prover.rs:377-384),
    // after the last OOD claim's value
    StirBegin sb{};
    sb.end_slot = rp.n_ood > 0 ? P->slots + P->n_svo_claims + (rp.n_ood - 1) : nullptr;
    sb.partials = P->partials;
    sb.n_part = 1u << om1;
    sb.n = (uint32_t)rp.n_stir;
    sb.bits = (uint32_t)rp.kfd;
    sb.idx_out = P->stir_idx;
    sb.count_out = P->stir_count;
    sb.eq_pool = P->ood_pool;
    sb.eq_stride = P->ood_stride;
    sb.n_eq = (uint32_t)rp.n_ood;
    sb.m0 = om0;
    sb.m1 = om1;
    sb.qa = P->qa;
    sb.query_scales = r > 0 ? 1u : 0u;  // later rounds' openings launch reads them
    hipLaunchKernelGGL(stir_begin_kernel, zg(dim3(1), P->jobs), dim3(SERIAL_THREADS), 0, s, P->tx, P->log, P->js, sb);
    // STIR openings (whir.rs:242-293); later rounds also build the pow tables
    // of combine_pows_scaled (prover.rs:419-425); STIR_EXTRA_BLOCKS more
    // workgroups scale by alpha what the combination reads next
    const StirScales ss{P->ood_pool, P->ood_stride, (uint32_t)rp.n_ood, (uint32_t)om0, (uint32_t)om1,
                        r == 0 ? P->qa.scale : nullptr};
    const dim3 qgrid((unsigned)rp.n_stir + STIR_EXTRA_BLOCKS);
    if (r == 0) {
      hipLaunchKernelGGL(stir_query_kernel, zg(qgrid, P->jobs), dim3(SERIAL_THREADS), 0, s, P->tx, P->js, P->log, P->stir_idx,
                         P->stir_count, P->cw0, W, P->tree0, (uint32_t)rp.kfd, 1, (uint32_t)f, P->qa, 0u, 0u,
                         (uint32_t*)nullptr, (Ext*)nullptr, ss);
    } else {
      hipLaunchKernelGGL(stir_query_kernel, zg(qgrid, P->jobs), dim3(SERIAL_THREADS), 0, s, P->tx, P->js, P->log, P->stir_idx,
                         P->stir_count, P->cw[(r - 1) & 1], 4 * W, P->tree[(r - 1) & 1], (uint32_t)rp.kfd, 2,
                         (uint32_t)f, P->qa, om0, om1, P->powL, P->powR, ss);
    }
    if (P->checks)  // the openings just written, before the next transcript kernel advances the log
      hipLaunchKernelGGL(opening_check_kernel,
                         zg(dim3((unsigned)((rp.n_stir + OPEN_CHECK_THREADS - 1) / OPEN_CHECK_THREADS)), P->jobs),
                         dim3(OPEN_CHECK_THREADS), 0, s, P->tx, P->log, P->stir_idx, P->stir_count,
                         r == 0 ? P->cw0 : P->cw[(r - 1) & 1], r == 0 ? W : 4 * W,
                         r == 0 ? P->tree0 : P->tree[(r - 1) & 1], (uint32_t)rp.kfd, fail);
    launch_combine(s, P->jobs, P->ew[cur], n, om0, P->ood_pool, (uint32_t)rp.n_ood, P->ood_stride, P->powL, P->powR,
                   r > 0 ? P->stir_count : nullptr, om1);
    if (r > 0 && n <= (1ull << SC_TAIL_MAX_LOG_N)) {
      // the round's f sumcheck rounds and folds in one workgroup (round_small.h)
      hipLaunchKernelGGL(sumcheck_tail_kernel, zg(dim3(1), P->jobs), dim3(SR_THREADS), 0, s, P->tx, P->log, P->js, P->ep[cur],
                         P->ew[cur], P->ep[cur ^ 1], P->ew[cur ^ 1], (uint32_t)n, (uint32_t)f, rs_idx);
      n >>= f;
      cur ^= f & 1;
      rs_idx += f;
      continue;
    }
    // round 0 (fold_with_code): the per-query state after the draw of step t
    // advances in QUERY_BLOCKS more workgroups of the next fold launch
    auto qfold = [&](int rho) {
      return QueryFold{P->qa, P->stir_count, (uint32_t)f, (uint32_t)rp.kfd, (uint32_t)rho, c->tw_lo(), c->tw_hi()};
    };
    // (and its query terms of the next step as QUERY_BLOCKS more partials)
    const unsigned qx = r == 0 ? QUERY_BLOCKS : 0u;
    const bool check_sc = P->checks && r > 0;
    if (check_sc) sumcheck_check(cur, n, nullptr);
    for (int t = 0; t < f; ++t) {
      unsigned g;
      if (t == 0) {
        g = std::min<unsigned>(1024, blocks_for(n / 2));
        if (r == 0)
          hipLaunchKernelGGL(coeffs_q_kernel, zg(dim3(g + qx), P->jobs), dim3(SC_THREADS), 0, s, P->ep[cur], P->ew[cur], n,
                             P->partials, qfold(0));
        else
          hipLaunchKernelGGL(coeffs_kernel, zg(dim3(g), P->jobs), dim3(SC_THREADS), 0, s, P->ep[cur], P->ew[cur], n,
                             P->partials);
      } else {
        g = std::min<unsigned>(1024, blocks_for(n / 4));
        if (r == 0)
          hipLaunchKernelGGL(fold_coeffs_q_kernel, zg(dim3(g + qx), P->jobs), dim3(SC_THREADS), 0, s, P->ep[cur], P->ew[cur], n,
                             P->js->rs + (rs_idx - 1), P->ep[cur ^ 1], P->ew[cur ^ 1], P->partials, qfold(t - 1));
        else
          hipLaunchKernelGGL(fold_coeffs_kernel, zg(dim3(g), P->jobs), dim3(SC_THREADS), 0, s, P->ep[cur], P->ew[cur], n,
                             P->js->rs + (rs_idx - 1), P->ep[cur ^ 1], P->ew[cur ^ 1], P->partials);
        n /= 2;
        cur ^= 1;
        if (check_sc) sumcheck_check(cur, n, P->js->rs + (rs_idx - 1));
      }
      hipLaunchKernelGGL(round_finalize_kernel, zg(dim3(1), P->jobs), dim3(SERIAL_THREADS), 0, s, P->tx, P->log, P->js,
                         P->partials, g + qx, rs_idx);
      ++rs_idx;
    }
    if (r == 0)
      hipLaunchKernelGGL(fold_q_kernel, zg(dim3(blocks_for(n / 2) + QUERY_BLOCKS), P->jobs), dim3(SC_THREADS), 0, s, P->ep[cur], P->ew[cur],
                         n, P->js->rs + (rs_idx - 1), P->ep[cur ^ 1], P->ew[cur ^ 1], qfold(f - 1));
    else
      hipLaunchKernelGGL(fold_kernel, zg(dim3(blocks_for(n / 2)), P->jobs), dim3(SC_THREADS), 0, s, P->ep[cur], P->ew[cur], n,
                         P->js->rs + (rs_idx - 1), P->ep[cur ^ 1], P->ew[cur ^ 1]);
    n /= 2;
    cur ^= 1;
    if (check_sc) sumcheck_check(cur, n, P->js->rs + (rs_idx - 1));
    if (r == 0) {
      // combine_pows_scaled(weights, ws, scales) at k - 2f variables (prover.rs:556)
      const int mm = rp.kc - f;
      const uint32_t pm0 = mm / 2, pm1 = mm - mm / 2;
      hipLaunchKernelGGL(pow_tables_kernel, zg(dim3(blocks_for((uint64_t)rp.n_stir * ((1ull << pm0) + (1ull << pm1)))), P->jobs),
                         dim3(SC_THREADS), 0, s, P->qa.ws, P->qa.scale, P->stir_count, pm0, pm1, P->powL, P->powR);
      launch_combine(s, P->jobs, P->ew[cur], n, pm0, nullptr, 0u, 0, P->powL, P->powR, P->stir_count, pm1);
    }
  }
  // ---- final polynomial, queries and the test's checkpoint (whir.rs:314-355, test.rs:233) ----
  {
    const int lr = pl.n_rounds - 1;
    const bool base = pl.n_rounds == 0;
    hipLaunchKernelGGL(final_phase_kernel, zg(dim3(1), P->jobs), dim3(SERIAL_THREADS), 0, s, P->tx, P->log, P->js, P->ep[cur],
                       (uint32_t)n, (uint32_t)pl.final_n_stir, (uint32_t)pl.final_kfd,
                       base ? P->cw0 : P->cw[lr & 1], base ? W : 4 * W, base ? P->tree0 : P->tree[lr & 1]);
  }
  HIP_TRY(hipGetLastError());
  return WHIR_OK;
}

// The log, the status words final_phase_kernel leaves after it and the
// transcript state, in one device-to-host copy.
// Finished proofs into pinned host memory (`host`, its device-mapped address;
// job z of a batch, blockIdx.z, at host + z * host_job_words): the log's used
// words (the first status word after the log's capacity `cap`; more than cap
// only on overflow), the status words and the transcript state at word
// tx_word. A kernel starts right after the proof's last kernel, where a DMA
// copy waited ~9 us for the copy engine, and only the written part of the log
// crosses PCIe, not the whole capacity.
static_assert(sizeof(DevTx) % 16 == 0, "the transcript state is copied in words; job slots stay 16-byte aligned");
constexpr uint32_t LOG_COPY_THREADS = 256;
__global__ void __launch_bounds__(LOG_COPY_THREADS) log_to_host_kernel(JP<const uint32_t> log_, uint32_t cap,
                                                                       uint32_t tx_word, uint32_t* __restrict__ host,
                                                                       uint64_t host_job_words) {
  const uint32_t* __restrict__ log = log_;
  host += (uint64_t)blockIdx.z * host_job_words;
  const uint32_t used = min(log[cap], cap);
  const uint32_t i = (blockIdx.x * LOG_COPY_THREADS + threadIdx.x) * 4;  // log and host slots are 16-byte aligned
  if (i + 4 <= used) {
    *reinterpret_cast<uint4*>(host + i) = *reinterpret_cast<const uint4*>(log + i);
  } else {
    for (uint32_t j = i; j < used && j < i + 4; ++j) host[j] = log[j];
  }
  if (blockIdx.x == 0) {
    if (threadIdx.x < LOG_TAIL) host[cap + threadIdx.x] = log[cap + threadIdx.x];
    for (uint32_t w = threadIdx.x; w < sizeof(DevTx) / 4; w += LOG_COPY_THREADS) host[tx_word + w] = log[tx_word + w];
  }
}

// The copy of `jobs` proofs' logs (P's layout) into host (device-mapped); a batch keeps
// its per-job DMA copies, which measured the same as one such launch.
static void enqueue_log_to_host(const whir_prover* P, uint32_t* host_dev, uint32_t jobs, hipStream_t s) {
  const uint32_t cap = (uint32_t)P->plan.log_words;
  hipLaunchKernelGGL(log_to_host_kernel, dim3((cap + 4 * LOG_COPY_THREADS - 1) / (4 * LOG_COPY_THREADS), 1, jobs),
                     dim3(LOG_COPY_THREADS), 0, s, JP<const uint32_t>(P->log), cap, (uint32_t)(P->tx_off / 4), host_dev,
                     (uint64_t)(P->host_log_bytes / 4));
}

static int enqueue_result_copy(whir_prover* P, hipStream_t s) {
  if (P->jobs == 1 && P->host_log_dev) {
    enqueue_log_to_host(P, P->host_log_dev, 1, s);
    HIP_TRY(hipGetLastError());
    return WHIR_OK;
  }
  HIP_TRY(hipMemcpyAsync(P->host_log, P->log, P->host_log_bytes, hipMemcpyDeviceToHost, s));
  return WHIR_OK;
}

static int enqueue_with_copies(whir_prover* P, const uint32_t* d_poly, hipStream_t s) {
  int rc = prover_enqueue(P, d_poly, s);
  if (rc) return rc;
  return enqueue_result_copy(P, s);
}

extern "C" int whir_prover_set_graph(whir_prover* P, int enable) {
  if (!P) return WHIR_EINVAL;
  P->graph_mode = enable != 0;
  return WHIR_OK;
}

static int prover_enqueue_impl(whir_prover* P, const uint32_t* d_poly, hipStream_t s) {
  if (!P->graph_mode) return enqueue_with_copies(P, d_poly, s);
  // The graph always proves the prover's own input buffer.
  if (P->timers) return WHIR_EINVAL;
  if (d_poly != P->input)
    HIP_TRY(hipMemcpyAsync(P->input, d_poly, (1ull << P->plan.k) * 4, hipMemcpyDeviceToDevice, s));
  if (!P->gexec) {
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_with_copies(P, P->input, s);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return rc ? rc : WHIR_EDEVICE;
    }
    hipGraphExec_t ge = nullptr;
    if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return WHIR_EDEVICE;
    }
    P->graph = g;
    P->gexec = ge;
    P->graph_stream = s;
  }
  HIP_TRY(hipGraphLaunch(P->gexec, s));
  return WHIR_OK;
}

extern "C" int whir_prover_enqueue(whir_prover* P, const uint32_t* d_poly, void* stream) {
  if (!P || !d_poly) return WHIR_EINVAL;
  hipStream_t s = pick_stream(P->ctx, stream);
  P->last_stream = s;
  P->result_ok = false;  // a failed enqueue leaves no result to read
  P->session = 0;
  if (!P->done) HIP_TRY(hipEventCreateWithFlags(&P->done, hipEventDisableTiming));
  int rc = prover_enqueue_impl(P, d_poly, s);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(P->done, s));
  P->result_ok = true;
  return WHIR_OK;
}

__global__ void zero_words_kernel(uint32_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ProverStack::layout's copy (prover.rs:114-118) into the prover's input, then the proof.
extern "C" int whir_prover_enqueue_stacked(whir_prover* P, const uint32_t* const* d_polys, void* stream) {
  if (!P || !d_polys) return WHIR_EINVAL;
  hipStream_t s = pick_stream(P->ctx, stream);
  const whir_host::StackPlan& S = P->stack;
  uint64_t used = 0;
  for (size_t id = 0; id < S.k_polys.size(); ++id) {
    if (!d_polys[id]) return WHIR_EINVAL;
    const uint64_t len = 1ull << S.k_polys[id];
    uint32_t* dst = P->input + S.offset[id];
    if (d_polys[id] != dst) HIP_TRY(hipMemcpyAsync(dst, d_polys[id], len * 4, hipMemcpyDeviceToDevice, s));
    used += len;
  }
  const uint64_t n = 1ull << S.k;  // the polys fill [0, used) from the largest down
  if (used < n)
    hipLaunchKernelGGL(zero_words_kernel, dim3(blocks_for(n - used)), dim3(SC_THREADS), 0, s, P->input + used,
                       n - used);
  HIP_TRY(hipGetLastError());
  return whir_prover_enqueue(P, P->input, stream);
}

// WHIR_OK when the last enqueued proof has finished, 7 while it is running.
extern "C" int whir_prover_query(whir_prover* P) {
  if (!P || !P->done) return WHIR_EINVAL;
  hipError_t e = hipEventQuery(P->done);
  if (e == hipSuccess) return WHIR_OK;
  if (e == hipErrorNotReady) return 7;
  return WHIR_EDEVICE;
}

extern "C" int whir_prover_result(whir_prover* P, uint8_t* proof, size_t cap, size_t* len, uint32_t* checkpoint,
                                  uint32_t* root8) {
  if (!P || !P->result_ok || !P->done) return WHIR_EINVAL;
  HIP_TRY(hipEventSynchronize(P->done));
  if (P->timer_pending) {
    for (int i = 0; i < WHIR_TIMER_COUNT; ++i) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, P->ev[i][0], P->ev[i][1]));
      P->timer_ms[i] += ms;
      P->timer_launches[i] += 1;
    }
    P->timer_pending = false;
  }
  const uint32_t* tail = P->host_log + P->plan.log_words;  // log_words, overflow, checkpoint, failed checks
  if (tail[1] || tail[0] > P->plan.log_words) return WHIR_ETRANSCRIPT;
  const size_t bytes = (size_t)tail[0] * 4;
  if (len) *len = bytes;
  if (checkpoint) *checkpoint = tail[2];
  if (root8) memcpy(root8, P->host_log, 32);  // the first transcript write is the root
  if (proof) {
    if (bytes > cap) return WHIR_ETRANSCRIPT;
    memcpy(proof, P->host_log, bytes);
  }
  return P->checks && tail[3] ? WHIR_ECHECK : WHIR_OK;
}

extern "C" int whir_prover_result_view(whir_prover* P, const uint8_t** proof, size_t* len, uint32_t* checkpoint,
                                       uint32_t* root8) {
  const int rc = whir_prover_result(P, nullptr, 0, len, checkpoint, root8);
  if (rc && rc != WHIR_ECHECK) return rc;
  if (proof) *proof = (const uint8_t*)P->host_log;
  return rc;
}

extern "C" int whir_prover_set_checks(whir_prover* P, int enable) {
  if (!P) return WHIR_EINVAL;
  if (P->checks != (enable != 0)) {  // a captured proof graph has the old launch sequence
    if (P->done) HIP_TRY(hipEventSynchronize(P->done));
    if (P->gexec) HIP_TRY(hipGraphExecDestroy(P->gexec));
    if (P->graph) HIP_TRY(hipGraphDestroy(P->graph));
    P->gexec = nullptr;
    P->graph = nullptr;
  }
  P->checks = enable != 0;
  return WHIR_OK;
}

extern "C" size_t whir_prover_device_bytes(const whir_prover* P) { return P ? P->arena_bytes : 0; }

extern "C" int whir_prove_host(whir_ctx* c, const whir_params* prm, const uint32_t* poly, uint8_t* proof, size_t cap,
                               size_t* len, uint32_t* checkpoint, uint32_t* root8) {
  if (!poly) return WHIR_EINVAL;
  whir_prover* P = nullptr;
  int rc = whir_prover_create(c, prm, &P);
  if (rc) return rc;
  if (hipMemcpy(P->input, poly, (1ull << P->plan.k) * 4, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    rc = WHIR_EDEVICE;
  }
  if (!rc) rc = whir_prover_enqueue(P, P->input, nullptr);
  if (!rc) rc = whir_prover_result(P, proof, cap, len, checkpoint, root8);
  const int rd = whir_prover_destroy(P);
  return rc ? rc : rd;
}

// ---- commit / eval / open with a caller's transcript ----

static bool canonical_word(uint32_t w) { return w < kb::P; }

// whir_challenger -> DevTx with an empty log of `cap` words; false if the
// state is not one a DuplexChallenger can be in (see whir_gpu.h).
static bool challenger_import(const whir_challenger* c, uint32_t cap, DevTx& t) {
  if (c->input_len > 7 || c->output_len > 8 || (c->input_len && c->output_len)) return false;
  memset(&t, 0, sizeof(t));
  for (int i = 0; i < 16; ++i) {
    if (!canonical_word(c->sponge_state[i])) return false;
    t.state[i] = c->sponge_state[i];
  }
  for (uint32_t i = 0; i < c->input_len; ++i) {
    if (!canonical_word(c->input_buffer[i])) return false;
    t.in_buf[i] = c->input_buffer[i];
  }
  // the output buffer is the sponge state's prefix left by the last duplex
  for (uint32_t i = 0; i < c->output_len; ++i)
    if (c->output_buffer[i] != c->sponge_state[i]) return false;
  for (int i = 0; i < 8; ++i) t.out_buf[i] = c->sponge_state[i];
  t.in_len = c->input_len;
  t.out_len = c->output_len;
  t.log_words = 0;
  t.log_cap = cap;
  t.overflow = 0;
  return true;
}

extern "C" int whir_prover_commit(whir_prover* P, const uint32_t* d_poly, const whir_challenger* tx_in, void* stream) {
  if (!P || !d_poly) return WHIR_EINVAL;
  DevTx init;
  if (tx_in && !challenger_import(tx_in, (uint32_t)P->plan.log_words, init)) return WHIR_EINVAL;
  hipStream_t s = pick_stream(P->ctx, stream);
  P->last_stream = s;
  P->result_ok = false;
  P->session = 0;
  int rc = enqueue_commit(P, d_poly, s, ClaimBegin{}, tx_in ? &init : nullptr);
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  P->session = 1;
  P->next_claim = 0;
  P->session_poly = d_poly;
  return WHIR_OK;
}

extern "C" int whir_prover_eval(whir_prover* P, const uint32_t* point, void* stream) {
  if (!P || P->session != 1) return WHIR_EINVAL;
  const int c = P->next_claim;
  const int n_user = (int)P->stack.claim_poly.size();
  if (c >= n_user) return WHIR_EINVAL;
  int g = 0;
  while (g + 1 < P->n_user_groups && (int)P->groups[g + 1].c0 <= c) ++g;
  const whir_prover::Group& G = P->groups[g];
  const bool lead = (int)G.c0 == c;
  if (!lead && point) return WHIR_EINVAL;  // a claim at the previous claim's point
  hipStream_t s = pick_stream(P->ctx, stream);
  const int f = P->plan.f;
  const ClaimBegin cb = group_begin(P, g);
  if (lead) {
    GivenPoint gp;
    memset(&gp, 0, sizeof(gp));
    if (point) {
      for (uint32_t j = 0; j < G.k; ++j)
        for (int q = 0; q < 4; ++q) {
          const uint32_t w = point[4 * j + q];
          if (!canonical_word(w)) return WHIR_EINVAL;
          gp.z[j].c[q] = w;
        }
    }
    hipLaunchKernelGGL(claim_begin_kernel, zg(dim3(1), P->jobs), dim3(SERIAL_THREADS), 0, s, P->tx, P->log, P->js, cb, gp,
                       point ? 1u : 0u);
  }
  // this claim alone: its compression with the group's eq tables, then its value
  FSWITCH(f, launch_compress_hi<FL>(dim3(1u << cb.m1, 1), s, P->session_poly, cb.eq0, cb.eq1, cb.m0, P->partials,
                                    P->desc, (uint32_t)c, P->jobs));
  hipLaunchKernelGGL(claim_end_kernel, zg(dim3(1), P->jobs), dim3(SERIAL_THREADS), 0, s, P->tx, P->log, P->js, P->slots + c, 1u,
                     P->partials, 1u << cb.m1, (uint32_t)f, ClaimBegin{});
  HIP_TRY(hipGetLastError());
  P->next_claim = c + 1;
  return WHIR_OK;
}

extern "C" int whir_prover_open(whir_prover* P, void* stream) {
  if (!P || P->session != 1 || P->next_claim != (int)P->stack.claim_poly.size()) return WHIR_EINVAL;
  hipStream_t s = pick_stream(P->ctx, stream);
  if (!P->done) HIP_TRY(hipEventCreateWithFlags(&P->done, hipEventDisableTiming));
  const int ng = (int)P->groups.size();
  int rc = WHIR_OK;
  if (P->n_user_groups < ng) {  // the OOD virtual claims (whir.rs:187-192), drawn
    hipLaunchKernelGGL(claim_begin_kernel, zg(dim3(1), P->jobs), dim3(SERIAL_THREADS), 0, s, P->tx, P->log, P->js,
                       group_begin(P, P->n_user_groups), GivenPoint{}, 0u);
    rc = enqueue_groups(P, P->session_poly, s, P->n_user_groups, ng);
  }
  if (!rc) rc = enqueue_open(P, P->session_poly, s);
  if (!rc) rc = enqueue_result_copy(P, s);
  P->session = 0;
  if (rc) return rc;
  HIP_TRY(hipEventRecord(P->done, s));
  P->result_ok = true;
  return WHIR_OK;
}

extern "C" int whir_prover_challenger(whir_prover* P, whir_challenger* out) {
  if (!P || !out || !P->result_ok || !P->done) return WHIR_EINVAL;
  HIP_TRY(hipEventSynchronize(P->done));
  const DevTx* t = (const DevTx*)((const char*)P->host_log + P->tx_off);
  memset(out, 0, sizeof(*out));
  for (int i = 0; i < 16; ++i) out->sponge_state[i] = t->state[i];
  out->input_len = t->in_len;
  out->output_len = t->out_len;
  for (uint32_t i = 0; i < t->in_len && i < 8; ++i) out->input_buffer[i] = t->in_buf[i];
  for (uint32_t i = 0; i < t->out_len && i < 8; ++i) out->output_buffer[i] = t->state[i];
  return WHIR_OK;
}

// ============================ seeded inputs ============================

// Poly::<F>::rand(&mut rng(seed), k) (crates/poly/src/poly.rs:47-52 with
// crates/common/src/test.rs:1-4): 2^k StandardUniform KoalaBear draws of
// StdRng::seed_from_u64(seed), as Montgomery words, on the host.
extern "C" int whir_poly_rand(uint64_t seed, uint32_t k, uint32_t* out) {
  if (!out || k > 32) return WHIR_EINVAL;
  whir_host::ChaChaRng r(seed);
  const uint64_t n = 1ull << k;
  for (uint64_t i = 0; i < n; ++i) out[i] = r.next_field();
  return WHIR_OK;
}

// ============================ memory helpers ============================

extern "C" int whir_device_alloc(size_t bytes, void** out) {
  if (!out) return WHIR_EINVAL;
  return hipMalloc(out, bytes) == hipSuccess ? WHIR_OK : WHIR_EOOM;
}
extern "C" int whir_device_free(void* p) {
  HIP_TRY(hipFree(p));
  return WHIR_OK;
}
extern "C" int whir_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return WHIR_OK;
}
extern "C" int whir_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return WHIR_OK;
}
extern "C" int whir_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return WHIR_OK;
}
extern "C" int whir_device_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return WHIR_OK;
}

#include "server.h"
#include "batch.h"
#include "components.h"
