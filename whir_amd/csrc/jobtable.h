// Commit work queue: the data-parallel part of many proofs' commitments, for
// jobs of uneven sizes, in a few persistent launches.
//
// A set of jobs (whir_multi) commits each polynomial as Whir::commit does
// (crates/whir/src/whir.rs:19-36: the coset LDE of the 2^f-wide rows, then the
// Poseidon2 tree of the codeword rows, poseidon_packed.rs:96-143, 268-304).
// Here every job's LDE passes, leaf hashes and wide tree layers are cut into
// workgroup-sized items and listed in a device job table; one launch per stage
// drains all jobs' items of that stage:
//   stage "ntt p"   : pass p of every job's transform (ntt_block, any radix)
//   stage "leaves"  : every job's leaf sponges (leaf_hash_row<16>)
//   stage "layer l" : layer l of every job's tree down to where that job's
//                     cooperative subtrees begin (layer_node)
// The kernels are persistent: a grid sized to the device's resident
// workgroups, each workgroup dequeuing items with one atomic on the stage's
// head counter until the list is empty. The kernel boundary between stages is
// the only synchronisation (pass p + 1 reads what pass p wrote, a layer reads
// the layer below it). Results are those of the per-job launches: the same
// device functions compute every row, node and butterfly.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <vector>
#include "merkle.h"
#include "ntt_plan.h"

namespace kb {

constexpr int JT_THREADS = 256;
constexpr uint32_t JT_MAX_STAGES = 64;  // head counters: <= 3 NTT passes x radix, the leaves, <= 28 layers
constexpr uint32_t JT_KIND_REG = 16;    // item kind: pass radix 2^l, + JT_KIND_REG for a register-only pass

struct JtItem {
  uint32_t job;   // index into the stage's job arrays
  uint32_t kind;  // NTT: pass radix (see above); otherwise 0
  uint32_t blk;   // the item's workgroup-sized block
  uint32_t nblk;  // 1
};

// One job's base tree: the codeword rows (16 words each: 2^f with f = 4) and
// its layers, leaves first (merkle.h layout).
struct JtTree {
  const uint32_t* rows;
  uint32_t* tree;
  uint64_t n_rows;
};

// Next item index of this workgroup (uniform): from the stage's head counter
// (one atomic per item; every thread has read the slot before the next call
// overwrites it), or with head == nullptr the static schedule prev + gridDim.x
// (first = blockIdx.x).
__device__ __forceinline__ uint32_t jt_next(uint32_t* head, uint32_t* s_slot, uint32_t prev, bool first) {
  if (!head) return first ? blockIdx.x : prev + gridDim.x;
  if (threadIdx.x == 0) *s_slot = atomicAdd(head, 1u);
  __syncthreads();
  const uint32_t it = __builtin_amdgcn_readfirstlane(*s_slot);
  __syncthreads();
  return it;
}

// Stage "ntt p, radix (LR1, LR2)": args[job] is pass p of that job's
// transform; one kernel instance per radix (the items of a pass are grouped by
// their radix), so each keeps its own register allocation.
template <int LR1, int LR2>
__global__ void __launch_bounds__(JT_THREADS) jt_ntt_kernel(const NttArgs* __restrict__ args,
                                                            const JtItem* __restrict__ items, uint32_t n_items,
                                                            uint32_t* head) {
  extern __shared__ uint32_t lds[];
  __shared__ uint32_t s_slot;
  for (uint32_t it = jt_next(head, &s_slot, 0, true); it < n_items; it = jt_next(head, &s_slot, it, false)) {
    const JtItem q = items[it];
    const NttArgs a = args[q.job];
    ntt_block<LR1, LR2, 2>(a, q.blk, lds);  // NTT items are single blocks
    if (LR2 > 0) __syncthreads();           // the next item refills the LDS tile
  }
}

typedef void (*JtNttFn)(const NttArgs*, const JtItem*, uint32_t, uint32_t*);
// The kernel of an item kind: radix 2^l as ntt_kernel_for / ntt_reg_kernel_for
// pick it (two words per line).
inline JtNttFn jt_ntt_fn(uint32_t kind) {
  switch (kind) {
    case 1: return jt_ntt_kernel<1, 0>;
    case 2: return jt_ntt_kernel<1, 1>;
    case 3: return jt_ntt_kernel<2, 1>;
    case 4: return jt_ntt_kernel<2, 2>;
    case 5: return jt_ntt_kernel<3, 2>;
    case 6: return jt_ntt_kernel<3, 3>;
    case 7: return jt_ntt_kernel<4, 3>;
    case 8: return jt_ntt_kernel<4, 4>;
    case JT_KIND_REG + 2: return jt_ntt_kernel<2, 0>;
    case JT_KIND_REG + 3: return jt_ntt_kernel<3, 0>;
    case JT_KIND_REG + 4: return jt_ntt_kernel<4, 0>;
    default: return nullptr;
  }
}

// Stage "leaves": leaf i of job j for the rows of each item's blocks.
__global__ void __launch_bounds__(JT_THREADS) jt_leaves_kernel(const JtTree* __restrict__ jobs,
                                                               const JtItem* __restrict__ items, uint32_t n_items,
                                                               uint32_t* head) {
  __shared__ uint32_t s_slot;
  for (uint32_t it = jt_next(head, &s_slot, 0, true); it < n_items; it = jt_next(head, &s_slot, it, false)) {
    const JtItem q = items[it];
    const JtTree J = jobs[q.job];
    const uint64_t i = (uint64_t)q.blk * JT_THREADS + threadIdx.x;
    if (i < J.n_rows) leaf_hash_row<16, true>(J.rows, i, J.tree);
  }
}

// Stage "layer l" (l >= 1): layer l - 1 of H >> (l - 1) nodes, stored at digest
// offset 2H - 2 (H >> (l - 1)), compressed into layer l right after it.
__global__ void __launch_bounds__(JT_THREADS) jt_layer_kernel(const JtTree* __restrict__ jobs,
                                                              const JtItem* __restrict__ items, uint32_t n_items,
                                                              uint32_t* head, uint32_t level) {
  __shared__ uint32_t s_slot;
  for (uint32_t it = jt_next(head, &s_slot, 0, true); it < n_items; it = jt_next(head, &s_slot, it, false)) {
    const JtItem q = items[it];
    const JtTree J = jobs[q.job];
    const uint64_t n_prev = J.n_rows >> (level - 1), half = n_prev >> 1;
    const uint32_t* prev = J.tree + (2 * J.n_rows - 2 * n_prev) * 8;
    uint32_t* next = J.tree + (2 * J.n_rows - n_prev) * 8;
    const uint64_t j = (uint64_t)q.blk * JT_THREADS + threadIdx.x;
    if (j < half) layer_node<true>(prev, half, j, next);
  }
}

__global__ void jt_reset_kernel(uint32_t* heads, uint32_t n) {
  if (threadIdx.x < n) heads[threadIdx.x] = 0u;
}

// ---- host: the job table of a set ----

// One job's commitment as the queue runs it.
struct JtJobSpec {
  const uint32_t* input;  // 2^(k-f) rows of 16 words
  uint32_t* cw;           // the codeword, 2^log_h rows of 16 words
  uint32_t* tmp;          // NTT scratch (the work array of a multi-pass transform)
  uint32_t* tree;
  uint64_t n_in;
  int log_h;
  int wide_levels;        // tree layers the queue computes above the leaves
};

enum JtStageKind { JT_STAGE_NTT = 0, JT_STAGE_LEAVES = 1, JT_STAGE_LAYER = 2 };
struct JtStage {
  int kind = JT_STAGE_NTT;
  uint32_t arg = 0;      // NTT pass or tree level
  uint32_t ntt_kind = 0; // NTT: the items' radix kind
  size_t items_off = 0;  // bytes into the table allocation
  size_t jobs_off = 0;   // NttArgs[] or JtTree[] of the stage
  uint32_t n_items = 0;
  unsigned grid = 0;
  size_t lds = 0;
};

struct JobTable {
  void* dev = nullptr;  // heads, then the stages' job arrays and items
  size_t bytes = 0;
  std::vector<JtStage> stages;
  uint32_t n_jobs = 0;
  // items dequeued through the head counters; false: the static schedule
  // (workgroup w takes items w, w + grid, ...; it measured the same on the
  // 64-job set: every item is one workgroup-sized block)
  bool dynamic = true;
};

// The pass plan the queue uses for a job (two words per line, 256-thread
// workgroups, passes of up to 256 rows, a register-only first pass when three
// passes are needed), each pass with its item kind: false if a pass does not
// fit the queue kernel.
inline bool jt_ntt_passes(const JtJobSpec& j, const uint32_t* tw_lo, const uint32_t* tw_hi,
                          std::vector<NttPass>& out, std::vector<uint32_t>& kinds) {
  NttPlan pl;
  if (!ntt_plan(j.n_in, j.log_h, 4, pl, 8, 4, 1, 0) || pl.lv != 1) return false;
  out.clear();
  kinds.clear();
  for (int p = 0; p < pl.P; ++p) {
    NttPass ps;
    ntt_pass_setup(pl, p, j.n_in, 8, j.input, j.cw, j.tmp, tw_lo, tw_hi, ps);
    const int l = pl.logl[p];
    if (ps.threads != (unsigned)JT_THREADS) return false;
    if (ps.kind == NTT_PASS_REG && l >= 2 && l <= 4) kinds.push_back(JT_KIND_REG + (uint32_t)l);
    else if (ps.kind == NTT_PASS_LDS && l >= 1 && l <= 8) kinds.push_back((uint32_t)l);
    else return false;
    out.push_back(ps);
  }
  return true;
}

// Items of one stage: per job, its blocks in runs of `chunk` (an item never
// spans two jobs), jobs in the order given.
inline void jt_items(const std::vector<std::pair<uint32_t, uint64_t>>& job_blocks,
                     const std::vector<uint32_t>& kinds, uint32_t chunk, std::vector<JtItem>& out) {
  out.clear();
  for (size_t i = 0; i < job_blocks.size(); ++i)
    for (uint64_t b = 0; b < job_blocks[i].second; b += chunk)
      out.push_back(JtItem{job_blocks[i].first, kinds[i], (uint32_t)b,
                           (uint32_t)std::min<uint64_t>(chunk, job_blocks[i].second - b)});
}

// Builds the table (device allocation included) for `jobs`, each of which
// jt_ntt_passes accepts, on a device with n_cu compute units.
inline int jt_build(const std::vector<JtJobSpec>& jobs, const uint32_t* tw_lo, const uint32_t* tw_hi, int n_cu,
                    JobTable& T) {
  T = JobTable();
  T.n_jobs = (uint32_t)jobs.size();
  std::vector<std::vector<NttPass>> passes(jobs.size());
  std::vector<std::vector<uint32_t>> pass_kinds(jobs.size());
  int max_p = 0, max_l = 0;
  for (size_t j = 0; j < jobs.size(); ++j) {
    if (!jt_ntt_passes(jobs[j], tw_lo, tw_hi, passes[j], pass_kinds[j])) return 1;
    max_p = std::max(max_p, (int)passes[j].size());
    max_l = std::max(max_l, jobs[j].wide_levels);
  }
  // host images of the stages' job arrays and items, laid out after the heads
  std::vector<uint8_t> img(JT_MAX_STAGES * 4, 0);
  auto append = [&](const void* src, size_t n) {
    const size_t off = (img.size() + 255) & ~(size_t)255;
    img.resize(off + n);
    memcpy(img.data() + off, src, n);
    return off;
  };
  // an item is one workgroup-sized block: a loop over several blocks per
  // dequeue makes the compiler keep twice the registers of the block body
  auto chunk_for = [&](uint64_t, int) { return 1u; };
  auto finish_stage = [&](JtStage& st, const std::vector<JtItem>& items, int occ) {
    st.items_off = append(items.data(), items.size() * sizeof(JtItem));
    st.n_items = (uint32_t)items.size();
    st.grid = (unsigned)std::min<uint64_t>(items.size(), (uint64_t)std::max(1, occ) * (uint64_t)n_cu);
    T.stages.push_back(st);
  };
  for (int p = 0; p < max_p; ++p) {
    // the jobs' pass p, one stage per radix kind
    std::vector<uint32_t> stage_kinds;
    for (size_t j = 0; j < jobs.size(); ++j)
      if ((int)passes[j].size() > p &&
          std::find(stage_kinds.begin(), stage_kinds.end(), pass_kinds[j][p]) == stage_kinds.end())
        stage_kinds.push_back(pass_kinds[j][p]);
    for (uint32_t kind : stage_kinds) {
      JtStage st;
      st.kind = JT_STAGE_NTT;
      st.arg = (uint32_t)p;
      st.ntt_kind = kind;
      std::vector<NttArgs> args;
      std::vector<std::pair<uint32_t, uint64_t>> jb;
      std::vector<uint32_t> kinds;
      uint64_t total = 0;
      for (size_t j = 0; j < jobs.size(); ++j) {
        if ((int)passes[j].size() <= p || pass_kinds[j][p] != kind) continue;
        const NttPass& ps = passes[j][p];
        jb.push_back({(uint32_t)args.size(), ps.grid});
        kinds.push_back(kind);
        args.push_back(ps.a);
        total += ps.grid;
        st.lds = std::max(st.lds, ps.lds);
      }
      int occ = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, jt_ntt_fn(kind), JT_THREADS, st.lds) != hipSuccess) {
        (void)hipGetLastError();
        return 2;
      }
      st.jobs_off = append(args.data(), args.size() * sizeof(NttArgs));
      std::vector<JtItem> items;
      jt_items(jb, kinds, chunk_for(total, occ), items);
      finish_stage(st, items, occ);
    }
  }
  std::vector<JtTree> trees(jobs.size());
  for (size_t j = 0; j < jobs.size(); ++j) trees[j] = JtTree{jobs[j].cw, jobs[j].tree, 1ull << jobs[j].log_h};
  const size_t trees_off = append(trees.data(), trees.size() * sizeof(JtTree));
  for (int l = 0; l <= max_l; ++l) {
    JtStage st;
    st.kind = l == 0 ? JT_STAGE_LEAVES : JT_STAGE_LAYER;
    st.arg = (uint32_t)l;
    st.jobs_off = trees_off;
    std::vector<std::pair<uint32_t, uint64_t>> jb;
    std::vector<uint32_t> kinds;
    uint64_t total = 0;
    for (size_t j = 0; j < jobs.size(); ++j) {
      if (jobs[j].wide_levels < l) continue;
      const uint64_t nodes = l == 0 ? (1ull << jobs[j].log_h) : (1ull << jobs[j].log_h) >> l;
      const uint64_t blocks = (nodes + JT_THREADS - 1) / JT_THREADS;
      jb.push_back({(uint32_t)j, blocks});
      kinds.push_back(0u);
      total += blocks;
    }
    int occ = 0;
    const hipError_t e = l == 0 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, jt_leaves_kernel, JT_THREADS, 0)
                                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, jt_layer_kernel, JT_THREADS, 0);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return 2;
    }
    std::vector<JtItem> items;
    jt_items(jb, kinds, chunk_for(total, occ), items);
    finish_stage(st, items, occ);
  }
  if (T.stages.size() > JT_MAX_STAGES) return 1;  // one head counter per stage
  if (hipMalloc(&T.dev, img.size()) != hipSuccess) {
    (void)hipGetLastError();
    T.dev = nullptr;
    return 3;
  }
  T.bytes = img.size();
  if (hipMemcpy(T.dev, img.data(), img.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    return 2;
  }
  return 0;
}

// The stages on stream s: the head counters are cleared by a kernel (the
// sequence may be captured into a graph and replayed), then one launch per
// stage.
inline void jt_enqueue(const JobTable& T, hipStream_t s) {
  uint8_t* base = (uint8_t*)T.dev;
  uint32_t* heads = (uint32_t*)base;
  hipLaunchKernelGGL(jt_reset_kernel, dim3(1), dim3(64), 0, s, heads, JT_MAX_STAGES);
  for (size_t i = 0; i < T.stages.size(); ++i) {
    const JtStage& st = T.stages[i];
    const JtItem* items = (const JtItem*)(base + st.items_off);
    if (st.n_items == 0) continue;
    if (st.kind == JT_STAGE_NTT)
      hipLaunchKernelGGL(jt_ntt_fn(st.ntt_kind), dim3(st.grid), dim3(JT_THREADS), st.lds, s,
                         (const NttArgs*)(base + st.jobs_off), items, st.n_items, T.dynamic ? heads + i : nullptr);
    else if (st.kind == JT_STAGE_LEAVES)
      hipLaunchKernelGGL(jt_leaves_kernel, dim3(st.grid), dim3(JT_THREADS), 0, s,
                         (const JtTree*)(base + st.jobs_off), items, st.n_items, T.dynamic ? heads + i : nullptr);
    else
      hipLaunchKernelGGL(jt_layer_kernel, dim3(st.grid), dim3(JT_THREADS), 0, s,
                         (const JtTree*)(base + st.jobs_off), items, st.n_items, T.dynamic ? heads + i : nullptr,
                         st.arg);
  }
}

}  // namespace kb
