// Lockstep batches: B independent proofs of one shape run as ONE kernel
// sequence. Every kernel of the proof is launched once with gridDim.z = B;
// workgroup z proves job z on its own copy of the workspace. The copies sit at
// a fixed stride in one allocation, and the kernels' pointer arguments carry
// that stride (job.h). The shape's claim tables are read-only and identical for
// every job, so all jobs share job 0's copy.
//
// A batch of B costs the launches of one proof (~90 at k = 24) instead of B
// times as many, and the transcript kernels (one workgroup per job) run all
// jobs' serial chains side by side. Each job's transcript, arithmetic and
// proof bytes are exactly those of the job proven alone.
#pragma once
#include "jobtable.h"

struct whir_batch {
  whir_prover* P = nullptr;  // job 0's layout; its workspace pointers carry the stride
  uint32_t jobs = 0;
  uint64_t stride = 0;       // bytes between jobs' workspaces (a multiple of 64 KiB)
  uint8_t* host = nullptr;   // pinned: per job the log, its status words and the transcript
  hipEvent_t done = nullptr;
  bool result_ok = false;
  bool graph_mode = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
};

extern "C" int whir_batch_destroy(whir_batch* B) {
  if (!B) return WHIR_OK;
  int rc = WHIR_OK;
  if (B->gexec) release_step(rc, hipGraphExecDestroy(B->gexec));
  if (B->graph) release_step(rc, hipGraphDestroy(B->graph));
  if (B->done) release_step(rc, hipEventDestroy(B->done));
  if (B->host) release_step(rc, hipHostFree(B->host));
  if (B->P) {
    const int rp = whir_prover_destroy(B->P);  // releases the batch allocation (P->arena)
    if (rc == WHIR_OK) rc = rp;
  }
  delete B;
  return rc;
}

extern "C" int whir_batch_create(whir_ctx* c, const whir_params* prm, int jobs, whir_batch** out) {
  if (!c || !prm || !out || jobs < 1 || jobs > 4096) return WHIR_EINVAL;
  whir_batch* B = new whir_batch();
  B->jobs = (uint32_t)jobs;
  whir_prover* P = new whir_prover();
  P->ctx = c;
  B->P = P;
  std::vector<ArenaItem> items;
  int rc = prover_layout(P, prm, nullptr, items);
  if (rc) {
    whir_batch_destroy(B);
    return rc;
  }
  B->stride = (arena_size(items) + (1ull << JOB_STRIDE_UNIT_LOG) - 1) & ~((1ull << JOB_STRIDE_UNIT_LOG) - 1);
  if ((B->stride >> JOB_STRIDE_UNIT_LOG) >= (1ull << (64 - JOB_TAG_SHIFT))) {
    whir_batch_destroy(B);
    return WHIR_EINVAL;  // a workspace of 4 GiB or more
  }
  const uint64_t total = B->stride * (uint64_t)jobs;
  if (hipMalloc(&P->arena, total) != hipSuccess) {
    (void)hipGetLastError();
    P->arena = nullptr;
    whir_batch_destroy(B);
    return WHIR_EOOM;
  }
  if (((uint64_t)P->arena + total) >> JOB_TAG_SHIFT) {  // addresses must leave the tag bits free
    whir_batch_destroy(B);
    return WHIR_EDEVICE;
  }
  P->arena_bytes = total;
  if ((rc = prover_bind(P, items))) {
    whir_batch_destroy(B);
    return rc;
  }
  // every workspace buffer is per job, except the shape's shared claim tables
  const void* const shared[] = {&P->desc, &P->group_start, &P->seg_start, &P->seg_claims};
  for (const ArenaItem& it : items) {
    bool is_shared = false;
    for (const void* sp : shared) is_shared |= (const void*)it.ptr == sp;
    if (!is_shared) *it.ptr = job_tag(*it.ptr, B->stride);
  }
  P->tx = (DevTx*)((char*)P->log + P->tx_off);
  P->jobs = B->jobs;
  if (hipHostMalloc((void**)&B->host, P->host_log_bytes * jobs, hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&B->done, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    whir_batch_destroy(B);
    return WHIR_EOOM;
  }
  if (c->autotune) prover_autotune(P);  // timed with the batch's job count, kept for that job-count class
  *out = B;
  return WHIR_OK;
}

extern "C" uint32_t* whir_batch_input(whir_batch* B, int job) {
  if (!B || job < 0 || (uint32_t)job >= B->jobs) return nullptr;
  return job_addr(B->P->input, (uint32_t)job);
}

extern "C" int whir_batch_set_graph(whir_batch* B, int enable) {
  if (!B) return WHIR_EINVAL;
  B->graph_mode = enable != 0;
  return WHIR_OK;
}

extern "C" int whir_batch_set_checks(whir_batch* B, int enable) {
  if (!B) return WHIR_EINVAL;
  if (B->P->checks != (enable != 0)) {  // a captured graph has the old launch sequence
    if (B->done) HIP_TRY(hipEventSynchronize(B->done));
    if (B->gexec) HIP_TRY(hipGraphExecDestroy(B->gexec));
    if (B->graph) HIP_TRY(hipGraphDestroy(B->graph));
    B->gexec = nullptr;
    B->graph = nullptr;
  }
  B->P->checks = enable != 0;
  return WHIR_OK;
}

static int batch_enqueue_proofs(whir_batch* B, hipStream_t s) {
  whir_prover* P = B->P;
  int rc = prover_enqueue(P, P->input, s);
  if (rc) return rc;
  for (uint32_t j = 0; j < B->jobs; ++j)
    HIP_TRY(hipMemcpyAsync(B->host + (size_t)j * P->host_log_bytes, job_addr(P->log, j), P->host_log_bytes,
                           hipMemcpyDeviceToHost, s));
  return WHIR_OK;
}

extern "C" int whir_batch_enqueue(whir_batch* B, const uint32_t* const* d_polys, void* stream) {
  if (!B) return WHIR_EINVAL;
  whir_prover* P = B->P;
  hipStream_t s = pick_stream(P->ctx, stream);
  B->result_ok = false;
  if (d_polys)
    for (uint32_t j = 0; j < B->jobs; ++j) {
      uint32_t* dst = job_addr(P->input, j);
      if (d_polys[j] && d_polys[j] != dst)
        HIP_TRY(hipMemcpyAsync(dst, d_polys[j], (1ull << P->plan.k) * 4, hipMemcpyDeviceToDevice, s));
    }
  if (!B->graph_mode) {
    int rc = batch_enqueue_proofs(B, s);
    if (rc) return rc;
  } else {
    if (!B->gexec) {
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int rc = batch_enqueue_proofs(B, s);
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
      B->graph = g;
      B->gexec = ge;
    }
    HIP_TRY(hipGraphLaunch(B->gexec, s));
  }
  HIP_TRY(hipEventRecord(B->done, s));
  B->result_ok = true;
  return WHIR_OK;
}

extern "C" int whir_batch_query(whir_batch* B) {
  if (!B || !B->result_ok) return WHIR_EINVAL;
  const hipError_t e = hipEventQuery(B->done);
  if (e == hipSuccess) return WHIR_OK;
  if (e == hipErrorNotReady) return WHIR_EBUSY;
  return WHIR_EDEVICE;
}

extern "C" int whir_batch_result_view(whir_batch* B, int job, const uint8_t** proof, size_t* len,
                                      uint32_t* checkpoint, uint32_t* root8) {
  if (!B || !B->result_ok || job < 0 || (uint32_t)job >= B->jobs) return WHIR_EINVAL;
  HIP_TRY(hipEventSynchronize(B->done));
  const whir_prover* P = B->P;
  const uint32_t* log = (const uint32_t*)(B->host + (size_t)job * P->host_log_bytes);
  const uint32_t* tail = log + P->plan.log_words;  // log_words, overflow, checkpoint, failed checks
  if (tail[1] || tail[0] > P->plan.log_words) return WHIR_ETRANSCRIPT;
  if (len) *len = (size_t)tail[0] * 4;
  if (checkpoint) *checkpoint = tail[2];
  if (root8) memcpy(root8, log, 32);
  if (proof) *proof = (const uint8_t*)log;
  return P->checks && tail[3] ? WHIR_ECHECK : WHIR_OK;
}

extern "C" int whir_batch_result(whir_batch* B, int job, uint8_t* proof, size_t cap, size_t* len,
                                 uint32_t* checkpoint, uint32_t* root8) {
  const uint8_t* src = nullptr;
  size_t bytes = 0;
  const int rc = whir_batch_result_view(B, job, &src, &bytes, checkpoint, root8);
  if (rc && rc != WHIR_ECHECK) return rc;
  if (len) *len = bytes;
  if (proof) {
    if (bytes > cap) return WHIR_ETRANSCRIPT;
    memcpy(proof, src, bytes);
  }
  return rc;
}

extern "C" size_t whir_batch_proof_capacity(const whir_batch* B) { return B ? B->P->plan.log_words * 4 : 0; }
extern "C" size_t whir_batch_device_bytes(const whir_batch* B) { return B ? B->P->arena_bytes : 0; }

// ---- many jobs of mixed shapes: lockstep batches per shape, side by side ----
// The jobs are grouped by proof shape into lockstep batches (largest first);
// the batches are spread over `streams` internal streams (each batch to the
// stream with the least work so far, work ~ the batch's workspace bytes) and
// joined back into the caller's stream. Graph mode 1 captures the whole set —
// every batch on its stream, the fork and the join — once as one hipGraph,
// replayed as one launch; mode 2 captures each batch as a graph of its own,
// launched on its stream.
struct whir_multi {
  whir_ctx* ctx = nullptr;
  std::vector<whir_batch*> batches;
  std::vector<int> batch_stream;              // internal stream of each batch
  std::vector<std::pair<int, int>> job_slot;  // per job: (batch, job within it)
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;             // fork event, then one join event per stream
  int graph_mode = 0;  // 0 eager, 1 one graph for the set, 2 one graph per batch
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t done = nullptr;
  bool result_ok = false;
  // commit work queue (jobtable.h): every eligible job's base LDE, leaves and
  // wide tree layers in persistent launches ahead of the batches, in groups
  // of batches whose queues run one after another on the caller's stream
  std::vector<JobTable> jt;         // one table per group
  std::vector<hipEvent_t> jt_done;  // per group: its queue has finished
  std::vector<int> batch_group;     // per batch: its group, -1 outside the queue
  int queue_jobs = 0;
};

extern "C" int whir_multi_destroy(whir_multi* M) {
  if (!M) return WHIR_OK;
  int rc = WHIR_OK;
  for (hipStream_t s : M->streams) release_step(rc, hipStreamSynchronize(s));
  if (M->gexec) release_step(rc, hipGraphExecDestroy(M->gexec));
  if (M->graph) release_step(rc, hipGraphDestroy(M->graph));
  for (JobTable& t : M->jt)
    if (t.dev) release_step(rc, hipFree(t.dev));
  for (hipEvent_t e : M->jt_done) release_step(rc, hipEventDestroy(e));
  for (whir_batch* b : M->batches) {
    const int rb = whir_batch_destroy(b);
    if (rc == WHIR_OK) rc = rb;
  }
  for (hipEvent_t e : M->events) release_step(rc, hipEventDestroy(e));
  if (M->done) release_step(rc, hipEventDestroy(M->done));
  for (hipStream_t s : M->streams) release_step(rc, hipStreamDestroy(s));
  delete M;
  return rc;
}

extern "C" int whir_multi_create(whir_ctx* c, const whir_params* params, int n_jobs, int streams, whir_multi** out) {
  if (!c || !params || !out || n_jobs < 1 || streams < 1 || streams > 64) return WHIR_EINVAL;
  whir_multi* M = new whir_multi();
  M->ctx = c;
  std::map<ShapeKey, std::vector<int>> groups;
  for (int i = 0; i < n_jobs; ++i) groups[shape_key(params[i])].push_back(i);
  std::vector<std::pair<uint64_t, std::vector<int>>> order;  // (workspace bytes, jobs), largest first
  for (auto& kv : groups) order.push_back({whir_prover_bytes_for(&params[kv.second[0]]), kv.second});
  std::sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  const int ns = std::min<int>(streams, (int)order.size());
  for (int i = 0; i < ns; ++i) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      whir_multi_destroy(M);
      return WHIR_EDEVICE;
    }
    M->streams.push_back(st);
  }
  for (int i = 0; i < ns + 1; ++i) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      whir_multi_destroy(M);
      return WHIR_EDEVICE;
    }
    M->events.push_back(e);
  }
  if (hipEventCreateWithFlags(&M->done, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    whir_multi_destroy(M);
    return WHIR_EDEVICE;
  }
  M->job_slot.assign((size_t)n_jobs, {-1, -1});
  std::vector<uint64_t> load((size_t)ns, 0);
  for (auto& g : order) {
    whir_batch* b = nullptr;
    int rc = whir_batch_create(c, &params[g.second[0]], (int)g.second.size(), &b);
    if (rc) {
      whir_multi_destroy(M);
      return rc;
    }
    const int bi = (int)M->batches.size();
    M->batches.push_back(b);
    const int si = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[si] += g.first * g.second.size();
    M->batch_stream.push_back(si);
    for (size_t j = 0; j < g.second.size(); ++j) M->job_slot[g.second[j]] = {bi, (int)j};
  }
  *out = M;
  return WHIR_OK;
}

extern "C" uint32_t* whir_multi_input(whir_multi* M, int job) {
  if (!M || job < 0 || job >= (int)M->job_slot.size()) return nullptr;
  return whir_batch_input(M->batches[M->job_slot[job].first], M->job_slot[job].second);
}

extern "C" int whir_multi_set_graph(whir_multi* M, int mode) {
  if (!M || mode < 0 || mode > 2) return WHIR_EINVAL;
  M->graph_mode = mode;
  for (whir_batch* b : M->batches) b->graph_mode = mode == 2;
  return WHIR_OK;
}

// Drops the captured graphs of the set and of its batches (their launch
// sequences change with the commit mode).
static int multi_drop_graphs(whir_multi* M) {
  for (hipStream_t st : M->streams) HIP_TRY(hipStreamSynchronize(st));
  if (M->done) HIP_TRY(hipEventSynchronize(M->done));
  if (M->gexec) HIP_TRY(hipGraphExecDestroy(M->gexec));
  if (M->graph) HIP_TRY(hipGraphDestroy(M->graph));
  M->gexec = nullptr;
  M->graph = nullptr;
  for (whir_batch* B : M->batches) {
    if (B->done) HIP_TRY(hipEventSynchronize(B->done));
    if (B->gexec) HIP_TRY(hipGraphExecDestroy(B->gexec));
    if (B->graph) HIP_TRY(hipGraphDestroy(B->graph));
    B->gexec = nullptr;
    B->graph = nullptr;
  }
  M->result_ok = false;
  return WHIR_OK;
}

static void multi_free_queue(whir_multi* M) {
  for (JobTable& t : M->jt)
    if (t.dev) (void)hipFree(t.dev);
  for (hipEvent_t e : M->jt_done) (void)hipEventDestroy(e);
  M->jt.clear();
  M->jt_done.clear();
  M->batch_group.clear();
  M->queue_jobs = 0;
  for (whir_batch* B : M->batches) B->P->ext_commit_levels = -1;
}

// Commit work queue. groups = 0: off. groups >= 1: every job whose shape the
// queue kernels take (folding 4: 16-word codeword rows; an LDE plan of
// 256-thread passes) gets its base LDE, leaf hashes and wide tree layers from
// persistent launches over a device job table (jobs largest first), and its
// batch starts its base tree at the cooperative stages. The eligible batches
// are split into `groups` groups of about equal commit work, one job table
// each; the groups' queues run one after another on the caller's stream and a
// group's batches start when its queue is done, so the openings of one group
// overlap the next group's commit. The proofs are byte-for-byte those of the
// batches alone. *n_jobs (optional): the number of jobs the queue covers.
extern "C" int whir_multi_set_commit_queue(whir_multi* M, int groups, int* n_jobs) {
  if (!M || groups < 0 || groups > 16) return WHIR_EINVAL;
  if (n_jobs) *n_jobs = 0;
  int rc = multi_drop_graphs(M);
  if (rc) return rc;
  multi_free_queue(M);
  if (groups == 0) return WHIR_OK;
  // eligible batches, largest commit work first, dealt to the least-loaded group
  struct Elig {
    int b;
    uint64_t work;
    int log_h, wide;
  };
  std::vector<Elig> el;
  for (size_t b = 0; b < M->batches.size(); ++b) {
    whir_prover* P = M->batches[b]->P;
    const Plan& pl = P->plan;
    if (pl.f != 4) continue;
    const int log_h = pl.k + pl.rate - pl.f;
    JtJobSpec sp{P->input, P->cw0, P->ntt_tmp, P->tree0, 1ull << (pl.k - pl.f), log_h, 0};
    std::vector<NttPass> probe;
    std::vector<uint32_t> kinds;
    if (!jt_ntt_passes(sp, M->ctx->tw_lo(), M->ctx->tw_hi(), probe, kinds)) continue;
    const int wide = merkle_wide_levels(log_h, merkle_config(M->ctx, log_h, 16, M->batches[b]->jobs));
    el.push_back({(int)b, (uint64_t)M->batches[b]->jobs << log_h, log_h, wide});
  }
  if (el.empty()) return WHIR_OK;
  std::stable_sort(el.begin(), el.end(), [](const Elig& x, const Elig& y) { return x.work > y.work; });
  const int G = std::min<int>(groups, (int)el.size());
  std::vector<uint64_t> load((size_t)G, 0);
  std::vector<std::vector<std::pair<uint64_t, JtJobSpec>>> specs((size_t)G);
  M->batch_group.assign(M->batches.size(), -1);
  for (const Elig& e : el) {
    const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[g] += e.work;
    M->batch_group[e.b] = g;
    whir_batch* B = M->batches[e.b];
    whir_prover* P = B->P;
    for (uint32_t j = 0; j < B->jobs; ++j)
      specs[g].push_back({1ull << e.log_h,
                          JtJobSpec{job_addr(P->input, j), job_addr(P->cw0, j), job_addr(P->ntt_tmp, j),
                                    job_addr(P->tree0, j), 1ull << (P->plan.k - P->plan.f), e.log_h, e.wide}});
  }
  int n_cu = 0;
  HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, M->ctx->device));
  M->jt.resize((size_t)G);
  for (int g = 0; g < G; ++g) {
    std::stable_sort(specs[g].begin(), specs[g].end(), [](const auto& x, const auto& y) { return x.first > y.first; });
    std::vector<JtJobSpec> jobs;
    for (auto& x : specs[g]) jobs.push_back(x.second);
    const int brc = jt_build(jobs, M->ctx->tw_lo(), M->ctx->tw_hi(), n_cu, M->jt[g]);
    hipEvent_t ev = nullptr;
    if (brc || hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      multi_free_queue(M);
      return brc == 3 ? WHIR_EOOM : WHIR_EDEVICE;
    }
    M->jt_done.push_back(ev);
    M->queue_jobs += (int)jobs.size();
  }
  for (const Elig& e : el) M->batches[e.b]->P->ext_commit_levels = e.wide;
  if (n_jobs) *n_jobs = M->queue_jobs;
  return WHIR_OK;
}

// Fork from s onto the internal streams, every batch's proofs, join back into
// s. Outside a capture each batch's completion is recorded on its own stream,
// so its proofs can be read while later batches still run. With the commit
// work queue on, the groups' queues run on s after the fork, each group's
// batches waiting for their group's queue.
static int multi_enqueue_all(whir_multi* M, hipStream_t s) {
  HIP_TRY(hipEventRecord(M->events[0], s));
  for (hipStream_t st : M->streams) HIP_TRY(hipStreamWaitEvent(st, M->events[0], 0));
  for (size_t g = 0; g < M->jt.size(); ++g) {
    jt_enqueue(M->jt[g], s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(M->jt_done[g], s));
  }
  for (size_t b = 0; b < M->batches.size(); ++b) {
    whir_batch* B = M->batches[b];
    hipStream_t bs = M->streams[M->batch_stream[b]];
    if (!M->batch_group.empty() && M->batch_group[b] >= 0)
      HIP_TRY(hipStreamWaitEvent(bs, M->jt_done[M->batch_group[b]], 0));
    const int rc = M->graph_mode == 2 ? whir_batch_enqueue(B, nullptr, bs) : batch_enqueue_proofs(B, bs);
    if (rc) return rc;
    if (M->graph_mode != 1) {
      HIP_TRY(hipEventRecord(B->done, bs));
      B->result_ok = true;
    }
  }
  for (size_t i = 0; i < M->streams.size(); ++i) {
    HIP_TRY(hipEventRecord(M->events[i + 1], M->streams[i]));
    HIP_TRY(hipStreamWaitEvent(s, M->events[i + 1], 0));
  }
  return WHIR_OK;
}

extern "C" int whir_multi_enqueue(whir_multi* M, const uint32_t* const* d_polys, void* stream) {
  if (!M) return WHIR_EINVAL;
  hipStream_t s = pick_stream(M->ctx, stream);
  M->result_ok = false;
  for (whir_batch* b : M->batches) b->result_ok = false;
  if (d_polys)
    for (size_t i = 0; i < M->job_slot.size(); ++i) {
      uint32_t* dst = whir_multi_input(M, (int)i);
      const whir_batch* b = M->batches[M->job_slot[i].first];
      if (d_polys[i] && d_polys[i] != dst)
        HIP_TRY(hipMemcpyAsync(dst, d_polys[i], (1ull << b->P->plan.k) * 4, hipMemcpyDeviceToDevice, s));
    }
  if (M->graph_mode != 1) {
    const int rc = multi_enqueue_all(M, s);
    if (rc) return rc;
  } else {
    if (!M->gexec) {
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const int rc = multi_enqueue_all(M, s);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s, &g);
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
      M->graph = g;
      M->gexec = ge;
    }
    HIP_TRY(hipGraphLaunch(M->gexec, s));
  }
  HIP_TRY(hipEventRecord(M->done, s));
  if (M->graph_mode == 1)
    for (whir_batch* b : M->batches) {
      b->result_ok = true;
      HIP_TRY(hipEventRecord(b->done, s));
    }
  M->result_ok = true;
  return WHIR_OK;
}

extern "C" int whir_multi_result(whir_multi* M, int job, uint8_t* proof, size_t cap, size_t* len,
                                 uint32_t* checkpoint, uint32_t* root8) {
  if (!M || !M->result_ok || job < 0 || job >= (int)M->job_slot.size()) return WHIR_EINVAL;
  return whir_batch_result(M->batches[M->job_slot[job].first], M->job_slot[job].second, proof, cap, len, checkpoint,
                           root8);
}

extern "C" int whir_multi_result_view(whir_multi* M, int job, const uint8_t** proof, size_t* len,
                                      uint32_t* checkpoint, uint32_t* root8) {
  if (!M || !M->result_ok || job < 0 || job >= (int)M->job_slot.size()) return WHIR_EINVAL;
  return whir_batch_result_view(M->batches[M->job_slot[job].first], M->job_slot[job].second, proof, len,
                                checkpoint, root8);
}

extern "C" size_t whir_multi_proof_capacity(const whir_multi* M, int job) {
  if (!M || job < 0 || job >= (int)M->job_slot.size()) return 0;
  return whir_batch_proof_capacity(M->batches[M->job_slot[job].first]);
}
