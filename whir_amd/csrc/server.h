// Proof server: continuous batching of independent WHIR proofs on one GPU.
//
// * slots: `max_slots` HIP streams; a running job owns one slot;
// * workspaces: prover arenas are keyed by proof shape and kept device-resident
//   after a job finishes (codewords, trees, eq tables and the captured hipGraph
//   of the whole proof are reused by the next job of that shape);
// * admission (FIFO): the head job starts when a slot is free and its shape's
//   workspace is idle, or a new one fits in the HBM budget after evicting idle
//   workspaces of other shapes; otherwise it waits for running jobs to finish;
// * completion: each job's last operation records an event; poll() never blocks.
// Jobs that run together overlap on the device (different streams), and each
// job's proof is computed by exactly the same kernel sequence as when alone.
#pragma once
#include <deque>
#include <map>
#include <tuple>

struct ServerJob {
  int64_t id = 0;
  whir_params params{};
  const uint32_t* d_poly = nullptr;
  int state = 0;  // 0 queued, 1 running, 2 finished
  whir_prover* prover = nullptr;
  int slot = -1;
  int status = WHIR_OK;
  std::vector<uint8_t> proof;
  uint32_t checkpoint = 0;
};

using ShapeKey = std::tuple<int, int, int, int, int, int, int>;
static ShapeKey shape_key(const whir_params& p) {
  return ShapeKey(p.k, p.folding, p.rate, p.initial_reduction, p.soundness, p.security_level, p.n_points);
}

struct whir_server {
  whir_ctx* ctx = nullptr;
  int max_slots = 0;
  uint64_t budget = 0;
  uint64_t used = 0;
  std::vector<hipStream_t> streams;
  std::vector<int64_t> slot_job;
  std::map<ShapeKey, std::vector<whir_prover*>> idle;
  std::map<int64_t, ServerJob> jobs;
  std::deque<int64_t> queued;
  int64_t next_id = 1;
  int running = 0;
};

extern "C" int whir_server_create(whir_ctx* ctx, int max_slots, uint64_t hbm_budget_bytes, whir_server** out) {
  if (!ctx || !out || max_slots < 1 || max_slots > 64) return WHIR_EINVAL;
  whir_server* S = new whir_server();
  S->ctx = ctx;
  S->max_slots = max_slots;
  S->budget = hbm_budget_bytes;
  for (int i = 0; i < max_slots; ++i) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      delete S;
      return WHIR_EDEVICE;
    }
    S->streams.push_back(st);
    S->slot_job.push_back(-1);
  }
  *out = S;
  return WHIR_OK;
}

extern "C" int whir_server_destroy(whir_server* S) {
  if (!S) return WHIR_OK;
  for (hipStream_t st : S->streams) (void)hipStreamSynchronize(st);
  for (auto& kv : S->jobs)
    if (kv.second.prover) whir_prover_destroy(kv.second.prover);
  for (auto& kv : S->idle)
    for (whir_prover* p : kv.second) whir_prover_destroy(p);
  for (hipStream_t st : S->streams) (void)hipStreamDestroy(st);
  delete S;
  return WHIR_OK;
}

extern "C" int whir_server_submit(whir_server* S, const whir_params* params, const uint32_t* d_poly,
                                  int64_t* job_id) {
  if (!S || !params || !d_poly) return WHIR_EINVAL;
  if (whir_prover_bytes_for(params) == 0) return WHIR_EINVAL;
  ServerJob j;
  j.id = S->next_id++;
  j.params = *params;
  j.d_poly = d_poly;
  S->jobs[j.id] = j;
  S->queued.push_back(j.id);
  if (job_id) *job_id = j.id;
  return WHIR_OK;
}

static bool server_evict_one(whir_server* S) {
  for (auto& kv : S->idle) {
    if (!kv.second.empty()) {
      whir_prover* p = kv.second.back();
      kv.second.pop_back();
      S->used -= p->arena_bytes;
      whir_prover_destroy(p);
      return true;
    }
  }
  return false;
}

// A workspace for the job's shape, or nullptr if it has to wait.
static whir_prover* server_acquire(whir_server* S, ServerJob& j, int& err) {
  err = WHIR_OK;
  auto it = S->idle.find(shape_key(j.params));
  if (it != S->idle.end() && !it->second.empty()) {
    whir_prover* p = it->second.back();
    it->second.pop_back();
    return p;
  }
  const uint64_t need = whir_prover_bytes_for(&j.params);
  while (S->used + need > S->budget && server_evict_one(S)) {
  }
  if (S->used + need > S->budget) {
    if (S->running == 0) err = WHIR_EOOM;  // can never fit
    return nullptr;
  }
  whir_prover* p = nullptr;
  int rc = whir_prover_create(S->ctx, &j.params, &p);
  if (rc) {
    if (rc == WHIR_EOOM && S->running > 0) return nullptr;  // retry when memory frees up
    err = rc;
    return nullptr;
  }
  whir_prover_set_graph(p, 1);
  S->used += p->arena_bytes;
  return p;
}

extern "C" int whir_server_poll(whir_server* S, int64_t* done_ids, int max, int* n_done) {
  if (!S) return WHIR_EINVAL;
  int nd = 0;
  // collect finished jobs
  for (int sl = 0; sl < S->max_slots; ++sl) {
    const int64_t id = S->slot_job[sl];
    if (id < 0) continue;
    ServerJob& j = S->jobs[id];
    const int q = whir_prover_query(j.prover);
    if (q == 7) continue;
    size_t len = 0;
    j.proof.resize(whir_prover_proof_capacity(j.prover));
    j.status = q == WHIR_OK ? whir_prover_result(j.prover, j.proof.data(), j.proof.size(), &len, &j.checkpoint,
                                                 nullptr)
                            : q;
    j.proof.resize(j.status == WHIR_OK ? len : 0);
    S->idle[shape_key(j.params)].push_back(j.prover);
    j.prover = nullptr;
    j.state = 2;
    S->slot_job[sl] = -1;
    S->running -= 1;
    if (done_ids && nd < max) done_ids[nd] = id;
    ++nd;
  }
  // admit queued jobs in arrival order
  while (!S->queued.empty()) {
    int sl = -1;
    for (int i = 0; i < S->max_slots; ++i)
      if (S->slot_job[i] < 0) {
        sl = i;
        break;
      }
    if (sl < 0) break;
    ServerJob& j = S->jobs[S->queued.front()];
    int err;
    whir_prover* p = server_acquire(S, j, err);
    if (!p) {
      if (err == WHIR_OK) break;  // wait for memory
      j.status = err;
      j.state = 2;
      S->queued.pop_front();
      if (done_ids && nd < max) done_ids[nd] = j.id;
      ++nd;
      continue;
    }
    S->queued.pop_front();
    j.prover = p;
    j.slot = sl;
    j.state = 1;
    S->slot_job[sl] = j.id;
    S->running += 1;
    int rc = whir_prover_enqueue(p, j.d_poly, S->streams[sl]);
    if (rc) return rc;
  }
  if (n_done) *n_done = nd < max ? nd : max;
  return WHIR_OK;
}

extern "C" int whir_server_take(whir_server* S, int64_t job_id, uint8_t* proof, size_t cap, size_t* len,
                                uint32_t* checkpoint) {
  if (!S) return WHIR_EINVAL;
  auto it = S->jobs.find(job_id);
  if (it == S->jobs.end() || it->second.state != 2) return WHIR_EINVAL;
  ServerJob& j = it->second;
  int st = j.status;
  if (st == WHIR_OK) {
    if (len) *len = j.proof.size();
    if (checkpoint) *checkpoint = j.checkpoint;
    if (proof) {
      if (j.proof.size() > cap) return WHIR_ETRANSCRIPT;
      memcpy(proof, j.proof.data(), j.proof.size());
    }
  }
  S->jobs.erase(it);
  return st;
}

extern "C" int whir_server_stats(const whir_server* S, int* running, int* queued, uint64_t* hbm_bytes) {
  if (!S) return WHIR_EINVAL;
  if (running) *running = S->running;
  if (queued) *queued = (int)S->queued.size();
  if (hbm_bytes) *hbm_bytes = S->used;
  return WHIR_OK;
}
