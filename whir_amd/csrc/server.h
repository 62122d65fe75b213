// Proof server: continuous batching of independent WHIR proofs on one GPU.
//
// * slots: `max_slots` HIP streams; a running job owns one slot;
// * workspaces: prover arenas are keyed by proof shape and kept device-resident
//   after a job finishes (codewords, trees, eq tables and the captured hipGraph
//   of the whole proof are reused by the next job of that shape);
// * admission (FIFO): the head job starts when a slot is free and its shape's
//   workspace is idle, or a new one fits in the HBM budget, or an idle
//   workspace of another shape is large enough to be re-laid out for it;
//   otherwise idle workspaces are released when nothing runs (a release waits
//   for the whole device), and until then the job waits;
// * completion: each job's last operation records an event; poll() never
//   blocks, and reports every finished job exactly once.
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
  std::deque<int64_t> done_pending;     // finished jobs not yet reported by poll
  std::vector<whir_prover*> retired;    // workspaces to release at the next idle point
  std::vector<void*> retired_host;      // pinned host logs to release at the next idle point
  int64_t next_id = 1;
  int running = 0;
};

// Device and pinned-host releases wait for the whole device, so the server
// makes them only when none of its jobs is running.
static void server_release_retired(whir_server* S) {
  for (whir_prover* p : S->retired) whir_prover_destroy(p);
  S->retired.clear();
  for (void* h : S->retired_host) (void)hipHostFree(h);
  S->retired_host.clear();
}

extern "C" int whir_server_create(whir_ctx* ctx, int max_slots, uint64_t hbm_budget_bytes, whir_server** out) {
  if (!ctx || !out || max_slots < 1 || max_slots > 64) return WHIR_EINVAL;
  whir_server* S = new whir_server();
  S->ctx = ctx;
  S->max_slots = max_slots;
  S->budget = hbm_budget_bytes;
  for (int i = 0; i < max_slots; ++i) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      for (hipStream_t made : S->streams) (void)hipStreamDestroy(made);
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
  server_release_retired(S);
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

static void server_retire(whir_server* S, whir_prover* p) {
  S->used -= p->arena_bytes;
  S->retired.push_back(p);
}

// A workspace for the job's shape, or nullptr if it has to wait (err ==
// WHIR_OK) or can never run (err set). In order: an idle workspace of the
// shape; a new one within the budget; an idle workspace of another shape whose
// arena holds this shape, re-laid out in place (no device allocation or
// release); and, only when no job is running, idle workspaces released until a
// new one fits.
static whir_prover* server_acquire(whir_server* S, ServerJob& j, int& err) {
  err = WHIR_OK;
  auto it = S->idle.find(shape_key(j.params));
  if (it != S->idle.end() && !it->second.empty()) {
    whir_prover* p = it->second.back();
    it->second.pop_back();
    return p;
  }
  const uint64_t need = whir_prover_bytes_for(&j.params);
  if (S->used + need > S->budget) {
    // the smallest idle arena that holds the new shape
    std::vector<whir_prover*>* best_list = nullptr;
    size_t best_i = 0;
    for (auto& kv : S->idle)
      for (size_t i = 0; i < kv.second.size(); ++i)
        if (kv.second[i]->arena_bytes >= need &&
            (!best_list || kv.second[i]->arena_bytes < (*best_list)[best_i]->arena_bytes)) {
          best_list = &kv.second;
          best_i = i;
        }
    if (best_list) {
      whir_prover* old = (*best_list)[best_i];
      whir_prover* p = prover_relayout(old, &j.params, S->retired_host);
      if (p) {
        best_list->erase(best_list->begin() + best_i);
        whir_prover_set_graph(p, 1);
        return p;
      }
    }
    if (S->running > 0) return nullptr;  // wait: releasing memory now would stall the running jobs
    server_release_retired(S);
    for (auto& kv : S->idle) {
      while (!kv.second.empty() && S->used + need > S->budget) {
        whir_prover* p = kv.second.back();
        kv.second.pop_back();
        S->used -= p->arena_bytes;
        whir_prover_destroy(p);
      }
    }
    if (S->used + need > S->budget) {
      err = WHIR_EOOM;  // can never fit
      return nullptr;
    }
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

// Collects finished jobs and admits queued ones; never blocks. Every finished
// job (also one that failed to start) is reported by exactly one poll; ids
// beyond `max` stay queued for the next one.
extern "C" int whir_server_poll(whir_server* S, int64_t* done_ids, int max, int* n_done) {
  if (n_done) *n_done = 0;
  if (!S) return WHIR_EINVAL;
  if (!done_ids) max = 0;
  // collect finished jobs
  for (int sl = 0; sl < S->max_slots; ++sl) {
    const int64_t id = S->slot_job[sl];
    if (id < 0) continue;
    ServerJob& j = S->jobs[id];
    const int q = whir_prover_query(j.prover);
    if (q == WHIR_EBUSY) continue;
    size_t len = 0;
    if (q == WHIR_OK) {
      j.proof.resize(whir_prover_proof_capacity(j.prover));
      j.status = whir_prover_result(j.prover, j.proof.data(), j.proof.size(), &len, &j.checkpoint, nullptr);
    } else {
      j.status = q;
    }
    j.proof.resize(j.status == WHIR_OK ? len : 0);
    if (j.status == WHIR_EDEVICE || q != WHIR_OK) server_retire(S, j.prover);  // not reused after a device error
    else S->idle[shape_key(j.params)].push_back(j.prover);
    j.prover = nullptr;
    j.state = 2;
    S->slot_job[sl] = -1;
    S->running -= 1;
    S->done_pending.push_back(id);
  }
  if (S->running == 0) server_release_retired(S);
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
      S->done_pending.push_back(j.id);
      continue;
    }
    S->queued.pop_front();
    const int rc = whir_prover_enqueue(p, j.d_poly, S->streams[sl]);
    if (rc) {  // reported as the job's status; the workspace is not reused
      server_retire(S, p);
      j.status = rc;
      j.state = 2;
      S->done_pending.push_back(j.id);
      continue;
    }
    j.prover = p;
    j.slot = sl;
    j.state = 1;
    S->slot_job[sl] = j.id;
    S->running += 1;
  }
  int nd = 0;
  while (nd < max && !S->done_pending.empty()) {
    done_ids[nd++] = S->done_pending.front();
    S->done_pending.pop_front();
  }
  if (n_done) *n_done = nd;
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
  // a job taken before poll reported it is not reported afterwards
  for (auto d = S->done_pending.begin(); d != S->done_pending.end(); ++d)
    if (*d == job_id) {
      S->done_pending.erase(d);
      break;
    }
  return st;
}

extern "C" int whir_server_stats(const whir_server* S, int* running, int* queued, uint64_t* hbm_bytes) {
  if (!S) return WHIR_EINVAL;
  if (running) *running = S->running;
  if (queued) *queued = (int)S->queued.size();
  if (hbm_bytes) *hbm_bytes = S->used;
  return WHIR_OK;
}
