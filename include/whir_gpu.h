/* whir_gpu.h — C ABI of the MI355X (gfx950) WHIR prover hot path.
 *
 * Drop-in boundary for the reference's prover plug points (kilic/whir):
 *   - TwoAdicSubgroupDft::dft_batch / dft_algebra_batch  (crates/whir/src/whir.rs:34,59)
 *   - MerkleTree(Ext)::commit / query                    (crates/merkle/src/lib.rs:19-83)
 *   - Whir::commit + ProverStack::eval + Whir::open       (crates/whir/src/whir.rs:127-359,
 *                                                         crates/whir/src/sumcheck/prover.rs:64-90,
 *                                                         driven as in crates/whir/src/test.rs:218-235)
 * Field elements cross the boundary as uint32 words holding the Montgomery
 * representative (x * 2^32 mod p, p = 2^31 - 2^24 + 1): the words the reference
 * serializes. Extension elements are 4 consecutive words (coefficient 0 first).
 * Pointers named d_* are device pointers; `stream` is a hipStream_t (NULL = the
 * library's default stream). Every function returns a whir_status.
 */
#ifndef WHIR_GPU_H
#define WHIR_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors common::Error {Transcript, Verify, IO} (crates/common/src/error.rs:1-6)
 * plus the device-side failures. */
typedef enum {
  WHIR_OK = 0,
  WHIR_ETRANSCRIPT = 1, /* proof buffer too small / transcript overflow */
  WHIR_EVERIFY = 2,
  WHIR_EIO = 3,
  WHIR_EINVAL = 4,      /* shape or parameter outside what the reference accepts */
  WHIR_EOOM = 5,        /* device allocation failed */
  WHIR_EDEVICE = 6,     /* HIP runtime error */
  WHIR_EBUSY = 7,       /* the queried proof is still running */
  WHIR_ECHECK = 8       /* a device self-check failed (whir_prover_set_checks) */
} whir_status;

/* SecurityAssumption (crates/whir/src/params.rs:6-21). */
enum { WHIR_UNIQUE_DECODING = 0, WHIR_JOHNSON_BOUND = 1, WHIR_CAPACITY_BOUND = 2 };

/* Whir::new(k, folding, rate, initial_reduction, soundness, security_level, ..)
 * (crates/whir/src/whir.rs:103-125) plus the number of evaluation claims drawn
 * after the commitment (crates/whir/src/test.rs:225-227). */
typedef struct {
  int32_t k;
  int32_t folding;
  int32_t rate;
  int32_t initial_reduction;
  int32_t soundness;
  int32_t security_level;
  int32_t n_points;
} whir_params;

typedef struct whir_ctx whir_ctx;       /* device, Poseidon2 constants, twiddle tables */
typedef struct whir_prover whir_prover; /* device workspace for one proof shape */

/* Creates the context on `device`: uploads the Poseidon2 KoalaBear constants of
 * Poseidon2KoalaBear<16|24>::new_from_rng_128(StdRng::seed_from_u64(seed))
 * (crates/whir/src/test.rs:195-196; the reference uses seed 1000). The
 * constants live in the device's constant memory: contexts on one device share
 * them, so they must use one seed (the last created context's constants are
 * the ones in effect). Thread-safe. */
int whir_ctx_create(int device, uint64_t poseidon_seed, whir_ctx** out);
/* Destroy calls release everything; a failed HIP release returns WHIR_EDEVICE. */
int whir_ctx_destroy(whir_ctx* ctx);

/* NTT tile shape of this context (no counterpart in the reference, whose DFT
 * has no device tiling): passes of at most 2^max_log_pass rows (1..8; raised
 * to what three passes need), workgroups of up to 2^log_threads threads (6..9;
 * 512 threads only for the three-stage passes of a two-pass transform, 256 for
 * the others). Results do not depend on it. Defaults 8, 8. */
int whir_ctx_set_ntt_tuning(whir_ctx* ctx, int max_log_pass, int log_threads);
/* NTT pass layout of this context (no counterpart in the reference): for a
 * transform that needs three passes, first_pass_log in 2..4 makes the first
 * pass an LDS-free radix-2^first_pass_log register pass (0: three even passes);
 * log_vec in 0..2 sets the words per lane and line (1, 2, 4). Results do not
 * depend on it. Defaults 4, 1. */
int whir_ctx_set_ntt_shape(whir_ctx* ctx, int first_pass_log, int log_vec);

/* NTT pass count of this context (no counterpart in the reference): with
 * two_pass = 1 a transform of 2^17..2^20 rows runs in two passes of up to 1024
 * rows (three register stages each), with 2 a transform of 2^18..2^20 rows of
 * at least 16 columns in two passes of 512 or 1024 rows (two register stages of
 * up to 32 points each, 256-thread workgroups), with 0 in three passes of
 * <= 256 rows. Results do not depend on it. Default 0. */
int whir_ctx_set_ntt_passes(whir_ctx* ctx, int two_pass);

/* Merkle tree shape of this context (no counterpart in the reference, whose
 * tree runs on CPU threads): two lanes per permutation for leaf layers and
 * compression layers of at most 2^pair_max_log rows / compressions (0..24), and
 * LDS-resident cooperative subtrees (5 levels per launch) from
 * 2^coop_from_log nodes down (5..24). Results do not depend on it. Defaults 15, 15. */
int whir_ctx_set_merkle_tuning(whir_ctx* ctx, int pair_max_log, int coop_from_log);

/* Autotuner (no counterpart in the reference, whose DFT plan and tree are
 * fixed, crates/whir/src/test.rs:215-216, crates/merkle/src/poseidon_packed.rs:14):
 * with enable = 1, each whir_prover_create / _create_stacked on this context
 * times every candidate NTT configuration (pass count, first-pass radix,
 * workgroup size, vector width) and Merkle configuration (two-lane threshold,
 * cooperative-subtree depth) for each transform and tree of its proof shape not
 * tuned yet, on the new workspace's buffers, and later launches of those sizes
 * on this context use the fastest. Proof bytes do not depend on it. Default 0. */
int whir_ctx_set_autotune(whir_ctx* ctx, int enable);
/* The tuner's choices so far as a JSON array (one object per transform /
 * tree: candidates timed, the choice, its time and the default's). *len gets
 * the length without the terminating NUL; buf may be NULL to query it. */
int whir_ctx_tuning(whir_ctx* ctx, char* buf, size_t cap, size_t* len);

/* ---- component entry points ---- */

/* Replaces TwoAdicSubgroupDft::dft_batch (whir.rs:34): each of the `width`
 * columns of the row-major matrix d_in (n_in_rows rows; rows n_in_rows..2^log_h
 * are the zero padding of whir.rs:32) is transformed to 2^log_h evaluations in
 * natural order. dft_algebra_batch (whir.rs:59) is the same call with width =
 * 4 * (extension columns). d_out has 2^log_h * width words. */
int whir_dft_batch(whir_ctx* ctx, const uint32_t* d_in, uint64_t n_in_rows, uint32_t log_h, uint32_t width,
                   uint32_t* d_out, void* stream);

/* Replaces MerkleTree::commit / MerkleTreeExt::commit (crates/merkle/src/lib.rs:22-29,
 * 55-62): leaves are PaddingFreeSponge<P24,24,16,8> of each row (row_words base
 * words, extension rows flattened coefficient-major), internal nodes
 * TruncatedPermutation<P16,2,8,16> of split-half pairs. d_layers receives all
 * 2 * 2^log_h - 1 digests (8 words each), leaves first; root_host (8 words) the root. */
int whir_merkle_commit(whir_ctx* ctx, const uint32_t* d_rows, uint32_t log_h, uint32_t row_words,
                       uint32_t* d_layers, uint32_t* root_host, void* stream);

/* Poseidon2 permutation of n independent states (width 16 or 24), in place.
 * width -16 / -64: width 16 in the cooperative forms, one wavefront per state
 * (-16: 16-lane rows, as in the Merkle tree tops; -64: the device
 * transcript's one-wavefront form). */
int whir_poseidon2_permute(whir_ctx* ctx, int width, uint32_t* d_states, uint64_t n, void* stream);

/* Replaces the sorted().dedup() of stir_indexes (crates/whir/src/whir.rs:63-77):
 * the distinct values of d_in[0..n) in ascending order into d_out, their
 * number into d_count (device words). n <= 1024; values below 2^32 - 1 (STIR
 * indices have at most 31 bits). */
int whir_sort_dedup(whir_ctx* ctx, const uint32_t* d_in, uint32_t n, uint32_t* d_out, uint32_t* d_count,
                    void* stream);

/* Replaces MerkleTree::query / MerkleTreeExt::query (crates/merkle/src/lib.rs:30-37,
 * poseidon_packed.rs:145-233) for n leaf indices: per index, the row (row_words
 * words of d_rows) then the log_h sibling digests bottom-up (sibling of the
 * node at level i = (index mod 2^(log_h-i)) ^ 2^(log_h-i-1)), i.e. the hint
 * words the reference writes. d_out: n * (row_words + 8 log_h) words. Indices
 * >= 2^log_h leave their slot untouched. */
int whir_merkle_open(whir_ctx* ctx, const uint32_t* d_layers, uint32_t log_h, const uint32_t* d_rows,
                     uint32_t row_words, const uint32_t* d_idx, uint32_t n, uint32_t* d_out, void* stream);

/* ---- sumcheck / eq functions (extension elements are 4 Montgomery words,
 * coefficient-major; points are k extension elements, bit j <-> z_j) ---- */

/* Replaces Point::eq / eq_packed (crates/poly/src/point.rs:110-126):
 * d_out[x] = scale * prod_j (x_j ? z_j : 1 - z_j), 2^k entries, k <= 32. */
int whir_eq_table(whir_ctx* ctx, const uint32_t* d_z, uint32_t k, const uint32_t scale[4], uint32_t* d_out,
                  void* stream);

/* Replaces SplitEq::eval_ext (crates/poly/src/eq.rs:179-250) / Poly::eval_ext
 * (poly.rs:89-95): d_out (1 ext) = <poly, eq(z)> for a 2^k extension polynomial. */
int whir_eval_ext(whir_ctx* ctx, const uint32_t* d_poly, uint32_t k, const uint32_t* d_z, uint32_t* d_out,
                  void* stream);

/* Replaces SplitEq::compress_hi (crates/poly/src/eq.rs:290-398): binds the top m
 * of the k variables of a base polynomial to z (m coordinates), d_out has
 * 2^(k-m) <= 32 extension values; m = k is SplitEq::eval_base (eq.rs:107-177). */
int whir_compress_hi(whir_ctx* ctx, const uint32_t* d_poly, uint32_t k, const uint32_t* d_z, uint32_t m,
                     uint32_t* d_out, void* stream);

/* No counterpart in the reference: whir_compress_hi for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). It allocates and releases nothing; its
 * scratch (the point's split eq tables and the per-block partial sums) is the
 * library's static device array of WHIR_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends. Calls needing more (m > 28 when k - m = 5) return
 * WHIR_EINVAL. Every call on one GPU shares that array: calls that overlap on
 * one GPU (other streams, other threads) race and must be serialised by the
 * caller. Same arguments and result as whir_compress_hi. */
#define WHIR_NOALLOC_SCRATCH_BYTES (16u << 20)
int whir_compress_hi_noalloc(whir_ctx* ctx, const uint32_t* d_poly, uint32_t k, const uint32_t* d_z, uint32_t m,
                             uint32_t* d_out, void* stream);

/* Replaces SplitEq::compress_lo (crates/poly/src/eq.rs:400-539): binds the low
 * m <= 5 variables of a base polynomial; d_out has 2^(k-m) extension values. */
int whir_compress_lo(whir_ctx* ctx, const uint32_t* d_poly, uint32_t k, const uint32_t* d_z, uint32_t m,
                     uint32_t* d_out, void* stream);

/* Replaces sumcheck coeffs (crates/whir/src/sumcheck/expr.rs:14-27): for two 2^k
 * extension arrays, d_c0c2 = (sum w[2i] f[2i], sum (2w[2i+1]-w[2i])(2f[2i+1]-f[2i])). */
int whir_sumcheck_coeffs(whir_ctx* ctx, const uint32_t* d_f, const uint32_t* d_w, uint32_t k, uint32_t* d_c0c2,
                         void* stream);

/* Replaces Poly::fix_lo_var_mut (crates/poly/src/poly.rs:131-142):
 * d_out[i] = a[2i] + r (a[2i+1] - a[2i]) for a 2^k extension array. */
int whir_fix_lo_var(whir_ctx* ctx, const uint32_t* d_poly, uint32_t k, const uint32_t r[4], uint32_t* d_out,
                    void* stream);

/* Replaces combine_eqs (crates/whir/src/sumcheck/combine.rs:178-203):
 * d_w[x] += sum_j alpha^j eq(points_j)[x] over n points of k coordinates. */
int whir_combine_eqs(whir_ctx* ctx, uint32_t* d_w, uint32_t k, const uint32_t* d_points, uint32_t n,
                     const uint32_t alpha[4], void* stream);

/* Replaces combine_pows (crates/whir/src/sumcheck/combine.rs:407-438):
 * d_w[i] += sum_j alpha^(shift+j) vars_j^i over n base-field vars. */
int whir_combine_pows(whir_ctx* ctx, uint32_t* d_w, uint32_t k, const uint32_t* d_vars, uint32_t n,
                      const uint32_t alpha[4], uint64_t shift, void* stream);

/* Replaces combine_pows_scaled (crates/whir/src/sumcheck/combine.rs:440-468):
 * d_w[i] += sum_j scale_j vars_j^i over n base-field vars with n extension
 * scales given on the device (d_scales: n x 4 words). */
int whir_combine_pows_scaled(whir_ctx* ctx, uint32_t* d_w, uint32_t k, const uint32_t* d_vars, uint32_t n,
                             const uint32_t* d_scales, void* stream);

/* ---- whole prover (Whir::commit + claims + Whir::open) ---- */

int whir_prover_create(whir_ctx* ctx, const whir_params* params, whir_prover** out);
int whir_prover_destroy(whir_prover* prover);
/* Upper bound of the serialized proof length for this shape, in bytes. */
size_t whir_prover_proof_capacity(const whir_prover* prover);
/* Device buffer of 2^k words owned by the prover, usable as d_poly. */
uint32_t* whir_prover_input(whir_prover* prover);
/* Enqueues the whole proof of d_poly (2^k base coefficients) on `stream`; no host
 * synchronisation. The proof is the PoseidonWriter byte stream of
 * run_whir_poseidon's prover (crates/whir/src/test.rs:218-235). */
int whir_prover_enqueue(whir_prover* prover, const uint32_t* d_poly, void* stream);
/* Waits for the enqueued proof and copies it out: `proof` (capacity `cap` bytes),
 * its length, the test's checkpoint draw (test.rs:233) and the commitment root. */
int whir_prover_result(whir_prover* prover, uint8_t* proof, size_t cap, size_t* len, uint32_t* checkpoint,
                       uint32_t* root8);
/* The same result without a copy: *proof points at the proof in the prover's
 * pinned host buffer, valid until the next enqueue on the prover. */
int whir_prover_result_view(whir_prover* prover, const uint8_t** proof, size_t* len, uint32_t* checkpoint,
                            uint32_t* root8);

/* ---- Whir::commit / ProverStack::eval / Whir::open with the caller's transcript ----
 * The reference threads one transcript through three calls
 * (crates/whir/src/whir.rs:127-138 commit, crates/whir/src/sumcheck/prover.rs:64-90
 * ProverStack::eval, whir.rs:168-180 open). These entry points do the same on
 * the device: the transcript is a DuplexChallenger<KoalaBear, Poseidon2<16>, 16, 8>
 * state handed in before the commitment and handed back after the opening.
 *
 * whir_challenger mirrors Plonky3's DuplexChallenger fields (sponge_state,
 * input_buffer, output_buffer; Montgomery words). A state is accepted when a
 * challenger can reach it: input_len <= 7, output_len <= 8, not both non-zero
 * (observe clears the output), every word canonical (< p), and
 * output_buffer[0..output_len) == sponge_state[0..output_len) (the output
 * buffer is the state's rate prefix after a duplex; sample pops from its end).
 * A zeroed struct is Challenger::new. */
typedef struct {
  uint32_t sponge_state[16];
  uint32_t input_buffer[8];
  uint32_t output_buffer[8];
  uint32_t input_len;
  uint32_t output_len;
} whir_challenger;

/* Whir::commit (whir.rs:127-138 -> commit_base, whir.rs:19-36): RS encode and
 * Merkle-commit d_poly (2^k words; for a stacked prover the stacked vector, see
 * whir_prover_stack_offset), then transcript.write(root). tx_in: the caller's
 * transcript (NULL: a fresh one). d_poly stays in use until whir_prover_open
 * has finished. Enqueued on `stream`, no host synchronisation. */
int whir_prover_commit(whir_prover* prover, const uint32_t* d_poly, const whir_challenger* tx_in, void* stream);
/* ProverStack::eval (prover.rs:64-90) for the next evaluation claim of the
 * prover's claim list: computes the value at the claim's point and writes it.
 * point (host, k_poly extension elements = 4 k_poly words): any point; NULL
 * draws point = expand(k_poly, transcript.draw()) as the reference's tests do
 * (crates/whir/src/test.rs:225-226). A claim that shares the previous claim's
 * point (whir_stack claim_fresh = 0) takes NULL. */
int whir_prover_eval(whir_prover* prover, const uint32_t* point, void* stream);
/* Whir::open (whir.rs:168-359) after every claim has been evaluated. Then
 * whir_prover_result returns the bytes written since the commitment's transcript
 * state (root, claim values, opening), root8 the commitment root, and
 * `checkpoint` the value the caller's next transcript.draw() returns. */
int whir_prover_open(whir_prover* prover, void* stream);
/* The transcript state after the opening (waits for it). */
int whir_prover_challenger(whir_prover* prover, whir_challenger* out);

/* ---- multi-polynomial stacking ----
 * Several multilinear polys committed and opened as one (ProverStack::register /
 * eval / layout, crates/whir/src/sumcheck/prover.rs:31-124; VerifierStack /
 * VerifierLayout, verifier.rs:100-212; Selector, mod.rs:30-58). The stacked
 * vector is the layout's poly (prover.rs:114-118): polys sorted by size, the
 * largest first (for equal sizes the last registered first), zero-padded to
 * 2^k, k = ceil(log2(sum of sizes)). The transcript is: the stacked
 * commitment's root, then each evaluation claim in order (a fresh point
 * expand(k_poly, draw()) or the previous claim's point, then its value), then
 * Whir::open(.., layout) (whir.rs:168-359). */
typedef struct {
  int32_t n_polys;
  const int32_t* k_polys;      /* log2 size of each registered poly, registration order; >= folding */
  int32_t n_claims;
  const int32_t* claim_poly;   /* poly id of each evaluation claim, transcript order */
  const uint8_t* claim_fresh;  /* 1: the claim draws its point; 0: the previous claim's point (same size) */
} whir_stack;

/* params->k must be the stacked k; params->n_points is ignored (n_claims is used). */
int whir_prover_create_stacked(whir_ctx* ctx, const whir_params* params, const whir_stack* stack,
                               whir_prover** out);
/* Offset (words) of poly `id` in whir_prover_input(prover); -1 if out of range. */
int64_t whir_prover_stack_offset(const whir_prover* prover, int32_t id);
/* Lays out d_polys[id] (2^k_polys[id] words each, device) into the prover's
 * stacked input with the zero tail (prover.rs:114-118), then enqueues the
 * proof of that vector as whir_prover_enqueue does. */
int whir_prover_enqueue_stacked(whir_prover* prover, const uint32_t* const* d_polys, void* stream);

/* Capture the whole proof (kernels + result copies) as a hipGraph on the first
 * enqueue and replay it afterwards; the graph proves the prover's own input
 * buffer (other d_poly are first copied into it on the stream). */
int whir_prover_set_graph(whir_prover* prover, int enable);
/* Device self-checks of each proof, the reference's debug assertions on the
 * prover path (MerkleTree::query's opening re-verification, crates/merkle/src/
 * poseidon_packed.rs:158-230; the sumcheck invariant sum == prod(), crates/whir/
 * src/sumcheck/expr.rs:68, for the rounds on the grid after round 0): extra
 * kernels beside the proof's own, the proof bytes unchanged. With checks on, a
 * result call reports WHIR_ECHECK (the proof still copied out) when one failed.
 * Off by default. */
int whir_prover_set_checks(whir_prover* prover, int enable);
/* WHIR_OK once the last enqueued proof has finished, WHIR_EBUSY while running. */
int whir_prover_query(whir_prover* prover);
/* Device bytes held by this prover / needed by a prover of this shape. */
size_t whir_prover_device_bytes(const whir_prover* prover);
size_t whir_prover_bytes_for(const whir_params* params);

/* Optional device timers: HIP events recorded on the prover's stream around one
 * kernel class of each proof (the reference's tracing spans "dft-base" and
 * "leaf_hash", crates/whir/src/whir.rs:33, crates/merkle/src/poseidon_packed.rs:285).
 * whir_prover_timer reports the accumulated milliseconds, the number of timed
 * proofs and the work per launch (Poseidon2-24 permutations for the leaf hash,
 * input + output words for the DFT). Enabling resets the accumulators. */
enum { WHIR_TIMER_LEAVES_BASE = 0, WHIR_TIMER_DFT_BASE = 1, WHIR_TIMER_COUNT = 2 };
int whir_prover_set_timers(whir_prover* prover, int enable);
int whir_prover_timer(const whir_prover* prover, int which, double* total_ms, uint64_t* launches,
                      uint64_t* work_units);

/* Host-to-host convenience: upload, prove, download. */
int whir_prove_host(whir_ctx* ctx, const whir_params* params, const uint32_t* poly, uint8_t* proof, size_t cap,
                    size_t* len, uint32_t* checkpoint, uint32_t* root8);

/* ---- proof server: continuous batching of independent proofs ----
 * No counterpart in the reference (its tests prove one polynomial at a time);
 * this is the serving form of the same prover. Jobs run on `max_slots` streams
 * and overlap on the device (on MI355X, k 16..22 jobs: 4 slots 1.66x the
 * one-by-one rate, 6 slots 2.48x, more no better); workspaces stay
 * device-resident per proof shape (graph-captured once); admission is FIFO and
 * gated on hbm_budget_bytes.
 * A job's proof bytes are the same as when it is proven alone. */
typedef struct whir_server whir_server;
int whir_server_create(whir_ctx* ctx, int max_slots, uint64_t hbm_budget_bytes, whir_server** out);
int whir_server_destroy(whir_server* server);
/* Queues a proof of the 2^k words at d_poly (device memory the caller keeps
 * valid until the job has finished). */
int whir_server_submit(whir_server* server, const whir_params* params, const uint32_t* d_poly, int64_t* job_id);
/* Collects finished jobs (up to `max` ids) and admits queued ones; never blocks. */
int whir_server_poll(whir_server* server, int64_t* done_ids, int max, int* n_done);
/* Proof of a finished job, which is then forgotten; returns the job's status. */
int whir_server_take(whir_server* server, int64_t job_id, uint8_t* proof, size_t cap, size_t* len,
                     uint32_t* checkpoint);
int whir_server_stats(const whir_server* server, int* running, int* queued, uint64_t* hbm_bytes);

/* ---- seeded inputs ----
 * Replaces Poly::<F>::rand(&mut rng(seed), k) (crates/poly/src/poly.rs:47-52,
 * rng = StdRng::seed_from_u64, crates/common/src/test.rs:1-4): out (host, 2^k
 * words) receives the polynomial the reference's tests prove for that seed. */
int whir_poly_rand(uint64_t seed, uint32_t k, uint32_t* out);

/* ---- lockstep batches of same-shape proofs ----
 * No counterpart in the reference (its tests prove one polynomial at a time).
 * A batch holds `jobs` workspaces of one proof shape and proves all of them as
 * one kernel sequence: every kernel is launched once for the batch, one slice
 * of its grid per job (the launches of a single proof for the whole batch, the
 * transcripts' serial chains side by side). Each job's proof bytes are those of
 * the job proven alone. The inputs are the batch's own device buffers
 * (whir_batch_input) or are copied from d_polys. */
typedef struct whir_batch whir_batch;
int whir_batch_create(whir_ctx* ctx, const whir_params* params, int jobs, whir_batch** out);
int whir_batch_destroy(whir_batch* batch);
/* Device buffer of 2^k words holding job `job`'s polynomial. */
uint32_t* whir_batch_input(whir_batch* batch, int job);
/* Capture the batch's proof as a hipGraph on the first enqueue, replay it after. */
int whir_batch_set_graph(whir_batch* batch, int enable);
/* whir_prover_set_checks for every job of the batch. */
int whir_batch_set_checks(whir_batch* batch, int enable);
/* Enqueues every job's proof on `stream` (d_polys: NULL, or per job a device
 * polynomial copied into the job's input first, NULL entries keep the input). */
int whir_batch_enqueue(whir_batch* batch, const uint32_t* const* d_polys, void* stream);
/* WHIR_OK once the last enqueued batch has finished, WHIR_EBUSY while running. */
int whir_batch_query(whir_batch* batch);
/* Job `job`'s proof, its checkpoint draw and commitment root (waits). */
int whir_batch_result(whir_batch* batch, int job, uint8_t* proof, size_t cap, size_t* len, uint32_t* checkpoint,
                      uint32_t* root8);
/* The same result without a copy: *proof points at the job's proof in the
 * library's pinned host buffer, valid until the next enqueue on the batch. */
int whir_batch_result_view(whir_batch* batch, int job, const uint8_t** proof, size_t* len, uint32_t* checkpoint,
                           uint32_t* root8);
size_t whir_batch_proof_capacity(const whir_batch* batch);
size_t whir_batch_device_bytes(const whir_batch* batch);

/* Many independent jobs of mixed shapes (no counterpart in the reference): the
 * jobs (params[0..n_jobs), one claim count each) are grouped by shape into
 * lockstep batches, the batches run side by side on up to `streams` internal
 * streams, forked from and joined back into the caller's stream. Graph mode 1
 * captures the whole set once as one hipGraph (each enqueue is one graph
 * launch); mode 2 captures each batch as its own graph, launched on its
 * stream; 0 launches eagerly. Job i's bytes are those of job i proven alone. */
typedef struct whir_multi whir_multi;
int whir_multi_create(whir_ctx* ctx, const whir_params* params, int n_jobs, int streams, whir_multi** out);
int whir_multi_destroy(whir_multi* multi);
uint32_t* whir_multi_input(whir_multi* multi, int job);
int whir_multi_set_graph(whir_multi* multi, int mode);
/* Commit work queue (replaces, for the whole set, the per-job calls of
 * Whir::commit's encode and tree, crates/whir/src/whir.rs:19-36 and
 * crates/merkle/src/poseidon_packed.rs:96-143, 268-304): with groups >= 1 every
 * job of folding 4 has its base LDE, leaf hashes and wide tree layers computed
 * by persistent kernels that drain a device job table (one launch per LDE pass
 * and per tree layer, items dequeued with an atomic counter); the batches go
 * on from the cooperative top of each tree. The batches are split into
 * `groups` groups of about equal commit work whose queues run one after
 * another, each group's batches starting when its queue is done (the openings
 * of one group overlap the next group's commit). groups = 0 turns it off.
 * Bytes are unchanged. Drops captured graphs. *n_jobs (may be NULL): the
 * number of jobs the queue covers. */
int whir_multi_set_commit_queue(whir_multi* multi, int groups, int* n_jobs);
int whir_multi_enqueue(whir_multi* multi, const uint32_t* const* d_polys, void* stream);
int whir_multi_result(whir_multi* multi, int job, uint8_t* proof, size_t cap, size_t* len, uint32_t* checkpoint,
                      uint32_t* root8);
int whir_multi_result_view(whir_multi* multi, int job, const uint8_t** proof, size_t* len, uint32_t* checkpoint,
                           uint32_t* root8);
size_t whir_multi_proof_capacity(const whir_multi* multi, int job);

/* ---- device memory helpers (plain HIP allocations) ---- */
int whir_device_alloc(size_t bytes, void** out);
int whir_device_free(void* ptr);
int whir_memcpy_h2d(void* dst, const void* src, size_t bytes);
int whir_memcpy_d2h(void* dst, const void* src, size_t bytes);
int whir_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int whir_device_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif /* WHIR_GPU_H */
