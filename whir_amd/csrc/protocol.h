// Transcript-ordered ("serial") kernels of the WHIR prover. Each one is a single
// workgroup whose first wavefront runs the device transcript (tx.h) while the
// other threads do the small parallel arithmetic around it. Together with the
// data-parallel kernels they reproduce Whir::commit / Whir::open
// (crates/whir/src/whir.rs:127-359) and ProverStack / ProverLayout /
// SumcheckProver (crates/whir/src/sumcheck/prover.rs:31-561) without any host
// round trip.
#pragma once
#include "field.h"
#include "tx.h"
#include "sumcheck.h"
#include "merkle.h"

namespace kb {

constexpr int MAX_K = 32;
constexpr int MAX_F = 5;
constexpr int MAX_CLAIMS = 1024;  // evaluation + virtual (OOD) claims of one proof
constexpr int MAX_STIR = 1024;
constexpr int SERIAL_THREADS = 256;

struct ClaimSlot {
  Ext point[MAX_K];
  Ext compressed[1 << MAX_F];
  Ext eval;
  uint32_t k;
  uint32_t pad[3];
};

struct JobScalars {
  Ext alpha;
  Ext rs[MAX_K + 8];  // sumcheck challenges in draw order (SVO first)
  Ext eq_rs[1 << MAX_F];
  uint32_t pending_words;  // proof words reserved by the last opening kernel
  uint32_t stir_count;
  uint32_t checkpoint;
  uint32_t pad;
  Ext chk_prod, chk_c0, chk_c2;  // self-checks: <f, w>, c0, c2 of the last checked arrays
};

// Per-query state (structure of arrays, MAX_STIR entries each).
struct QueryArrays {
  JP<uint32_t> index;
  JP<uint32_t> ws;     // omega^index, squared each fold round (prover.rs:463-469)
  JP<uint32_t> ts;     // omega^-(index mod mid)            (prover.rs:471-481)
  JP<uint32_t> var;    // pow-claim variable omega^index     (whir.rs:242-246)
  JP<Ext> scale;       // alpha^(n_eq + q), updated by fold rounds
  JP<Ext> G;           // fold-with-code values, 2^d per query
};


__device__ __forceinline__ bool is_wave0() { return threadIdx.x < 64; }

__device__ __forceinline__ void tx_begin(WaveTx& tx, DevTx* t, uint32_t* log, JobScalars* js,
                                         const CoopLane* pre = nullptr) {
  tx.load(t, log, pre);
  const uint32_t pend = js->pending_words;
  tx.log_words += pend;
  if (tx.log_words > tx.log_cap) tx.overflow = 1;
}
__device__ __forceinline__ void tx_end(WaveTx& tx, DevTx* t, JobScalars* js) {
  tx.store(t);
  if (threadIdx.x == 0) js->pending_words = 0;
}

// lagrange_weights_012 (sumcheck/mod.rs:17-23)
__device__ __forceinline__ void lagrange012(const Ext& r, Ext (&l)[3]) {
  const Ext one = eone(), two = efrom(TWO);
  l[0] = emul_base(emul(esub(r, one), esub(r, two)), HALF);
  l[1] = emul(r, esub(two, r));
  l[2] = emul_base(emul(r, esub(r, one)), HALF);
}

// Where the next claim's draw goes (a claim kernel can finish one claim and
// begin the next in the same launch).
struct ClaimBegin {
  JP<ClaimSlot> slot;  // first slot of the group; nullptr: nothing to begin
  JP<Ext> eq0;
  JP<Ext> eq1;
  uint32_t k, fsv, m0, m1;
  uint32_t count;   // claims sharing the point (slots slot[0 .. count))
};

struct ClaimScratch {
  Ext pt[MAX_K];
  EqPairScratch eq;
};

// Extension squaring spread over the four lanes of a quad: lane q holds and
// returns coefficient q. With x^4 = W, coefficient q of a^2 is
//   sum_{i <= q} a_i a_(q-i)  +  W sum_{i > q} a_i a_(q-i+4),
// each sum below 4 p^2, so one reduce_sum4 each; the factors come from DPP
// quad broadcasts and rotations (no memory). The same canonical words as esqr
// with half its instructions and two short chains instead of one long one
// (esqr on one lane is ~260 ns, and a claim point takes k in a row).
constexpr int DPP_QUAD_BCAST(int i) { return i | (i << 2) | (i << 4) | (i << 6); }
constexpr int DPP_QUAD_PREV = 3 | (0 << 2) | (1 << 4) | (2 << 6);  // lane a <- a-1 (mod 4)
__device__ __forceinline__ uint32_t esqr_quad(uint32_t a, uint32_t q) {
  const uint32_t a0 = dpp<DPP_QUAD_BCAST(0)>(a), a1 = dpp<DPP_QUAD_BCAST(1)>(a);
  const uint32_t a2 = dpp<DPP_QUAD_BCAST(2)>(a), a3 = dpp<DPP_QUAD_BCAST(3)>(a);
  // b_i = a_((q - i) mod 4)
  const uint32_t b1 = dpp<DPP_QUAD_PREV>(a), b2 = dpp<DPP_QUAD_XOR2>(a), b3 = dpp<DPP_QUAD_NEXT>(a);
  const uint32_t u1 = q >= 1 ? b1 : 0u, u2 = q >= 2 ? b2 : 0u, u3 = q >= 3 ? b3 : 0u;
  const uint32_t v1 = b1 - u1, v2 = b2 - u2, v3 = b3 - u3;  // the wrapped factors, 0 when unwrapped
  const uint64_t U = (uint64_t)a0 * a + (uint64_t)a1 * u1 + (uint64_t)a2 * u2 + (uint64_t)a3 * u3;
  const uint64_t V = (uint64_t)a1 * v1 + (uint64_t)a2 * v2 + (uint64_t)a3 * v3;
  const uint32_t u = reduce_sum4(U), v = reduce_sum4(V);
  return add(u, mul_w(v));
}

// pt[i] = z^(2^i) for i < k (expand, whir.rs:227-229 and test.rs:225-226),
// called by every lane of one wavefront with the same z; lanes 0..3 store.
__device__ __forceinline__ void expand_point_wave(const Ext& z, uint32_t k, Ext* pt) {
  const uint32_t q = threadIdx.x & 3;
  uint32_t a = q == 0 ? z.c[0] : q == 1 ? z.c[1] : q == 2 ? z.c[2] : z.c[3];
  for (uint32_t i = 0; i < k; ++i) {
    if ((threadIdx.x & 63) < 4) pt[i].c[q] = a;
    a = esqr_quad(a, q);
  }
}

// A claim point given by the caller (ProverStack::eval takes any point,
// prover.rs:64-90): k extension coordinates, passed by value.
struct GivenPoint {
  Ext z[MAX_K];
};

// Draw the claim point: point = expand(k, draw()) (test.rs:225-226, prover.rs:240,
// whir.rs:227-229), or take the caller's `given` point (no draw), then build
// the eq tables of point[fsv..k) split at m0.
__device__ void claim_begin_body(WaveTx& tx, const ClaimBegin& cb, ClaimScratch& sc,
                                 const GivenPoint* given = nullptr) {
  if (given) {
    for (uint32_t i = threadIdx.x; i < cb.k * 4; i += blockDim.x) sc.pt[i / 4].c[i % 4] = given->z[i / 4].c[i % 4];
  } else if (is_wave0()) {
    const Ext z = tx.sample_ext();
    expand_point_wave(z, cb.k, sc.pt);
  }
  __syncthreads();
  for (uint32_t it = threadIdx.x; it < cb.count * cb.k; it += blockDim.x) cb.slot[it / cb.k].point[it % cb.k] = sc.pt[it % cb.k];
  for (uint32_t c = threadIdx.x; c < cb.count; c += blockDim.x) cb.slot[c].k = cb.k;
  build_eq_pair(cb.eq0, sc.pt + cb.fsv, (int)cb.m0, eone(), cb.eq1, sc.pt + cb.fsv + cb.m0, (int)cb.m1, eone(),
                sc.eq);
}

// transcript.write(root) (poseidon_packed.rs:136-140): 8 hint words, then
// observe the 8 elements. Wave 0 only.
__device__ __forceinline__ void write_root_body(WaveTx& tx, const uint32_t* root) {
  const uint32_t v = root[threadIdx.x & 7];
  for (int j = 0; j < 8; ++j) tx.hint_word((uint32_t)__shfl((int)v, j, 8));
  for (int j = 0; j < 8; ++j) tx.observe((uint32_t)__shfl((int)v, j, 8));
}

// The last cooperative launch of a commitment's tree (the top n0 <= 32 nodes
// down to the root, merkle_coop_subtrees_kernel with G = 1), then its
// transcript step in the same launch: transcript.write(root) and, when cb.slot
// is set, the first claim's draw (claim_begin_body). fresh_cap != 0 starts a
// new transcript with a log of that many words (the proof's first commitment).
static_assert(COOP_THREADS == SERIAL_THREADS, "one workgroup shape for the tree top and the claim");
__global__ void __launch_bounds__(SERIAL_THREADS) tree_top_claim_kernel(JP<uint32_t> layer_, uint32_t n0, JP<DevTx> t_,
                                                                        JP<uint32_t> log_, JP<JobScalars> js_,
                                                                        ClaimBegin cb, uint32_t fresh_cap) {
  uint32_t* layer = layer_;
  DevTx* t = t_;
  uint32_t* log = log_;
  JobScalars* js = js_;
  __shared__ uint32_t s_node[COOP_MAX_NODES * 8];
  __shared__ ClaimScratch sc;
  CoopP16 K;  // the tree's compressions and then the transcript's
  K.load();
  coop_subtrees_body(layer, n0, 1, 0, K, s_node);
  WaveTx tx;
  if (is_wave0()) {
    if (fresh_cap) tx.fresh(fresh_cap, log, &K);
    else tx_begin(tx, t, log, js, &K);
    write_root_body(tx, s_node);
  }
  if (cb.slot) claim_begin_body(tx, cb, sc);
  if (is_wave0()) tx_end(tx, t, js);
}

// A claim group's begin on its own (the commit / eval / open entry points,
// where the caller decides between a drawn and a given point after the
// commitment): the point, written to the group's slots, and its eq tables.
__global__ void __launch_bounds__(SERIAL_THREADS) claim_begin_kernel(JP<DevTx> t_, JP<uint32_t> log_, JP<JobScalars> js_,
                                                                     ClaimBegin cb, GivenPoint gp, uint32_t given) {
  DevTx* t = t_;
  uint32_t* log = log_;
  JobScalars* js = js_;
  __shared__ ClaimScratch sc;
  WaveTx tx;
  if (is_wave0()) tx_begin(tx, t, log, js);
  claim_begin_body(tx, cb, sc, given ? &gp : nullptr);
  if (is_wave0()) tx_end(tx, t, js);
}

// Continue a caller's transcript: the challenger state `init` (sponge state,
// input and output buffers) with an empty proof log of init.log_cap words.
__global__ void tx_import_kernel(JP<DevTx> t, JP<JobScalars> js, DevTx init) {
  if (threadIdx.x == 0) {
    *t.get() = init;
    js->pending_words = 0;
  }
}

// Reduce the evaluation partials of a group of `count` claims drawn at one
// point and write their values in order (prover.rs:64-90,276; whir.rs:230).
// fsv > 0: claim c's partials hold 2^fsv compressed values per block, at
// partials[(c * n_part + p) * 2^fsv + i] (SvoPoint::eval, svo.rs:77-91);
// fsv == 0: one value per block (eval_ext). Then begin the next group when
// `next.slot` is set.
struct ClaimEndScratch {
  Ext sh[SERIAL_THREADS];
  Ext eq[1 << MAX_F];
};

// The values of the `count` claims of a group into s_eval[] (every thread
// returns after the last barrier); slots get their compressed values and eval.
// n_part is a power of two.
__device__ void claim_end_values(ClaimSlot* slot, uint32_t count, const Ext* partials, uint32_t n_part, uint32_t fsv,
                                 Ext* s_eval, ClaimEndScratch& cs) {
  Ext* sh = cs.sh;
  Ext* s_eq = cs.eq;
  const uint32_t W = 1u << fsv;
  const uint32_t tid = threadIdx.x;
  // eq(svo)[i] for i < W, svo = point[0..fsv) (the group's shared point)
  if (tid < W) {
    Ext e = eone();
    for (uint32_t j = 0; j < fsv; ++j) {
      const Ext z = slot[0].point[j];
      e = emul(e, ((tid >> j) & 1) ? z : esub(eone(), z));
    }
    s_eq[tid] = e;
  }
  // A pass covers cpp claims with S = 256 / cpp threads each (S, W powers of
  // two, W <= S). Thread (claim c, sub) sums elements sub, sub + S, ... of
  // claim c's n_part * W partials (all in column sub mod W): consecutive
  // threads read consecutive 16 B, eight loads in flight. The columns are then
  // summed by a butterfly over the claim's lanes of each wavefront (span
  // lanes) and across its wpc wavefronts through LDS.
  const uint32_t per_claim = n_part * W;
  const uint32_t cpp = per_claim >= SERIAL_THREADS ? 1u : SERIAL_THREADS / per_claim;
  const uint32_t S = SERIAL_THREADS / cpp;
  const uint32_t span = S < 64 ? S : 64, wpc = S / span;
  for (uint32_t base = 0; base < count; base += cpp) {
    const uint32_t c = base + tid / S, sub = tid % S;
    Ext acc = ezero();
    if (c < count) {
      const Ext* pc = partials + (uint64_t)c * per_claim;
      uint32_t e = sub;
      for (; e + 7 * S < per_claim; e += 8 * S) {
        Ext v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = ld_ext(pc + e + u * S);
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = eadd(acc, v[u]);
      }
      for (; e < per_claim; e += S) acc = eadd(acc, ld_ext(pc + e));
    }
    for (uint32_t m = W; m < span; m <<= 1) {
#pragma unroll
      for (int q = 0; q < 4; ++q) acc.c[q] = add(acc.c[q], (uint32_t)__shfl_xor((int)acc.c[q], (int)m));
    }
    if (sub % span < W) sh[(tid / S) * wpc * W + (sub / span) * W + sub % W] = acc;
    __syncthreads();
    // thread ci * W + i: column i of claim base + ci; eval by a butterfly over
    // the claim's W consecutive lanes
    const uint32_t ci = tid / W, i = tid % W;
    const bool mine = ci < cpp && base + ci < count;
    Ext prod = ezero();
    if (mine) {
      Ext comp = sh[ci * wpc * W + i];
      for (uint32_t j = 1; j < wpc; ++j) comp = eadd(comp, sh[ci * wpc * W + j * W + i]);
      slot[base + ci].compressed[i] = comp;
      prod = emul(comp, s_eq[i]);
    }
    for (uint32_t m = 1; m < W; m <<= 1) {
#pragma unroll
      for (int q = 0; q < 4; ++q) prod.c[q] = add(prod.c[q], (uint32_t)__shfl_xor((int)prod.c[q], (int)m));
    }
    if (mine && i == 0) {
      s_eval[base + ci] = prod;
      slot[base + ci].eval = prod;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(SERIAL_THREADS) claim_end_kernel(JP<DevTx> t_, JP<uint32_t> log_, JP<JobScalars> js_,
                                                                   JP<ClaimSlot> slot_, uint32_t count,
                                                                   JP<const Ext> partials_, uint32_t n_part, uint32_t fsv,
                                                                   ClaimBegin next) {
  DevTx* t = t_;
  uint32_t* log = log_;
  JobScalars* js = js_;
  ClaimSlot* slot = slot_;
  const Ext* partials = partials_;
  __shared__ ClaimEndScratch cs;
  __shared__ Ext s_eval[MAX_CLAIMS];
  __shared__ ClaimScratch sc;
  WaveTx tx;
  if (is_wave0()) tx_begin(tx, t, log, js);  // the sponge's loads overlap the reduction
  claim_end_values(slot, count, partials, n_part, fsv, s_eval, cs);
  if (is_wave0())
    for (uint32_t c = 0; c < count; ++c) tx.write_e(s_eval[c]);
  if (next.slot) claim_begin_body(tx, next, sc);
  if (is_wave0()) tx_end(tx, t, js);
}

// ProverLayout::new_prover (prover.rs:279-353) with Svo::calculate_accumulators
// (svo.rs:224-278) and the expression scales (prover.rs:184-208):
//   draw alpha; f SVO rounds (write c0, c2; draw r); eq(rs) into js->eq_rs;
//   claim c's scale alpha^rank * eq(svo_c, rs) into claim_scale[c] (the
//   expression build applies it to the group's eq1 table).
// The accumulators are linear in the claims' partial evaluations, so a group's
// claims (one point) are combined with their alpha powers first and the
// accumulators are taken once per group, as the reference does per point
// (prover.rs:300-313). Claim c has alpha power alpha^desc[c].rank (user claims
// in layout order, then the virtual claims).
// Scratch: pe_buf (n_claims x PE_PER), gpe_buf (n_groups x PE_PER); eq1_pref:
// prefix sums of the claims' eq1 sizes (n_claims + 1).
__global__ void __launch_bounds__(SERIAL_THREADS) svo_kernel(JP<DevTx> t_, JP<uint32_t> log_, JP<JobScalars> js_,
                                                             JP<const ClaimSlot> slots_, JP<const ClaimDesc> desc_,
                                                             uint32_t n_claims, JP<const uint32_t> group_start_,
                                                             uint32_t n_groups, uint32_t f, JP<Ext> pe_buf_,
                                                             JP<Ext> gpe_buf_, JP<Ext> claim_scale_) {
  DevTx* t = t_;
  uint32_t* log = log_;
  JobScalars* js = js_;
  const ClaimSlot* slots = slots_;
  const ClaimDesc* desc = desc_;
  const uint32_t* group_start = group_start_;
  Ext* pe_buf = pe_buf_;
  Ext* gpe_buf = gpe_buf_;
  Ext* claim_scale = claim_scale_;
  // partial evals pe[c][2^l - 2 + a] (l = 1..f), accumulators tot[ri][side][u]
  constexpr int PE_PER = (2 << MAX_F) - 2;
  constexpr int ACC_PER = 243 - 1;  // sum_{ri<5} 2 * 3^ri
  __shared__ Ext s_tot[ACC_PER];
  __shared__ Ext s_part[SERIAL_THREADS];
  __shared__ Ext s_apow[MAX_CLAIMS];  // alpha^rank, then the expression scales
  __shared__ Ext s_rs[MAX_F];
  __shared__ Ext s_lag[MAX_F][3];
  __shared__ Ext s_alpha;
  const uint32_t tid = threadIdx.x;
  const uint32_t W = 1u << f;
  uint32_t acc_per = 0, pow3 = 1;
  for (uint32_t ri = 0; ri < f; ++ri) {
    acc_per += 2 * pow3;
    pow3 *= 3;
  }
  const uint32_t pe_per = 2 * W - 2;
  WaveTx tx;
  if (is_wave0()) tx_begin(tx, t, log, js);

  // (1) partial evals: pe[l][a] = sum_b comp[b*2^l + a] * eq(svo[l..f))[b].
  // One product per thread, item (claim c, level l, a, b) at lane offset
  // (l - 1) W + a + b 2^l of claim c's f W items; the sum over b is a butterfly
  // over the lanes 2^l, ..., W/2 apart inside the W-aligned group of (c, l).
  const uint32_t n_items = n_claims * f * W;
  for (uint32_t base = 0; base < n_items; base += SERIAL_THREADS) {
    const uint32_t it = base + tid;
    const bool valid = it < n_items;
    const uint32_t c = valid ? it / (f * W) : 0u, rem = it % (f * W);
    const uint32_t l = rem / W + 1, inner = rem % W;
    const uint32_t a = inner & ((1u << l) - 1), b = inner >> l;
    Ext v = ezero();
    if (valid) {
      const ClaimSlot& sl = slots[c];
      Ext e = eone();
      for (uint32_t j = 0; j + l < f; ++j) {
        const Ext z = sl.point[l + j];
        e = emul(e, ((b >> j) & 1) ? z : esub(eone(), z));
      }
      v = emul(sl.compressed[(b << l) + a], e);
    }
    for (uint32_t m = 1; m < W; m <<= 1) {
      Ext o;
#pragma unroll
      for (int q = 0; q < 4; ++q) o.c[q] = (uint32_t)__shfl_xor((int)v.c[q], (int)m);
      if (m >= (1u << l)) v = eadd(v, o);
    }
    if (valid && b == 0) pe_buf[(uint64_t)c * PE_PER + (1u << l) - 2 + a] = v;
  }

  // (2) draw alpha; alpha^rank per claim
  if (is_wave0()) {
    const Ext alpha = tx.sample_ext();
    if (tid == 0) {
      s_alpha = alpha;
      js->alpha = alpha;
    }
  }
  __syncthreads();
  for (uint32_t c = tid; c < n_claims; c += SERIAL_THREADS) s_apow[c] = epow(s_alpha, desc[c].rank);
  __syncthreads();

  // (3) per group: gpe[g][o] = sum_{c in g} alpha^rank_c * pe[c][o]
  for (uint32_t it = tid; it < n_groups * pe_per; it += SERIAL_THREADS) {
    const uint32_t g = it / pe_per, o = it % pe_per;
    Ext s = ezero();
    for (uint32_t c = group_start[g]; c < group_start[g + 1]; ++c)
      s = eadd(s, emul(s_apow[c], pe_buf[(uint64_t)c * PE_PER + o]));
    gpe_buf[(uint64_t)g * PE_PER + o] = s;
  }
  __syncthreads();

  // (4) accumulators, summed over groups: `chunks` threads per entry split the groups
  const uint32_t chunks = acc_per < SERIAL_THREADS ? SERIAL_THREADS / acc_per : 1;
  for (uint32_t o0 = 0; o0 < acc_per; o0 += SERIAL_THREADS / chunks) {
    const uint32_t oi = o0 + tid % (SERIAL_THREADS / chunks), ch = tid / (SERIAL_THREADS / chunks);
    Ext total = ezero();
    if (oi < acc_per && ch < chunks) {
      uint32_t o = oi, ri = 0, p3 = 1;
      while (o >= 2 * p3) {
        o -= 2 * p3;
        p3 *= 3;
        ++ri;
      }
      const uint32_t side = o / p3, u = o % p3;
      const uint32_t l = ri + 1;
      uint32_t pt[MAX_F];
      uint32_t uu = u;
      for (uint32_t j = 0; j < ri; ++j) {
        pt[j] = uu % 3;
        uu /= 3;
      }
      pt[ri] = side ? 2u : 0u;
      for (uint32_t g = ch; g < n_groups; g += chunks) {
        const ClaimSlot& sl = slots[group_start[g]];
        // eval_eq_xy(pt, svo[0..l)) with pt_j in {0,1,2}: 1-y, y, 3y-1
        Ext exy = eone();
        for (uint32_t j = 0; j < l; ++j) {
          const Ext y = sl.point[j];
          Ext term;
          if (pt[j] == 0) term = esub(eone(), y);
          else if (pt[j] == 1) term = y;
          else term = esub(eadd(edbl(y), y), eone());
          exy = emul(exy, term);
        }
        // <gpe[l], eq(pt)>; eq(pt)[a] = prod_j (a_j ? pt_j : 1 - pt_j) is 0 unless
        // a_j = pt_j wherever pt_j < 2; a position with pt_j = 2 gives 2 (a_j = 1)
        // or -1 (a_j = 0), so the terms are +-2^e gp[a] over subsets of those
        const Ext* gp = gpe_buf + (uint64_t)g * PE_PER + (1u << l) - 2;
        uint32_t fixed = 0, freem = 0;
        for (uint32_t j = 0; j < l; ++j) {
          if (pt[j] == 1) fixed |= 1u << j;
          else if (pt[j] == 2) freem |= 1u << j;
        }
        const uint32_t n_free = (uint32_t)__builtin_popcount(freem);
        Ext dot = ezero();
        uint32_t sub = 0;
        do {
          Ext term = gp[fixed | sub];
          const uint32_t ones = (uint32_t)__builtin_popcount(sub);
          for (uint32_t e = 0; e < ones; ++e) term = edbl(term);
          dot = ((n_free - ones) & 1) ? esub(dot, term) : eadd(dot, term);
          sub = (sub - freem) & freem;
        } while (sub != 0);
        total = eadd(total, emul(exy, dot));
      }
    }
    s_part[tid] = total;
    __syncthreads();
    if (tid < SERIAL_THREADS / chunks && o0 + tid < acc_per) {
      Ext s = ezero();
      for (uint32_t c = 0; c < chunks; ++c) s = eadd(s, s_part[c * (SERIAL_THREADS / chunks) + tid]);
      s_tot[o0 + tid] = s;
    }
    __syncthreads();
  }
  // (5) SVO rounds
  if (is_wave0()) {
    uint32_t off = 0, p3 = 1;
    for (uint32_t ri = 0; ri < f; ++ri) {
      Ext c0 = ezero(), c2 = ezero();
      for (uint32_t u = tid; u < p3; u += 64) {
        Ext w = eone();
        uint32_t uu = u;
        for (uint32_t j = 0; j < ri; ++j) {
          w = emul(w, s_lag[j][uu % 3]);
          uu /= 3;
        }
        c0 = eadd(c0, emul(s_tot[off + u], w));
        c2 = eadd(c2, emul(s_tot[off + p3 + u], w));
      }
      c0 = wave_sum_ext(c0);
      c2 = wave_sum_ext(c2);
      tx.write_e(c0);
      tx.write_e(c2);
      const Ext r = tx.sample_ext();
      if (tid == 0) {
        s_rs[ri] = r;
        js->rs[ri] = r;
      }
      if (tid < 3) {  // lagrange012(r), one weight per lane
        const Ext one = eone(), two = efrom(TWO);
        const Ext a = tid == 0 ? esub(r, one) : r;
        const Ext b = tid == 0 ? esub(r, two) : tid == 1 ? esub(two, r) : esub(r, one);
        const Ext m = emul(a, b);
        s_lag[ri][tid] = tid == 1 ? m : emul_base(m, HALF);
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      off += 2 * p3;
      p3 *= 3;
    }
    tx_end(tx, t, js);
  }
  __syncthreads();
  // (6) eq(rs) and claim scales alpha^rank * eq_xy(svo_c, rs)
  if (tid < W) {
    Ext e = eone();
    for (uint32_t j = 0; j < f; ++j) e = emul(e, ((tid >> j) & 1) ? s_rs[j] : esub(eone(), s_rs[j]));
    js->eq_rs[tid] = e;
  }
  for (uint32_t c = tid; c < n_claims; c += SERIAL_THREADS) {
    Ext e = s_apow[c];
    for (uint32_t j = 0; j < f; ++j) {
      const Ext x = slots[c].point[j], y = s_rs[j];
      // (y x) * 2 - x - y + 1
      e = emul(e, eadd(esub(esub(edbl(emul(y, x)), x), y), eone()));
    }
    claim_scale[c] = e;  // expr_build_kernel applies it to the group's eq1 table
  }
}

// Bitonic sort of s_v[0..np) (np a power of two, values at and past n read
// as 0xFFFFFFFF) with the values in registers: thread t holds elements
// t + k T (k < E, E T >= np). A compare-exchange partner x ^ stride is in the
// same thread when stride >= T (another register), in the same wavefront when
// stride < 64 (a lane shuffle) and in another wavefront otherwise (through
// s_v, two workgroup barriers). Writes the sorted values back to s_v.
template <int E>
__device__ void bitonic_regs(uint32_t* s_v, uint32_t n, uint32_t np) {
  const uint32_t tid = threadIdx.x, T = blockDim.x;
  uint32_t v[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const uint32_t x = tid + (uint32_t)k * T;
    v[k] = x < n ? s_v[x] : 0xFFFFFFFFu;  // > any index (bits <= 31)
  }
  for (uint32_t size = 2; size <= np; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= T) {
        const uint32_t kx = stride / T;
#pragma unroll
        for (int k = 0; k < E; ++k) {
          const uint32_t kp = (uint32_t)k ^ kx;
          if ((uint32_t)k < kp && kp < (uint32_t)E) {
            const uint32_t x = tid + (uint32_t)k * T;
            const bool up = (x & size) == 0;
            const uint32_t a = v[k], b = v[kp];
            if ((a > b) == up) {
              v[k] = b;
              v[kp] = a;
            }
          }
        }
        continue;
      }
      uint32_t o[E];
      if (stride < 64) {
#pragma unroll
        for (int k = 0; k < E; ++k) o[k] = (uint32_t)__shfl_xor((int)v[k], (int)stride, 64);
      } else {
#pragma unroll
        for (int k = 0; k < E; ++k) s_v[tid + (uint32_t)k * T] = v[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < E; ++k) o[k] = s_v[(tid + (uint32_t)k * T) ^ stride];
        __syncthreads();
      }
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const uint32_t x = tid + (uint32_t)k * T;
        const bool keep_min = ((x & stride) == 0) == ((x & size) == 0);
        v[k] = keep_min ? (v[k] < o[k] ? v[k] : o[k]) : (v[k] > o[k] ? v[k] : o[k]);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const uint32_t x = tid + (uint32_t)k * T;
    if (x < np) s_v[x] = v[k];
  }
}

// Sort s_v[0..n) and write its distinct values in ascending order to out[],
// with the whole workgroup (s_scan: blockDim.x words; blockDim.x a multiple
// of 64 with 4 blockDim.x >= n). A bitonic network over the next power of two
// (bitonic_regs; padding sorts last), then a workgroup prefix count over
// "first of its value" flags gives each distinct value its rank. Returns the
// number of distinct values (in every thread). n <= MAX_STIR. Measured on
// MI355X (stir_begin at k = 24): n = 287, 10.9 us with a round trip through
// LDS and a workgroup barrier per stage, 8.8 us as here, 12.8 us on one
// wavefront with 8 values per lane; n = 133, 8.4 / 5.3 / 7.4 us.
__device__ uint32_t sort_dedup_block(uint32_t* s_v, uint32_t n, uint32_t* s_scan, uint32_t* out) {
  const uint32_t tid = threadIdx.x, T = blockDim.x;
  uint32_t np = 1;
  while (np < n) np <<= 1;
  __syncthreads();  // s_v complete
  if (np <= T) bitonic_regs<1>(s_v, n, np);
  else if (np <= 2 * T) bitonic_regs<2>(s_v, n, np);
  else bitonic_regs<4>(s_v, n, np);
  __syncthreads();
  // per thread: a contiguous chunk of the sorted values; count the firsts
  const uint32_t chunk = (n + T - 1) / T;
  const uint32_t b0 = tid * chunk, b1 = b0 + chunk < n ? b0 + chunk : n;
  uint32_t cnt = 0;
  for (uint32_t i = b0; i < b1; ++i) cnt += (i == 0 || s_v[i] != s_v[i - 1]) ? 1u : 0u;
  // exclusive prefix: an inclusive scan in each wavefront, then the
  // wavefronts' totals through LDS
  const uint32_t lane = tid & 63, wave = tid >> 6, nw = T >> 6;
  uint32_t inc = cnt;
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t o = (uint32_t)__shfl_up((int)inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) s_scan[wave] = inc;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (uint32_t w = 0; w < nw; ++w) {
    const uint32_t c = s_scan[w];
    total += c;
    if (w < wave) before += c;
  }
  uint32_t rank = before + inc - cnt;
  for (uint32_t i = b0; i < b1; ++i)
    if (i == 0 || s_v[i] != s_v[i - 1]) out[rank++] = s_v[i];
  __syncthreads();
  return total;
}

// The STIR step of a round up to the fold's alpha, in one kernel:
//   [the last OOD claim's value: write it (whir.rs:230)]
//   n draws of `bits` bits -> sorted distinct indexes (whir.rs:63-77, 238);
//   draw alpha (prover.rs:377-384; the openings written afterwards are hints,
//   which the sponge does not absorb, so alpha can be drawn before them);
//   with query_scales, the query scales alpha^(n_eq + q) (combine.rs:37-64,
//   prover.rs:419-425), which the same round's openings launch reads; the
//   rest of alpha's scaling is stir_scales_block's.
struct StirBegin {
  JP<ClaimSlot> end_slot;  // OOD claim whose value this kernel writes first (nullptr: none)
  JP<const Ext> partials;  // its eval_ext partials
  uint32_t n_part;
  uint32_t n, bits;
  JP<uint32_t> idx_out;
  JP<uint32_t> count_out;
  JP<Ext> eq_pool;         // OOD claims' split eq tables (eq0: 2^m0 entries, then eq1)
  uint64_t eq_stride;
  uint32_t n_eq, m0, m1;
  QueryArrays qa;
  uint32_t query_scales;  // 0: stir_query_kernel's extra workgroups write them
};

__global__ void __launch_bounds__(SERIAL_THREADS) stir_begin_kernel(JP<DevTx> t_, JP<uint32_t> log_, JP<JobScalars> js_,
                                                                    StirBegin a) {
  DevTx* t = t_;
  uint32_t* log = log_;
  JobScalars* js = js_;
  __shared__ ClaimEndScratch cs;
  __shared__ Ext s_eval[1];
  __shared__ uint32_t s_v[MAX_STIR];
  __shared__ uint32_t s_scan[SERIAL_THREADS];
  __shared__ Ext s_alpha;
  const uint32_t tid = threadIdx.x;
  WaveTx tx;
  if (is_wave0()) tx_begin(tx, t, log, js);
  if (a.end_slot) claim_end_values(a.end_slot, 1, a.partials, a.n_part, 0, s_eval, cs);
  if (is_wave0()) {
    if (a.end_slot) tx.write_e(s_eval[0]);
    tx.sample_bits_many(a.n, a.bits, s_v);
    const Ext al = tx.sample_ext();
    tx_end(tx, t, js);
    if (tid == 0) {
      s_alpha = al;
      js->alpha = al;
    }
  }
  const uint32_t count = sort_dedup_block(s_v, a.n, s_scan, a.idx_out);  // barriers make s_alpha visible
  if (tid == 0) {
    *a.count_out = count;
    js->stir_count = count;
  }
  if (!a.query_scales) return;
  // query scales: thread t takes q = t + 256 i as alpha^(n_eq + t) alpha^(256 i)
  // (alpha^256 only when there are more than 256 queries: its eight squarings
  // would otherwise follow the power on every thread's chain)
  Ext sq = epow(s_alpha, (uint64_t)a.n_eq + tid);
  if (tid < count) a.qa.scale[tid] = sq;
  if (count > SERIAL_THREADS) {
    Ext step = s_alpha;
    for (int i = 0; i < 8; ++i) step = esqr(step);  // alpha^256 (SERIAL_THREADS)
    for (uint32_t q = tid + SERIAL_THREADS; q < count; q += SERIAL_THREADS) {
      sq = emul(sq, step);
      a.qa.scale[q] = sq;
    }
  }
}

// What the STIR step's alpha scales and no transcript kernel reads before the
// next claim combination: eq1 of OOD claim c >= 1 times alpha^c (combine.rs:
// 37-64), and in round 0 the query scales alpha^(n_eq + q) (read first by
// round_finalize_kernel). STIR_EXTRA_BLOCKS workgroups appended to the
// openings launch do it, off the transcript's chain.
constexpr uint32_t STIR_EXTRA_BLOCKS = 4;
struct StirScales {
  JP<Ext> eq_pool;  // OOD claims' split eq tables (eq0: 2^m0 entries, then eq1)
  uint64_t eq_stride;
  uint32_t n_eq, m0, m1;
  JP<Ext> scale;    // query scales to write; nullptr: stir_begin_kernel wrote them
};
__device__ void stir_scales_block(const StirScales& ss, const Ext alpha, uint32_t count, uint32_t eb) {
  constexpr uint32_t S = STIR_EXTRA_BLOCKS * SERIAL_THREADS;
  const uint32_t tid = eb * SERIAL_THREADS + threadIdx.x;
  if (ss.scale)
    for (uint32_t q = tid; q < count; q += S) ss.scale[q] = epow(alpha, (uint64_t)ss.n_eq + q);
  // four loads in flight
  const uint64_t n0 = 1ull << ss.m0, n1 = 1ull << ss.m1;
  const uint64_t n_sc = ss.n_eq > 1 ? (uint64_t)(ss.n_eq - 1) * n1 : 0;
  for (uint64_t it0 = tid; it0 < n_sc; it0 += 4 * S) {
    Ext v[4];
    Ext* p[4];
    uint32_t c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint64_t it = it0 + (uint64_t)u * S;
      c[u] = (uint32_t)(1 + it / n1);
      p[u] = it < n_sc ? ss.eq_pool + c[u] * ss.eq_stride + n0 + it % n1 : nullptr;
      if (p[u]) v[u] = ld_ext(p[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (p[u]) st_ext(p[u], emul(v[u], epow(alpha, c[u])));
  }
}

// Word w of query `index`'s opening: the row (row_words words), then the
// `depth` sibling digests bottom-up (sibling at level i of the node on the
// path = (index mod 2^(depth-i)) ^ 2^(depth-i-1); layers stored leaves first).
__device__ __forceinline__ uint32_t opening_word(const uint32_t* cw, uint32_t row_words, const uint32_t* tree,
                                                 uint32_t depth, uint32_t index, uint32_t w) {
  if (w < row_words) return cw[(uint64_t)index * row_words + w];
  const uint32_t e = w - row_words, i = e >> 3;
  const uint32_t mid = 1u << (depth - i - 1);
  const uint32_t sib = (index & ((mid << 1) - 1)) ^ mid;
  const uint64_t layer_off = (2ull << depth) - (2ull << (depth - i));  // sum_{j<i} 2^(depth-j)
  return tree[(layer_off + sib) * 8 + (e & 7)];
}

// Merkle openings of the sorted queries (MerkleTree::query, poseidon_packed.rs:
// 145-233): per query the flattened row, then `depth` sibling digests bottom-up,
// as transcript hints. One block per query; offsets are fixed per query, so
// blocks write in parallel and the next transcript kernel advances the log.
// mode 1 (round 0, fold_with_code): also G[t] = <eq(rs), row(index mod M + t M)>
// for t < 2^d (prover.rs:483-506) and ws/ts (prover.rs:463-481).
// mode 2 (later rounds): pow-claim variable var = omega^index (whir.rs:242-246)
// and the query's pow tables for combine_kernel: powL[q][e] = var^e (e < 2^m0),
// powR[q][e] = scale_q var^(e 2^m0) (e < 2^m1).
__global__ void __launch_bounds__(SERIAL_THREADS) stir_query_kernel(JP<const DevTx> t_, JP<JobScalars> js_, JP<uint32_t> log_,
                                                                    JP<const uint32_t> idx_, JP<const uint32_t> count_ptr_,
                                                                    JP<const uint32_t> cw_, uint32_t row_words,
                                                                    JP<const uint32_t> tree_, uint32_t depth, int mode,
                                                                    uint32_t d, QueryArrays qa, uint32_t m0,
                                                                    uint32_t m1, JP<uint32_t> powL_, JP<Ext> powR_,
                                                                    StirScales ss) {
  const DevTx* t = t_;
  JobScalars* js = js_;
  uint32_t* log = log_;
  const uint32_t* idx = idx_;
  const uint32_t* count_ptr = count_ptr_;
  const uint32_t* cw = cw_;
  const uint32_t* tree = tree_;
  uint32_t* powL = powL_;
  Ext* powR = powR_;
  const uint32_t count = *count_ptr;
  const uint32_t n_open = gridDim.x - STIR_EXTRA_BLOCKS;  // workgroups for openings
  if (blockIdx.x >= n_open) {
    stir_scales_block(ss, js->alpha, count, blockIdx.x - n_open);
    return;
  }
  const uint32_t q = blockIdx.x;
  const uint32_t per = row_words + 8 * depth;
  if (blockIdx.x == 0 && threadIdx.x == 0) js->pending_words = count * per;
  if (q >= count) return;
  const uint32_t index = idx[q];
  const uint64_t base = (uint64_t)t->log_words + (uint64_t)q * per;
  const bool fits = base + per <= t->log_cap;
  if (fits)
    for (uint32_t w = threadIdx.x; w < per; w += SERIAL_THREADS)
      log[base + w] = opening_word(cw, row_words, tree, depth, index, w);
  if (mode == 1) {
    // rows index mod M + t*M, M = 2^(depth - d); each row has 2^f = row_words base words
    const uint32_t M = 1u << (depth - d);
    const uint32_t b = index & (M - 1);
    for (uint32_t tt = threadIdx.x; tt < (1u << d); tt += SERIAL_THREADS) {
      const uint32_t* row = cw + (uint64_t)(b + tt * M) * row_words;
      Ext acc = ezero();
      for (uint32_t i = 0; i < row_words; ++i) acc = eadd(acc, emul_base(js->eq_rs[i], row[i]));
      qa.G[q * (1u << MAX_F) + tt] = acc;
    }
    if (threadIdx.x == 0) {
      const uint32_t om = two_adic_generator(depth);
      qa.index[q] = index;
      qa.ws[q] = pow(om, index);
      // omega^-(index mod mid) = omega^(2^depth - index mod mid): om has order 2^depth
      qa.ts[q] = pow(om, (1ull << depth) - (index & ((1u << (depth - 1)) - 1)));
    }
  } else if (mode == 2) {
    const uint32_t var = pow(two_adic_generator(depth), index);
    if (threadIdx.x == 0) {
      qa.index[q] = index;
      qa.var[q] = var;
    }
    const uint32_t n0 = 1u << m0, n1 = 1u << m1;
    const Ext scale = qa.scale[q];
    for (uint32_t e = threadIdx.x; e < n0 + n1; e += SERIAL_THREADS) {
      if (e < n0) powL[(uint64_t)q * n0 + e] = pow(var, e);
      else st_ext(powR + (uint64_t)q * n1 + (e - n0), emul_base(scale, pow(var, (uint64_t)(e - n0) << m0)));
    }
  }
}

// Reduce the (c0, c2) partials, write c0, c2 and draw r (expr.rs:52-70). In
// round 0 the fold_with_code query terms (prover.rs:509-534) are among the
// partials (query_terms_partial, by the query workgroups of the launch before)
// and the per-query state advances with r (prover.rs:542-553) in the next
// fold launch (query_fold_block): both off the transcript's chain.
__global__ void __launch_bounds__(SERIAL_THREADS) round_finalize_kernel(JP<DevTx> t_, JP<uint32_t> log_, JP<JobScalars> js_,
                                                                        JP<const Ext> partials_, uint32_t n_part,
                                                                        uint32_t rs_index) {
  DevTx* t = t_;
  uint32_t* log = log_;
  JobScalars* js = js_;
  const Ext* partials = partials_;
  __shared__ Ext sh[SERIAL_THREADS];
  const uint32_t tid = threadIdx.x;
  WaveTx tx;
  if (is_wave0()) tx_begin(tx, t, log, js);
  Ext c0 = ezero(), c2 = ezero();
  for (uint32_t p0 = tid; p0 < n_part; p0 += 4 * SERIAL_THREADS) {  // four pairs in flight
    Ext a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t p = p0 + (uint32_t)u * SERIAL_THREADS;
      a[u] = p < n_part ? ld_ext(partials + 2 * p) : ezero();
      b[u] = p < n_part ? ld_ext(partials + 2 * p + 1) : ezero();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      c0 = eadd(c0, a[u]);
      c2 = eadd(c2, b[u]);
    }
  }
  const Ext C0 = block_reduce_ext<SERIAL_THREADS>(c0, sh);
  __syncthreads();
  const Ext C2 = block_reduce_ext<SERIAL_THREADS>(c2, sh);
  if (is_wave0()) {
    tx.write_e(C0);
    tx.write_e(C2);
    const Ext r = tx.sample_ext();
    tx_end(tx, t, js);
    if (tid == 0) js->rs[rs_index] = r;
  }
}

// The per-query state of fold_with_code after a draw r at fold step rho
// (prover.rs:542-553): scale *= 1 + r (w - 1), w and t squared (t negated by
// the index bit), and G_{rho+1}[tt] = fold(G[tt], G[tt + h], omega^-i, r) for
// every (query, tt), four items' loads in flight. Then, with `partials`, the
// query terms of step rho + 1 (query_terms_partial). QUERY_BLOCKS workgroups
// of SERIAL_THREADS threads, workgroup qb taking the queries q = qb mod
// QUERY_BLOCKS (so that a query's terms read only its own workgroup's
// writes), run as the last workgroups of the next fold launch: the next
// round_finalize_kernel reads their partials after that launch ends.
struct QueryFold {
  QueryArrays qa;
  JP<const uint32_t> count_ptr;
  uint32_t d, k_data, rho;
  JP<const uint32_t> tw_lo;
  JP<const uint32_t> tw_hi;
};

constexpr uint32_t QUERY_BLOCKS = 4;
static_assert(SC_THREADS == SERIAL_THREADS, "the query workgroups run in a fold launch");

// The fold_with_code terms of step rho (prover.rs:509-534) summed over the
// queries q = qb mod QUERY_BLOCKS: c0 += s even, c2 += s (2 odd - even)(2 w - 1),
// written as partial pair pidx. sh: SERIAL_THREADS / 64 entries.
__device__ void query_terms_partial(const QueryArrays& qa, uint32_t count, uint32_t qb, uint32_t d, uint32_t k_data,
                                    uint32_t rho, Ext* partials, uint32_t pidx, Ext* sh) {
  const uint32_t GS = 1u << MAX_F;
  const uint32_t half = 1u << (d - rho - 1);  // t_hi = t_lo + half
  Ext c0 = ezero(), c2 = ezero();
  for (uint32_t q = qb + QUERY_BLOCKS * threadIdx.x; q < count; q += QUERY_BLOCKS * SERIAL_THREADS) {
    const uint32_t index = qa.index[q];
    const uint32_t t_lo = (index >> (k_data - d)) & (half - 1);
    const Ext g_lo = qa.G[q * GS + t_lo], g_hi = qa.G[q * GS + t_lo + half];
    const Ext even = emul_base(eadd(g_lo, g_hi), HALF);
    const Ext odd = emul_base(emul_base(esub(g_lo, g_hi), HALF), qa.ts[q]);
    const Ext sc = qa.scale[q];
    c0 = eadd(c0, emul(sc, even));
    const uint32_t wt = sub(dbl(qa.ws[q]), ONE);
    c2 = eadd(c2, emul_base(emul(sc, esub(edbl(odd), even)), wt));
  }
  const Ext s0 = block_reduce_ext<SERIAL_THREADS>(c0, sh);
  __syncthreads();
  const Ext s2 = block_reduce_ext<SERIAL_THREADS>(c2, sh);
  if (threadIdx.x == 0) {
    st_ext(partials + 2 * pidx, s0);
    st_ext(partials + 2 * pidx + 1, s2);
  }
}

__device__ void query_fold_block(const QueryFold& qf, const Ext r, uint32_t qb, Ext* partials, uint32_t pidx,
                                 Ext* sh) {
  const QueryArrays& qa = qf.qa;
  const uint32_t count = *qf.count_ptr;
  const uint32_t d = qf.d, k_data = qf.k_data, rho = qf.rho;
  const uint32_t GS = 1u << MAX_F;
  // this workgroup's queries: q = qb + QUERY_BLOCKS ql, ql < nq
  const uint32_t nq = count > qb ? (count - qb + QUERY_BLOCKS - 1) / QUERY_BLOCKS : 0u;
  for (uint32_t ql = threadIdx.x; ql < nq; ql += SERIAL_THREADS) {
    const uint32_t q = qb + QUERY_BLOCKS * ql;
    const uint32_t index = qa.index[q];
    const uint32_t ws = qa.ws[q];
    qa.scale[q] = emul(qa.scale[q], eadd(eone(), emul(r, esub(efrom(ws), eone()))));
    qa.ws[q] = sqr(ws);
    uint32_t ts = sqr(qa.ts[q]);
    if (rho + 2 <= k_data && ((index >> (k_data - rho - 2)) & 1)) ts = neg(ts);
    qa.ts[q] = ts;
  }
  if (rho + 1 < d) {
    // G_{rho+1}[tt] = fold(G[tt], G[tt + h], omega_{2^(k_data-rho)}^{-i}, r), i = b + tt*M
    const uint32_t h = 1u << (d - rho - 1);
    const uint32_t M = 1u << (k_data - d);
    const uint32_t lg = k_data - rho;
    const uint32_t lh = d - rho - 1;  // log2 h
    const uint32_t n_items = nq << lh;
    for (uint32_t it0 = threadIdx.x; it0 < n_items; it0 += 4 * SERIAL_THREADS) {
      Ext g_lo[4], g_hi[4];
      uint32_t tw[4], gi[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t it = it0 + (uint32_t)u * SERIAL_THREADS;
        gi[u] = 0xFFFFFFFFu;
        if (it < n_items) {
          const uint32_t q = qb + QUERY_BLOCKS * (it >> lh), tt = it & (h - 1);
          const uint32_t i = (qa.index[q] & (M - 1)) + tt * M;
          const uint32_t e = ((1u << lg) - i) & ((1u << lg) - 1);
          const uint32_t e24 = e << (24 - lg);
          gi[u] = q * GS + tt;
          tw[u] = mul(qf.tw_lo[e24 & 4095u], qf.tw_hi[e24 >> 12]);
          g_lo[u] = qa.G[gi[u]];
          g_hi[u] = qa.G[gi[u] + h];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (gi[u] == 0xFFFFFFFFu) continue;
        const Ext even = emul_base(eadd(g_lo[u], g_hi[u]), HALF);
        const Ext odd = emul_base(emul_base(esub(g_lo[u], g_hi[u]), HALF), tw[u]);
        qa.G[gi[u]] = eadd(even, emul(r, esub(odd, even)));
      }
    }
  }
  if (!partials) return;
  __syncthreads();  // this workgroup's queries are at step rho + 1
  query_terms_partial(qa, count, qb, d, k_data, rho + 1, partials, pidx, sh);
}

// coeffs_kernel / fold_coeffs_kernel / fold_kernel with QUERY_BLOCKS more
// workgroups, the last ones: the query terms of step 0 as partials g..g+3
// (coeffs_q_kernel), or query_fold_block with the same r (its terms as
// partials g..g+3 in fold_coeffs_q_kernel; none after the last fold).
__global__ void __launch_bounds__(SC_THREADS) coeffs_q_kernel(JP<const Ext> f_, JP<const Ext> w_,
                                                          uint64_t n, JP<Ext> partials_, QueryFold qf) {
  const Ext* __restrict__ f = f_;
  const Ext* __restrict__ w = w_;
  Ext* __restrict__ partials = partials_;
  __shared__ Ext sh[SC_THREADS];
  const uint32_t nf = gridDim.x - QUERY_BLOCKS;
  if (blockIdx.x >= nf)
    query_terms_partial(qf.qa, *qf.count_ptr, blockIdx.x - nf, qf.d, qf.k_data, 0, partials, blockIdx.x, sh);
  else coeffs_body(f, w, n, partials, blockIdx.x, nf, sh);
}
__global__ void __launch_bounds__(SC_THREADS) fold_coeffs_q_kernel(JP<const Ext> f_, JP<const Ext> w_,
                                                               uint64_t n, JP<const Ext> r_ptr_,
                                                               JP<Ext> fo_, JP<Ext> wo_,
                                                               JP<Ext> partials_, QueryFold qf) {
  const Ext* __restrict__ f = f_;
  const Ext* __restrict__ w = w_;
  const Ext* __restrict__ r_ptr = r_ptr_;
  Ext* __restrict__ fo = fo_;
  Ext* __restrict__ wo = wo_;
  Ext* __restrict__ partials = partials_;
  __shared__ Ext sh[SC_THREADS];
  const Ext r = ld_ext(r_ptr);
  const uint32_t nf = gridDim.x - QUERY_BLOCKS;
  if (blockIdx.x >= nf) query_fold_block(qf, r, blockIdx.x - nf, partials, blockIdx.x, sh);
  else fold_coeffs_body(f, w, n, r, fo, wo, partials, blockIdx.x, nf, sh);
}
__global__ void __launch_bounds__(SC_THREADS) fold_q_kernel(JP<const Ext> f_, JP<const Ext> w_,
                                                        uint64_t n, JP<const Ext> r_ptr_,
                                                        JP<Ext> fo_, JP<Ext> wo_, QueryFold qf) {
  const Ext* __restrict__ f = f_;
  const Ext* __restrict__ w = w_;
  const Ext* __restrict__ r_ptr = r_ptr_;
  Ext* __restrict__ fo = fo_;
  Ext* __restrict__ wo = wo_;
  __shared__ Ext sh[SC_THREADS];
  const Ext r = ld_ext(r_ptr);
  const uint32_t nf = gridDim.x - QUERY_BLOCKS;
  if (blockIdx.x >= nf) query_fold_block(qf, r, blockIdx.x - nf, nullptr, 0, sh);
  else fold_body(f, w, n, r, fo, wo, blockIdx.x);
}

}  // namespace kb
