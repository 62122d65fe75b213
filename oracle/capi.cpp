// ORACLE — test infrastructure only (see field.h). ctypes entry points used by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. All field
// values cross this boundary as Montgomery u32 words.
#include <cstring>
#include <omp.h>
#include "field.h"
#include "rng.h"
#include "poseidon2.h"
#include "transcript.h"
#include "whir.h"

using namespace oracle;

static Point to_point(const uint32_t* z, int k) {
  Point p(k);
  for (int i = 0; i < k; ++i)
    for (int c = 0; c < 4; ++c) p[i].c[c] = F{z[4 * i + c]};
  return p;
}
static void put_e(uint32_t* out, const E& e) {
  for (int c = 0; c < 4; ++c) out[c] = e.c[c].v;
}
static E get_e(const uint32_t* in) {
  E e;
  for (int c = 0; c < 4; ++c) e.c[c] = F{in[c]};
  return e;
}

extern "C" {

void oracle_set_threads(int n) { omp_set_num_threads(n); }
int oracle_get_threads() { return omp_get_max_threads(); }

uint32_t oracle_fmul(uint32_t a, uint32_t b) { return fmul(F{a}, F{b}).v; }
uint32_t oracle_fadd(uint32_t a, uint32_t b) { return fadd(F{a}, F{b}).v; }
uint32_t oracle_fsub(uint32_t a, uint32_t b) { return fsub(F{a}, F{b}).v; }
uint32_t oracle_finv(uint32_t a) { return finv(F{a}).v; }
uint32_t oracle_to_monty(uint32_t c) { return F::from_u32(c).v; }
uint32_t oracle_from_monty(uint32_t m) { return F{m}.canonical(); }
uint32_t oracle_two_adic_generator(unsigned bits) { return two_adic_generator(bits).v; }
void oracle_emul(const uint32_t* a, const uint32_t* b, uint32_t* out) { put_e(out, emul(get_e(a), get_e(b))); }

void oracle_chacha_block(const uint32_t* key, uint64_t counter, uint64_t stream, int rounds, uint32_t* out) {
  chacha_block(key, counter, stream, rounds, out);
}
void oracle_rng_u32(uint64_t seed, size_t n, uint32_t* out) {
  StdRng r = StdRng::seed_from_u64(seed);
  for (size_t i = 0; i < n; ++i) out[i] = r.next_u32();
}
void oracle_stdrng_key(uint64_t seed, uint32_t* out) {  // the ChaCha key seed_from_u64 derives
  StdRng r = StdRng::seed_from_u64(seed);
  memcpy(out, r.key, 32);
}
void oracle_rand_poly(uint64_t seed, int k, uint32_t* out) {
  std::vector<F> v = rand_poly(seed, k);
  memcpy(out, v.data(), v.size() * 4);
}
void oracle_rand_ext(uint64_t seed, size_t n, uint32_t* out) {
  StdRng r = StdRng::seed_from_u64(seed);
  for (size_t i = 0; i < n; ++i) put_e(out + 4 * i, r.next_e());
}

// Poseidon2 constants as used by the reference: width 16 or 24, seed 1000.
void oracle_poseidon2_constants(int width, uint64_t seed, uint32_t* ext_initial, uint32_t* ext_terminal,
                                uint32_t* internal, uint32_t* diag) {
  Poseidon2 p = Poseidon2::make(width, seed);
  for (int i = 0; i < 4 * width; ++i) {
    ext_initial[i] = p.ext_initial[i].v;
    ext_terminal[i] = p.ext_terminal[i].v;
  }
  for (int i = 0; i < p.rounds_p; ++i) internal[i] = p.internal[i].v;
  for (int i = 0; i < width; ++i) diag[i] = p.diag[i].v;
}
void oracle_poseidon2_permute(int width, uint32_t* state) {
  const Hashers& H = Hashers::get();
  const Poseidon2& p = width == 16 ? H.p16 : H.p24;
  p.permute(reinterpret_cast<F*>(state));
}
void oracle_mds_light(int width, uint32_t* state) {
  const Hashers& H = Hashers::get();
  (width == 16 ? H.p16 : H.p24).mds_light(reinterpret_cast<F*>(state));
}
void oracle_internal_layer(int width, uint32_t* state) {
  const Hashers& H = Hashers::get();
  (width == 16 ? H.p16 : H.p24).internal_layer(reinterpret_cast<F*>(state));
}
void oracle_sponge24(const uint32_t* in, size_t n, uint32_t* out8) {
  sponge24(Hashers::get().p24, reinterpret_cast<const F*>(in), n, reinterpret_cast<F*>(out8));
}
void oracle_compress16(const uint32_t* l, const uint32_t* r, uint32_t* out8) {
  compress16(Hashers::get().p16, reinterpret_cast<const F*>(l), reinterpret_cast<const F*>(r),
             reinterpret_cast<F*>(out8));
}

void oracle_dft_batch(uint32_t* mat, size_t h, size_t w) { dft_batch(reinterpret_cast<F*>(mat), h, w); }
void oracle_dft_batch_naive(uint32_t* mat, size_t h, size_t w) { dft_batch_naive(reinterpret_cast<F*>(mat), h, w); }

// Merkle: writes root (8 words) and, if `layers_out` is non-null, all digests
// layer by layer (leaves first): 2*height-1 digests of 8 words.
void oracle_merkle_commit(const uint32_t* data, size_t height, size_t row_base, uint32_t* root,
                          uint32_t* layers_out) {
  std::vector<F> d(height * row_base);
  memcpy(d.data(), data, d.size() * 4);
  Tree t = merkle_commit(std::move(d), height, row_base);
  memcpy(root, t.root().d, 32);
  if (layers_out) {
    size_t off = 0;
    for (auto& L : t.layers)
      for (auto& dg : L) {
        memcpy(layers_out + off, dg.d, 32);
        off += 8;
      }
  }
}

// Challenger probe: runs `n_obs` observes of `obs` then `n_samples` samples.
void oracle_challenger_probe(const uint32_t* obs, size_t n_obs, size_t n_samples, uint32_t* out) {
  Challenger ch(&Hashers::get().p16);
  for (size_t i = 0; i < n_obs; ++i) ch.observe(F{obs[i]});
  for (size_t i = 0; i < n_samples; ++i) out[i] = ch.sample().v;
}

// ---- poly ----
void oracle_eq_table(const uint32_t* z, int k, const uint32_t* scale, uint32_t* out) {
  std::vector<E> t = eq_table(to_point(z, k), get_e(scale));
  memcpy(out, t.data(), t.size() * 16);
}
void oracle_eval_base(const uint32_t* poly, int k, const uint32_t* z, uint32_t* out) {
  std::vector<F> p((size_t)1 << k);
  memcpy(p.data(), poly, p.size() * 4);
  put_e(out, eval_base(p, to_point(z, k)));
}
void oracle_eval_ext(const uint32_t* poly, int k, const uint32_t* z, uint32_t* out) {
  std::vector<E> p((size_t)1 << k);
  memcpy(p.data(), poly, p.size() * 16);
  put_e(out, eval_ext(p, to_point(z, k)));
}
void oracle_compress_hi(const uint32_t* poly, int k, const uint32_t* z, int m, uint32_t* out) {
  std::vector<F> p((size_t)1 << k);
  memcpy(p.data(), poly, p.size() * 4);
  std::vector<E> o = compress_hi(p, to_point(z, m));
  memcpy(out, o.data(), o.size() * 16);
}
void oracle_compress_lo(const uint32_t* poly, int k, const uint32_t* z, int m, uint32_t* out) {
  std::vector<F> p((size_t)1 << k);
  memcpy(p.data(), poly, p.size() * 4);
  std::vector<E> o = compress_lo(p, to_point(z, m));
  memcpy(out, o.data(), o.size() * 16);
}
void oracle_coeffs(const uint32_t* poly, const uint32_t* w, int k, uint32_t* c0c2) {
  std::vector<E> p((size_t)1 << k), ww((size_t)1 << k);
  memcpy(p.data(), poly, p.size() * 16);
  memcpy(ww.data(), w, ww.size() * 16);
  E c0, c2;
  coeffs(p, ww, c0, c2);
  put_e(c0c2, c0);
  put_e(c0c2 + 4, c2);
}
void oracle_fix_lo_var(uint32_t* poly, int k, const uint32_t* r) {
  std::vector<E> p((size_t)1 << k);
  memcpy(p.data(), poly, p.size() * 16);
  fix_lo_var(p, get_e(r));
  memcpy(poly, p.data(), p.size() * 16);
}
void oracle_combine_eqs(uint32_t* out, int k, const uint32_t* points, int n, const uint32_t* alpha) {
  std::vector<E> o((size_t)1 << k);
  memcpy(o.data(), out, o.size() * 16);
  std::vector<Point> pts;
  for (int i = 0; i < n; ++i) pts.push_back(to_point(points + (size_t)i * 4 * k, k));
  combine_eqs(o, pts, get_e(alpha));
  memcpy(out, o.data(), o.size() * 16);
}
void oracle_combine_pows(uint32_t* out, int k, const uint32_t* vars, int n, const uint32_t* alpha, int shift) {
  std::vector<E> o((size_t)1 << k);
  memcpy(o.data(), out, o.size() * 16);
  std::vector<F> v(n);
  for (int i = 0; i < n; ++i) v[i] = F{vars[i]};
  combine_pows(o, v, get_e(alpha), shift);
  memcpy(out, o.data(), o.size() * 16);
}
void oracle_combine_pows_scaled(uint32_t* out, int k, const uint32_t* vars, int n, const uint32_t* scales) {
  std::vector<E> o((size_t)1 << k);
  memcpy(o.data(), out, o.size() * 16);
  std::vector<F> v(n);
  std::vector<E> sc(n);
  for (int i = 0; i < n; ++i) {
    v[i] = F{vars[i]};
    sc[i] = get_e(scales + 4 * i);
  }
  combine_pows_scaled(o, v, sc);
  memcpy(out, o.data(), o.size() * 16);
}
void oracle_eval_univariate(const uint32_t* poly, size_t n, const uint32_t* x, uint32_t* out) {
  std::vector<E> p(n);
  memcpy(p.data(), poly, n * 16);
  put_e(out, eval_univariate_e(p, get_e(x)));
}

// ---- params ----
int oracle_n_ood_queries(int s, int sec, int d, int r) { return n_ood_queries(s, sec, d, r, 124); }
int oracle_n_stir_queries(int s, int sec, int r) { return n_stir_queries(s, sec, r); }
void oracle_number_of_rounds(int f, int k, int* n_rounds, int* final_rounds) {
  number_of_rounds(f, k, *n_rounds, *final_rounds);
}

// ---- WHIR end to end ----
// Returns 0 on success; -1 if `cap` is too small (len holds the needed size).
int oracle_whir_prove(int k, int folding, int rate, int ir, int soundness, int security, int n_points,
                      const uint32_t* poly, uint8_t* proof, size_t cap, size_t* len, uint32_t* checkpoint,
                      uint32_t* root) {
  Params p{k, folding, rate, ir, soundness, security};
  std::vector<F> v((size_t)1 << k);
  memcpy(v.data(), poly, v.size() * 4);
  ProveResult r = whir_prove(p, v, n_points);
  *len = r.proof.size();
  if (checkpoint) *checkpoint = r.checkpoint.v;
  if (root) memcpy(root, r.root.d, 32);
  if (r.proof.size() > cap) return -1;
  memcpy(proof, r.proof.data(), r.proof.size());
  return 0;
}

int oracle_whir_verify(int k, int folding, int rate, int ir, int soundness, int security, int n_points,
                       const uint8_t* proof, size_t len, uint32_t* checkpoint) {
  Params p{k, folding, rate, ir, soundness, security};
  F cp{0};
  int s = whir_verify(p, proof, len, n_points, cp);
  if (checkpoint) *checkpoint = cp.v;
  return s;
}

// ---- multi-polynomial stacking ----
// Polys cross as one concatenated array in id order; claims as parallel arrays
// (poly id, 1 = draws a fresh point). Exceptions (a malformed stack) return -2.
static std::vector<int> k_vec(int n, const int* k) { return std::vector<int>(k, k + n); }
static std::vector<std::vector<F>> polys_from(int n, const int* k, const uint32_t* data) {
  std::vector<std::vector<F>> out;
  size_t off = 0;
  for (int i = 0; i < n; ++i) {
    const size_t len = (size_t)1 << k[i];
    std::vector<F> v(len);
    memcpy(v.data(), data + off, len * 4);
    off += len;
    out.push_back(std::move(v));
  }
  return out;
}
static std::vector<StackClaim> claims_from(int n, const int* poly, const uint8_t* fresh) {
  std::vector<StackClaim> out((size_t)n);
  for (int i = 0; i < n; ++i) out[i] = StackClaim{poly[i], fresh[i] > 2 ? 1 : (int)fresh[i]};
  return out;
}

int oracle_stack_layout(int n_polys, const int* k_polys, int k_svo, int* k_out, uint64_t* offsets, int* order) {
  try {
    const StackLayout L = stack_layout(k_vec(n_polys, k_polys), k_svo);
    *k_out = L.k;
    for (int i = 0; i < n_polys; ++i) {
      offsets[i] = L.offset[i];
      order[i] = L.order[i];
    }
    return 0;
  } catch (const std::exception&) {
    return -2;
  }
}

int oracle_stack_claim_ranks(int n_polys, const int* k_polys, int k_svo, int n_claims, const int* claim_poly,
                             const uint8_t* claim_new, uint64_t* ranks) {
  try {
    const StackLayout L = stack_layout(k_vec(n_polys, k_polys), k_svo);
    const auto r = stack_claim_ranks(L, claims_from(n_claims, claim_poly, claim_new));
    for (int i = 0; i < n_claims; ++i) ranks[i] = r[i];
    return 0;
  } catch (const std::exception&) {
    return -2;
  }
}

int oracle_stack_virtual_check(int n_polys, const int* k_polys, const uint32_t* polys, int k_svo,
                               const uint32_t* x) {
  try {
    return stack_virtual_check(polys_from(n_polys, k_polys, polys), k_svo, get_e(x));
  } catch (const std::exception&) {
    return -2;
  }
}

int oracle_whir_prove_stacked(int k, int folding, int rate, int ir, int soundness, int security, int n_polys,
                              const int* k_polys, const uint32_t* polys, int n_claims, const int* claim_poly,
                              const uint8_t* claim_new, uint8_t* proof, size_t cap, size_t* len,
                              uint32_t* checkpoint, uint32_t* root) {
  try {
    Params p{k, folding, rate, ir, soundness, security};
    ProveResult r = whir_prove_stacked(p, polys_from(n_polys, k_polys, polys),
                                       claims_from(n_claims, claim_poly, claim_new));
    *len = r.proof.size();
    if (checkpoint) *checkpoint = r.checkpoint.v;
    if (root) memcpy(root, r.root.d, 32);
    if (r.proof.size() > cap) return -1;
    memcpy(proof, r.proof.data(), r.proof.size());
    return 0;
  } catch (const std::exception&) {
    return -2;
  }
}

int oracle_whir_verify_stacked(int k, int folding, int rate, int ir, int soundness, int security, int n_polys,
                               const int* k_polys, int n_claims, const int* claim_poly, const uint8_t* claim_new,
                               const uint8_t* proof, size_t len, uint32_t* checkpoint) {
  try {
    Params p{k, folding, rate, ir, soundness, security};
    F cp{0};
    int s = whir_verify_stacked(p, proof, len, k_vec(n_polys, k_polys),
                                claims_from(n_claims, claim_poly, claim_new), cp);
    if (checkpoint) *checkpoint = cp.v;
    return s;
  } catch (const std::exception&) {
    return -2;
  }
}

// ---- a proof inside a caller's transcript session (whir.h Session) ----
// claim_new: 0 shared, 1 drawn, 2 given (the next k_poly extension elements of
// `given`); prefix: words written before the commitment. challenger_out (40
// words): sponge state 16, input buffer 8, output buffer 8, input length,
// output length, 6 zero words — the state open hands back.
static Session session_from(int n_prefix, const uint32_t* prefix, int n_claims, const int* claim_poly,
                            const uint8_t* claim_new, const int* k_polys, const uint32_t* given) {
  Session S;
  for (int i = 0; i < n_prefix; ++i) S.prefix.push_back(F{prefix[i]});
  size_t off = 0;
  for (int c = 0; c < n_claims; ++c)
    if (claim_new[c] == 2) {
      const int k = k_polys[claim_poly[c]];
      S.given.push_back(to_point(given + off, k));
      off += 4 * (size_t)k;
    }
  return S;
}

int oracle_whir_prove_session(int k, int folding, int rate, int ir, int soundness, int security, int n_polys,
                              const int* k_polys, const uint32_t* polys, int n_claims, const int* claim_poly,
                              const uint8_t* claim_new, int n_prefix, const uint32_t* prefix, const uint32_t* given,
                              uint8_t* proof, size_t cap, size_t* len, uint32_t* checkpoint, uint32_t* root,
                              uint32_t* challenger_out) {
  try {
    Params p{k, folding, rate, ir, soundness, security};
    const Session S = session_from(n_prefix, prefix, n_claims, claim_poly, claim_new, k_polys, given);
    Challenger after(nullptr);
    ProveResult r = whir_prove_stacked(p, polys_from(n_polys, k_polys, polys),
                                       claims_from(n_claims, claim_poly, claim_new), &S, &after);
    *len = r.proof.size();
    if (checkpoint) *checkpoint = r.checkpoint.v;
    if (root) memcpy(root, r.root.d, 32);
    if (challenger_out) {
      memset(challenger_out, 0, 40 * 4);
      for (int i = 0; i < 16; ++i) challenger_out[i] = after.state[i].v;
      for (size_t i = 0; i < after.in.size() && i < 8; ++i) challenger_out[16 + i] = after.in[i].v;
      for (size_t i = 0; i < after.out.size() && i < 8; ++i) challenger_out[24 + i] = after.out[i].v;
      challenger_out[32] = (uint32_t)after.in.size();
      challenger_out[33] = (uint32_t)after.out.size();
    }
    if (r.proof.size() > cap) return -1;
    memcpy(proof, r.proof.data(), r.proof.size());
    return 0;
  } catch (const std::exception&) {
    return -2;
  }
}

int oracle_whir_verify_session(int k, int folding, int rate, int ir, int soundness, int security, int n_polys,
                               const int* k_polys, int n_claims, const int* claim_poly, const uint8_t* claim_new,
                               int n_prefix, const uint32_t* prefix, const uint32_t* given, const uint8_t* proof,
                               size_t len, uint32_t* checkpoint) {
  try {
    Params p{k, folding, rate, ir, soundness, security};
    const Session S = session_from(n_prefix, prefix, n_claims, claim_poly, claim_new, k_polys, given);
    F cp{0};
    int st = whir_verify_stacked(p, proof, len, k_vec(n_polys, k_polys), claims_from(n_claims, claim_poly, claim_new),
                                 cp, &S);
    if (checkpoint) *checkpoint = cp.v;
    return st;
  } catch (const std::exception&) {
    return -2;
  }
}

}  // extern "C"
