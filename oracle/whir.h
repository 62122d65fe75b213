// ORACLE — test infrastructure only (see field.h).
//
// CPU restatement of the reference's WHIR prover and verifier for one
// multilinear polynomial with n evaluation claims (the path exercised by
// `run_whir_poseidon`, crates/whir/src/test.rs:174-254), and for a stack of
// several polynomials committed as one (ProverStack / ProverLayout):
//   params      crates/whir/src/params.rs:6-202, whir.rs:140-166
//   poly / eq   crates/poly/src/{point,poly,eq,lib}.rs
//   DFT         Plonky3 Radix2DFTSmallBatch semantics (natural order, subgroup)
//   Merkle      crates/merkle/src/{poseidon,poseidon_packed}.rs (split-half tree)
//   sumcheck    crates/whir/src/sumcheck/{mod,expr,svo,combine,prover,verifier}.rs
//   protocol    crates/whir/src/whir.rs:19-522
// Each function cites the reference lines it follows in whir.cpp.
#pragma once
#include <cstdint>
#include <vector>
#include "field.h"
#include "poseidon2.h"
#include "transcript.h"

namespace oracle {

using Point = std::vector<E>;

enum Soundness { UNIQUE_DECODING = 0, JOHNSON_BOUND = 1, CAPACITY_BOUND = 2 };

struct Params {
  int k, folding, rate, initial_reduction;
  int soundness, security_level;
};

int n_ood_queries(int soundness, int security, int log_degree, int log_inv_rate, int field_bits);
int n_stir_queries(int soundness, int security, int log_inv_rate);
void number_of_rounds(int factor, int num_variables, int& n_rounds, int& final_rounds);
void round_params(const Params& p, int round, int& k_folded_domain, int& this_rate, int& prev_rate);

// ---- poly ----
std::vector<E> eq_table(const Point& z, E scale);
Point expand(int k, E x);
E eval_eq_xy(const Point& x, const Point& y);
E eval_pow_xy(const std::vector<F>& x, const Point& y);
E eval_base(const std::vector<F>& poly, const Point& z);
E eval_ext(const std::vector<E>& poly, const Point& z);
std::vector<E> compress_hi(const std::vector<F>& poly, const Point& z);
std::vector<E> compress_lo(const std::vector<F>& poly, const Point& z);
std::vector<E> compress_lo_ext(const std::vector<E>& poly, const Point& z);
void fix_lo_var(std::vector<E>& poly, E r);
E eval_univariate_e(const std::vector<E>& poly, E x);
E eval_univariate_f(const std::vector<F>& poly, E x);
void coeffs(const std::vector<E>& poly, const std::vector<E>& w, E& c0, E& c2);
E extrapolate_012(E e0, E e1, E e2, E r);
void combine_eqs(std::vector<E>& out, const std::vector<Point>& points, E alpha);
void combine_pows(std::vector<E>& out, const std::vector<F>& vars, E alpha, size_t shift);
void combine_pows_scaled(std::vector<E>& out, const std::vector<F>& vars, const std::vector<E>& scales);

// ---- DFT: each column of a row-major height x width matrix, in place ----
void dft_batch(F* mat, size_t height, size_t width);
void dft_batch_naive(F* mat, size_t height, size_t width);

// ---- Merkle tree (row = `row_base` base elements, flattened) ----
struct Digest { F d[8]; };
struct Tree {
  size_t height = 0;
  size_t row_base = 0;
  std::vector<F> data;                      // height * row_base
  std::vector<std::vector<Digest>> layers;  // layers[0] leaves .. layers[k] root
  Digest root() const { return layers.back()[0]; }
  int k() const { return (int)layers.size() - 1; }
};
Tree merkle_commit(std::vector<F>&& data, size_t height, size_t row_base);
void merkle_query(Writer& w, const Tree& t, size_t index);
int merkle_verify(Reader& r, const Digest& comm, size_t index, size_t row_base, int k,
                  std::vector<F>& row);

// ---- WHIR ----
struct ProveResult {
  std::vector<uint8_t> proof;
  F checkpoint;
  Digest root;
};
ProveResult whir_prove(const Params& p, const std::vector<F>& poly, int n_points);
int whir_verify(const Params& p, const uint8_t* proof, size_t len, int n_points, F& checkpoint);

// ---- multi-polynomial stacking (sumcheck/prover.rs:31-124,137-277; verifier.rs:100-212; mod.rs:30-58) ----
// One evaluation claim: the poly it is on and whether it draws a fresh point
// (expand(k_poly, draw())) or reuses the previous claim's point (then the two
// polys have the same k). The list is in transcript order.
struct StackClaim {
  int poly;
  int new_point;  // 0: the previous claim's point, 1: expand(k_poly, draw()), 2: the next given point
};
struct StackLayout {
  int k = 0;                      // log2_ceil(sum of poly sizes)
  std::vector<size_t> offset;     // per poly id: first index in the stacked vector
  std::vector<int> sel_k;         // per poly id: selector variables, k - k_poly
  std::vector<size_t> sel_index;  // per poly id: selector index, offset >> k_poly
  std::vector<int> order;         // poly ids in Selector order (the BTreeMap's iteration)
};
bool selector_less(int ka, size_t ia, int kb, size_t ib);
StackLayout stack_layout(const std::vector<int>& k_polys, int k_svo);
std::vector<F> stack_polys(const StackLayout& L, const std::vector<std::vector<F>>& polys);
// alpha exponent of each claim: its position in claims_in_layout_order (prover.rs:155-174)
std::vector<size_t> stack_claim_ranks(const StackLayout& L, const std::vector<StackClaim>& claims);
// A session around the proof, as a caller of Whir::commit / ProverStack::eval /
// Whir::open drives it (whir.rs:127-180, prover.rs:64-90): `prefix` words are
// written to the transcript before the commitment (their bytes start the
// proof), and a claim with new_point == 2 takes the next of `given` (any
// point of k_poly coordinates) instead of expand(k_poly, draw()). `after`
// receives the challenger state open hands back (before the checkpoint draw).
struct Session {
  std::vector<F> prefix;
  std::vector<Point> given;
};
ProveResult whir_prove_stacked(const Params& p, const std::vector<std::vector<F>>& polys,
                               const std::vector<StackClaim>& claims, const Session* session = nullptr,
                               Challenger* after = nullptr);
int whir_verify_stacked(const Params& p, const uint8_t* proof, size_t len, const std::vector<int>& k_polys,
                        const std::vector<StackClaim>& claims, F& checkpoint, const Session* session = nullptr);
// write_virtual_eval's value and SVO accumulators computed per stacked poly
// (prover.rs:241-256) against one claim on the stacked vector (its debug check,
// prover.rs:258-266): returns 1 when both agree.
int stack_virtual_check(const std::vector<std::vector<F>>& polys, int k_svo, E x);

std::vector<F> rand_poly(uint64_t seed, int k);  // Poly::<F>::rand(&mut rng(seed), k)

}  // namespace oracle
