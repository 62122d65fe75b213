// Host-side plan of a multi-polynomial stack: where each registered poly sits in
// the stacked vector and the alpha order of the evaluation claims.
//   ProverStack::layout        crates/whir/src/sumcheck/prover.rs:92-124
//   VerifierLayout::new        crates/whir/src/sumcheck/verifier.rs:152-174
//   Selector (derived Ord)     crates/whir/src/sumcheck/mod.rs:30-58
//   claims_in_layout_order     crates/whir/src/sumcheck/prover.rs:155-174
// Polys are sorted by size (stable) and placed from the largest down, so for
// equal sizes the last registered poly gets the lowest offset; every offset is
// a multiple of its poly's size. The layout (a BTreeMap keyed by Selector)
// iterates selectors in lexicographic order of their bits (index >> i) & 1,
// i = 0, 1, ..., with a proper prefix first; a claim's alpha power is its
// position in that order (a poly's claims in the order they were made).
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace whir_host {

struct StackPlan {
  int k = 0;                       // ceil(log2(sum of sizes))
  std::vector<int> k_polys;        // per poly id
  std::vector<uint64_t> offset;    // per poly id: first index in the stacked vector
  std::vector<int> order;          // poly ids in Selector order
  std::vector<int> claim_poly;     // per claim (transcript order)
  std::vector<uint8_t> claim_fresh;
  std::vector<uint32_t> rank;      // per claim: alpha exponent
};

inline bool selector_before(int ka, uint64_t ia, int kb, uint64_t ib) {
  const int n = std::min(ka, kb);
  for (int i = 0; i < n; ++i) {
    const int a = (int)((ia >> i) & 1), b = (int)((ib >> i) & 1);
    if (a != b) return a < b;
  }
  return ka < kb;
}

// false when the stack or its claims are malformed: no polys, a poly smaller
// than the folding factor (SvoPoint needs k_poly >= k_svo) or larger than 2^30,
// a first claim that reuses a point, a reused point across sizes.
inline bool make_stack_plan(const int* k_polys, int n_polys, const int* claim_poly, const uint8_t* fresh,
                            int n_claims, int f, StackPlan& out) {
  if (n_polys < 1 || n_claims < 0 || !k_polys || (n_claims > 0 && (!claim_poly || !fresh))) return false;
  StackPlan S;
  S.k_polys.assign(k_polys, k_polys + n_polys);
  uint64_t total = 0;
  for (int kp : S.k_polys) {
    if (kp < f || kp > 30) return false;
    total += 1ull << kp;
  }
  while ((1ull << S.k) < total) ++S.k;
  std::vector<int> ids((size_t)n_polys);
  for (int i = 0; i < n_polys; ++i) ids[i] = i;
  std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) { return S.k_polys[a] < S.k_polys[b]; });
  S.offset.assign((size_t)n_polys, 0);
  uint64_t off = 0;
  for (size_t j = ids.size(); j-- > 0;) {
    S.offset[ids[j]] = off;
    off += 1ull << S.k_polys[ids[j]];
  }
  auto sel_k = [&](int id) { return S.k - S.k_polys[id]; };
  auto sel_index = [&](int id) { return S.offset[id] >> S.k_polys[id]; };
  S.order = ids;
  std::sort(S.order.begin(), S.order.end(),
            [&](int a, int b) { return selector_before(sel_k(a), sel_index(a), sel_k(b), sel_index(b)); });
  for (int c = 0; c < n_claims; ++c) {
    const int id = claim_poly[c];
    if (id < 0 || id >= n_polys) return false;
    if (!fresh[c] && (c == 0 || S.k_polys[claim_poly[c - 1]] != S.k_polys[id])) return false;
    S.claim_poly.push_back(id);
    S.claim_fresh.push_back(fresh[c] ? 1 : 0);
  }
  S.rank.assign((size_t)n_claims, 0);
  uint32_t r = 0;
  for (int id : S.order)
    for (int c = 0; c < n_claims; ++c)
      if (S.claim_poly[c] == id) S.rank[c] = r++;
  out = std::move(S);
  return true;
}

}  // namespace whir_host
