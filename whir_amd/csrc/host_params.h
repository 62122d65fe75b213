// Host-side protocol parameters and the proof schedule (no device code).
//   soundness formulas: crates/whir/src/params.rs:6-202
//   round parameters:   crates/whir/src/whir.rs:140-166
// These fix the proof shape (OOD/STIR counts, number of rounds), so the f64
// arithmetic follows the reference expression by expression.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace whir_host {

enum { UD = 0, JB = 1, CB = 2 };
constexpr double LOG2_10 = 3.32192809488736234787031942948939018;
constexpr int EXT_FIELD_BITS = 124;  // Ext::bits() for KoalaBear^4

inline double log_eta(int s, int r) {
  if (s == UD) return 0.;
  if (s == JB) return -(0.5 * (double)r + LOG2_10 + 1.);
  return -((double)r + LOG2_10 + 1.);
}
inline double list_size_bits(int s, int d, int r) {
  const double le = log_eta(s, r);
  if (s == UD) return 0.;
  if (s == JB) return (double)r / 2. - (1. + le);
  return (double)(d + r) - le;
}
inline double log_1_delta(int s, int r) {
  const double le = log_eta(s, r);
  const double eta = std::pow(2.0, le);
  const double rate = 1. / (double)(1 << r);
  double delta;
  if (s == UD) delta = 0.5 * (1. - rate);
  else if (s == JB) delta = 1. - std::sqrt(rate) - eta;
  else delta = 1. - rate - eta;
  return std::log2(1. - delta);
}
inline int n_stir_queries(int s, int sec, int r) { return (int)std::ceil(-(double)sec / log_1_delta(s, r)); }
inline double ood_error(int s, int d, int r, int fb, int n) {
  if (s == UD) return 0.;
  const double lsb = list_size_bits(s, d, r);
  const double err = 2. * lsb + (double)(d * n);
  return (double)(n * fb) + 1. - err;
}
inline int n_ood_queries(int s, int sec, int d, int r) {
  if (s == UD) return 0;
  for (int n = 1; n < 64; ++n)
    if (ood_error(s, d, r, EXT_FIELD_BITS, n) >= (double)sec) return n;
  return -1;
}
inline void number_of_rounds(int factor, int nv, int& n_rounds, int& final_rounds) {
  if (nv <= 6) {
    n_rounds = 0;
    final_rounds = nv - factor;
    return;
  }
  const int nr = (nv - 6 + factor - 1) / factor;
  final_rounds = nv - nr * factor;
  n_rounds = nr - 1;
}

struct RoundPlan {
  int kc;        // variables of the polynomial committed in this round
  int kfd;       // k_folded_domain: depth of the tree the STIR queries open
  int this_rate, prev_rate;
  int log_rows;  // rows of this round's codeword (2^f extension elements each)
  int n_ood, n_stir;
};

struct Plan {
  int k, f, rate, ir, soundness, security, n_points;
  int n_ood0, n_rounds, final_k;
  std::vector<RoundPlan> rounds;
  int final_kfd, final_n_stir;
  uint64_t log_words;  // proof length upper bound (words)
  bool ok = false;
};

inline Plan make_plan(int k, int f, int rate, int ir, int soundness, int security, int n_points) {
  Plan p;
  p.k = k; p.f = f; p.rate = rate; p.ir = ir; p.soundness = soundness; p.security = security; p.n_points = n_points;
  if (k < 1 || f < 1 || f > k || rate < 0 || ir < 1 || ir > f || soundness < 0 || soundness > 2 || security < 1 ||
      n_points < 0)
    return p;
  p.n_ood0 = n_ood_queries(soundness, security, k, rate);
  if (p.n_ood0 < 0) return p;
  int fin;
  number_of_rounds(f, k, p.n_rounds, fin);
  auto round_params = [&](int round, int& kfd, int& this_rate, int& prev_rate) {
    const int prev_red = round > 0 ? ir + (round - 1) : 0;
    const int this_red = ir + round;
    kfd = rate + k - prev_red - f;
    prev_rate = rate + round * f - prev_red;
    this_rate = rate + (round + 1) * f - this_red;
  };
  const uint64_t base_row = 1ull << f, ext_row = 4ull << f;
  uint64_t words = 8 + 4ull * (n_points + p.n_ood0) + 8ull * f;
  int kc = k - f;
  for (int r = 0; r < p.n_rounds; ++r) {
    RoundPlan rp;
    rp.kc = kc;
    round_params(r, rp.kfd, rp.this_rate, rp.prev_rate);
    rp.log_rows = kc - f + rp.this_rate;
    rp.n_ood = n_ood_queries(soundness, security, kc, rp.this_rate);
    rp.n_stir = n_stir_queries(soundness, security, rp.prev_rate);
    if (rp.n_ood < 0) return p;
    words += 8 + 4ull * rp.n_ood + (uint64_t)rp.n_stir * ((r == 0 ? base_row : ext_row) + 8ull * rp.kfd) + 8ull * f;
    p.rounds.push_back(rp);
    kc -= f;
  }
  p.final_k = kc;
  if (kc != fin) return p;
  int this_rate, prev_rate;
  round_params(p.n_rounds, p.final_kfd, this_rate, prev_rate);
  p.final_n_stir = n_stir_queries(soundness, security, prev_rate);
  words += 4ull << kc;
  words += (uint64_t)p.final_n_stir * ((p.n_rounds > 0 ? ext_row : base_row) + 8ull * p.final_kfd);
  words += 1;
  p.log_words = words;
  p.ok = true;
  return p;
}

}  // namespace whir_host
