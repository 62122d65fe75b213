// ORACLE — test infrastructure only (see field.h).
// CPU restatement of the reference WHIR prover/verifier; reference lines cited
// per function. OpenMP parallelises the large loops; all arithmetic is exact, so
// results do not depend on the thread count or reduction order.
#include "whir.h"
#include <algorithm>
#include <array>
#include <functional>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <omp.h>
#include "rng.h"

namespace oracle {

// ======================= params (crates/whir/src/params.rs) =======================
static const double LOG2_10 = 3.32192809488736234787031942948939018;

static double log_eta(int s, int r) {  // params.rs:30-40
  switch (s) {
    case UNIQUE_DECODING: return 0.;
    case JOHNSON_BOUND: return -(0.5 * (double)r + LOG2_10 + 1.);
    default: return -((double)r + LOG2_10 + 1.);
  }
}
static double list_size_bits(int s, int d, int r) {  // params.rs:44-58
  double le = log_eta(s, r);
  switch (s) {
    case UNIQUE_DECODING: return 0.;
    case JOHNSON_BOUND: return (double)r / 2. - (1. + le);
    default: return (double)(d + r) - le;
  }
}
static double log_1_delta(int s, int r) {  // params.rs:103-116
  double le = log_eta(s, r);
  double eta = std::pow(2.0, le);
  double rate = 1. / (double)(1 << r);
  double delta;
  switch (s) {
    case UNIQUE_DECODING: delta = 0.5 * (1. - rate); break;
    case JOHNSON_BOUND: delta = 1. - std::sqrt(rate) - eta; break;
    default: delta = 1. - rate - eta; break;
  }
  return std::log2(1. - delta);
}
int n_stir_queries(int s, int sec, int r) {  // params.rs:118-122
  double q = -(double)sec / log_1_delta(s, r);
  return (int)std::ceil(q);
}
static double ood_error(int s, int d, int r, int fb, int n) {  // params.rs:137-155
  if (s == UNIQUE_DECODING) return 0.;
  double lsb = list_size_bits(s, d, r);
  double err = 2. * lsb + (double)(d * n);
  return (double)(n * fb) + 1. - err;
}
int n_ood_queries(int s, int sec, int d, int r, int fb) {  // params.rs:160-180
  if (s == UNIQUE_DECODING) return 0;
  for (int n = 1; n < 64; ++n)
    if (ood_error(s, d, r, fb, n) >= (double)sec) return n;
  throw std::runtime_error("Could not find an appropriate number of OOD samples");
}
void number_of_rounds(int factor, int nv, int& n_rounds, int& final_rounds) {  // params.rs:190-202
  if (nv <= 6) {
    n_rounds = 0;
    final_rounds = nv - factor;
    return;
  }
  int nr = (nv - 6 + factor - 1) / factor;
  final_rounds = nv - nr * factor;
  n_rounds = nr - 1;
}
void round_params(const Params& p, int round, int& kfd, int& this_rate, int& prev_rate) {
  // whir.rs:149-166
  int prev_red = round > 0 ? p.initial_reduction + (round - 1) : 0;
  int this_red = p.initial_reduction + round;
  int k_domain = p.rate + p.k - prev_red;
  kfd = k_domain - p.folding;
  prev_rate = p.rate + round * p.folding - prev_red;
  this_rate = p.rate + (round + 1) * p.folding - this_red;
}
static const int EXT_BITS = 124;  // Ext::bits(): bit length of p^4

// ======================= poly (crates/poly/src) =======================
std::vector<E> eq_table(const Point& z, E scale) {  // point.rs:110-126
  size_t k = z.size();
  std::vector<E> eq((size_t)1 << k, ezero());
  eq[0] = scale;
  for (size_t i = 0; i < k; ++i) {
    size_t h = (size_t)1 << i;
    for (size_t j = 0; j < h; ++j) {
      eq[j + h] = emul(eq[j], z[i]);
      eq[j] = esub(eq[j], eq[j + h]);
    }
  }
  return eq;
}

Point expand(int k, E x) {  // point.rs:66-75
  Point p(k);
  for (int i = 0; i < k; ++i) {
    p[i] = x;
    x = esq(x);
  }
  return p;
}

E eval_eq_xy(const Point& x, const Point& y) {  // poly/src/lib.rs:17-23
  E acc = eone();
  for (size_t i = 0; i < x.size(); ++i) {
    E t = esub(esub(edbl(emul(y[i], x[i])), x[i]), y[i]);
    acc = emul(acc, eadd(t, eone()));
  }
  return acc;
}

E eval_pow_xy(const std::vector<F>& x, const Point& y) {  // poly/src/lib.rs:25-31
  E acc = eone();
  for (size_t i = 0; i < x.size(); ++i) {
    E t = esub(emul_base(y[i], x[i]), y[i]);
    acc = emul(acc, eadd(t, eone()));
  }
  return acc;
}

E eval_base(const std::vector<F>& poly, const Point& z) {  // eq.rs:107-177 (= eval_poly_reference)
  std::vector<E> eq = eq_table(z, eone());
  E acc = ezero();
#pragma omp parallel
  {
    E loc = ezero();
#pragma omp for nowait
    for (long long i = 0; i < (long long)poly.size(); ++i) loc = eadd(loc, emul_base(eq[i], poly[i]));
#pragma omp critical
    acc = eadd(acc, loc);
  }
  return acc;
}

E eval_ext(const std::vector<E>& poly, const Point& z) {  // eq.rs:179-250
  std::vector<E> eq = eq_table(z, eone());
  E acc = ezero();
#pragma omp parallel
  {
    E loc = ezero();
#pragma omp for nowait
    for (long long i = 0; i < (long long)poly.size(); ++i) loc = eadd(loc, emul(eq[i], poly[i]));
#pragma omp critical
    acc = eadd(acc, loc);
  }
  return acc;
}

std::vector<E> compress_hi(const std::vector<F>& poly, const Point& z) {
  // eq.rs:12-29 (reference) / 290-398: bind the top |z| index bits to z.
  std::vector<E> eq = eq_table(z, eone());
  size_t inner = poly.size() / eq.size();
  std::vector<E> out(inner, ezero());
#pragma omp parallel
  {
    std::vector<E> loc(inner, ezero());
#pragma omp for nowait
    for (long long h = 0; h < (long long)eq.size(); ++h)
      for (size_t i = 0; i < inner; ++i) loc[i] = eadd(loc[i], emul_base(eq[h], poly[h * inner + i]));
#pragma omp critical
    for (size_t i = 0; i < inner; ++i) out[i] = eadd(out[i], loc[i]);
  }
  return out;
}

std::vector<E> compress_lo(const std::vector<F>& poly, const Point& z) {
  // eq.rs:31-49 / 400-539: bind the low |z| index bits to z.
  std::vector<E> eq = eq_table(z, eone());
  size_t m = eq.size(), n = poly.size() / m;
  std::vector<E> out(n);
#pragma omp parallel for
  for (long long j = 0; j < (long long)n; ++j) {
    E acc = ezero();
    for (size_t i = 0; i < m; ++i) acc = eadd(acc, emul_base(eq[i], poly[j * m + i]));
    out[j] = acc;
  }
  return out;
}

std::vector<E> compress_lo_ext(const std::vector<E>& poly, const Point& z) {
  std::vector<E> eq = eq_table(z, eone());
  size_t m = eq.size(), n = poly.size() / m;
  std::vector<E> out(n);
  for (size_t j = 0; j < n; ++j) {
    E acc = ezero();
    for (size_t i = 0; i < m; ++i) acc = eadd(acc, emul(eq[i], poly[j * m + i]));
    out[j] = acc;
  }
  return out;
}

void fix_lo_var(std::vector<E>& a, E r) {  // poly.rs:131-142
  size_t n = a.size() / 2;
  std::vector<E> out(n);
#pragma omp parallel for
  for (long long i = 0; i < (long long)n; ++i)
    out[i] = eadd(emul(esub(a[2 * i + 1], a[2 * i]), r), a[2 * i]);
  a.swap(out);
}

E eval_univariate_e(const std::vector<E>& poly, E x) {  // utils.rs:93-115 (Horner)
  E acc = ezero();
  for (size_t i = poly.size(); i-- > 0;) acc = eadd(emul(acc, x), poly[i]);
  return acc;
}
E eval_univariate_f(const std::vector<F>& poly, E x) {
  E acc = ezero();
  for (size_t i = poly.size(); i-- > 0;) acc = eadd(emul(acc, x), efrom(poly[i]));
  return acc;
}

void coeffs(const std::vector<E>& f, const std::vector<E>& w, E& c0, E& c2) {  // expr.rs:14-27
  E a0 = ezero(), a2 = ezero();
  size_t n = f.size() / 2;
#pragma omp parallel
  {
    E l0 = ezero(), l2 = ezero();
#pragma omp for nowait
    for (long long i = 0; i < (long long)n; ++i) {
      l0 = eadd(l0, emul(w[2 * i], f[2 * i]));
      E wd = esub(edbl(w[2 * i + 1]), w[2 * i]);
      E fd = esub(edbl(f[2 * i + 1]), f[2 * i]);
      l2 = eadd(l2, emul(wd, fd));
    }
#pragma omp critical
    {
      a0 = eadd(a0, l0);
      a2 = eadd(a2, l2);
    }
  }
  c0 = a0;
  c2 = a2;
}

static void lagrange_weights_012(E r, E out[3]) {  // sumcheck/mod.rs:17-23
  E inv2 = efrom(finv(F::from_u32(2)));
  E one = eone(), two = efrom(F::from_u32(2));
  out[0] = emul(emul(esub(r, one), esub(r, two)), inv2);
  out[1] = emul(r, esub(two, r));
  out[2] = emul(emul(r, esub(r, one)), inv2);
}

E extrapolate_012(E e0, E e1, E e2, E r) {  // sumcheck/mod.rs:25-28
  E w[3];
  lagrange_weights_012(r, w);
  return eadd(eadd(emul(e0, w[0]), emul(e1, w[1])), emul(e2, w[2]));
}

static E horner(const std::vector<E>& v, E x) {  // common/src/utils.rs:125-127
  E acc = ezero();
  for (size_t i = v.size(); i-- > 0;) acc = eadd(emul(acc, x), v[i]);
  return acc;
}

void combine_eqs(std::vector<E>& out, const std::vector<Point>& points, E alpha) {
  // sumcheck/combine.rs:178-203: out += sum_j alpha^j eq(z_j)
  E a = eone();
  for (const Point& z : points) {
    std::vector<E> eq = eq_table(z, a);
#pragma omp parallel for
    for (long long i = 0; i < (long long)out.size(); ++i) out[i] = eadd(out[i], eq[i]);
    a = emul(a, alpha);
  }
}

void combine_pows_scaled(std::vector<E>& out, const std::vector<F>& vars, const std::vector<E>& scales) {
  // sumcheck/combine.rs:440-468: out[i] += sum_j scales_j * vars_j^i
  const long long n = (long long)out.size();
  const long long B = 1024;
#pragma omp parallel for
  for (long long b = 0; b < (n + B - 1) / B; ++b) {
    long long lo = b * B, hi = std::min(n, lo + B);
    for (size_t j = 0; j < vars.size(); ++j) {
      F pw = fpow(vars[j], (uint64_t)lo);
      for (long long i = lo; i < hi; ++i) {
        out[i] = eadd(out[i], emul_base(scales[j], pw));
        pw = fmul(pw, vars[j]);
      }
    }
  }
}

void combine_pows(std::vector<E>& out, const std::vector<F>& vars, E alpha, size_t shift) {
  // sumcheck/combine.rs:407-438: out[i] += sum_j alpha^(shift+j) vars_j^i
  std::vector<E> scales(vars.size());
  E a = epow(alpha, shift);
  for (size_t j = 0; j < vars.size(); ++j) {
    scales[j] = a;
    a = emul(a, alpha);
  }
  combine_pows_scaled(out, vars, scales);
}

// ======================= DFT =======================
void dft_batch(F* mat, size_t h, size_t w) {
  // Plonky3 TwoAdicSubgroupDft::dft_batch: out[i][c] = sum_j in[j][c] * omega^(ij),
  // omega = two_adic_generator(log2 h), natural order in and out.
  if (h <= 1) return;
  int logh = 0;
  while (((size_t)1 << logh) < h) ++logh;
  std::vector<F> tw(h / 2);
  F g = two_adic_generator(logh);
  tw[0] = fone();
  for (size_t i = 1; i < h / 2; ++i) tw[i] = fmul(tw[i - 1], g);
#pragma omp parallel for
  for (long long c = 0; c < (long long)w; ++c) {
    std::vector<F> a(h);
    for (size_t j = 0; j < h; ++j) a[j] = mat[j * w + c];
    for (size_t i = 1, j = 0; i < h; ++i) {  // bit reversal
      size_t bit = h >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= h; len <<= 1) {
      size_t half = len / 2, step = h / len;
      for (size_t i = 0; i < h; i += len)
        for (size_t j = 0; j < half; ++j) {
          F u = a[i + j], v = fmul(a[i + j + half], tw[j * step]);
          a[i + j] = fadd(u, v);
          a[i + j + half] = fsub(u, v);
        }
    }
    for (size_t j = 0; j < h; ++j) mat[j * w + c] = a[j];
  }
}

void dft_batch_naive(F* mat, size_t h, size_t w) {
  int logh = 0;
  while (((size_t)1 << logh) < h) ++logh;
  F g = two_adic_generator(logh);
  std::vector<F> out(h * w, F{0});
  for (size_t i = 0; i < h; ++i) {
    F gi = fpow(g, i), x = fone();
    for (size_t j = 0; j < h; ++j) {
      for (size_t c = 0; c < w; ++c) out[i * w + c] = fadd(out[i * w + c], fmul(mat[j * w + c], x));
      x = fmul(x, gi);
    }
  }
  memcpy(mat, out.data(), h * w * sizeof(F));
}

// ======================= Merkle =======================
Tree merkle_commit(std::vector<F>&& data, size_t height, size_t row_base) {
  // merkle/src/poseidon.rs:61-74,98-130; poseidon_packed.rs:96-143,268-304:
  // leaves = sponge24(row), parent[j] = C(layer[j], layer[j + n/2]).
  const Hashers& H = Hashers::get();
  Tree t;
  t.height = height;
  t.row_base = row_base;
  t.data = std::move(data);
  t.layers.emplace_back(height);
  auto& leaves = t.layers[0];
#pragma omp parallel for
  for (long long i = 0; i < (long long)height; ++i) sponge24(H.p24, &t.data[i * row_base], row_base, leaves[i].d);
  while (t.layers.back().size() > 1) {
    const auto& prev = t.layers.back();
    size_t half = prev.size() / 2;
    std::vector<Digest> next(half);
#pragma omp parallel for
    for (long long j = 0; j < (long long)half; ++j) compress16(H.p16, prev[j].d, prev[j + half].d, next[j].d);
    t.layers.push_back(std::move(next));
  }
  return t;
}

void merkle_query(Writer& w, const Tree& t, size_t index) {
  // poseidon.rs:192-242 / poseidon_packed.rs:145-233: hint the flattened row,
  // then k sibling digests bottom-up (sibling = index ^ mid, index &= mid - 1).
  for (size_t i = 0; i < t.row_base; ++i) w.hint_f(t.data[index * t.row_base + i]);
  int k = t.k();
  size_t idx = index;
  for (int i = 0; i < k; ++i) {
    size_t mid = (size_t)1 << (k - i - 1);
    w.hint_digest(t.layers[i][idx ^ mid].d);
    idx &= mid - 1;
  }
}

int merkle_verify(Reader& r, const Digest& comm, size_t index, size_t row_base, int k,
                  std::vector<F>& row) {
  // poseidon.rs:11-30,244-263 (verify_merkle_proof)
  const Hashers& H = Hashers::get();
  row.resize(row_base);
  for (size_t i = 0; i < row_base; ++i)
    if (r.hint_f(row[i]) != OK) return ERR_TRANSCRIPT;
  std::vector<Digest> wit(k);
  for (int i = 0; i < k; ++i)
    if (r.hint_digest(wit[i].d) != OK) return ERR_TRANSCRIPT;
  Digest acc;
  sponge24(H.p24, row.data(), row_base, acc.d);
  size_t idx = index;
  for (int i = 0; i < k; ++i) {
    size_t mid = (size_t)1 << (k - i - 1);
    Digest nxt;
    if (idx >= mid) compress16(H.p16, wit[i].d, acc.d, nxt.d);
    else compress16(H.p16, acc.d, wit[i].d, nxt.d);
    acc = nxt;
    idx &= mid - 1;
  }
  for (int i = 0; i < 8; ++i)
    if (acc.d[i] != comm.d[i]) return ERR_VERIFY;
  return OK;
}

// ======================= SVO (sumcheck/svo.rs) =======================
struct SvoClaim {
  Point svo, rest;
  std::vector<std::vector<E>> partial;  // partial[l-1] has 2^l entries
  E eval;
};

static SvoClaim svo_eval(int f, const Point& point, const std::vector<F>& poly) {
  // svo.rs:64-91 (SvoPoint::new / eval) and svo.rs:18-38 (svo_partial_evals)
  SvoClaim c;
  c.svo.assign(point.begin(), point.begin() + f);
  c.rest.assign(point.begin() + f, point.end());
  std::vector<E> comp = compress_hi(poly, c.rest);
  std::vector<E> eqs = eq_table(c.svo, eone());
  c.eval = ezero();
  for (size_t i = 0; i < comp.size(); ++i) c.eval = eadd(c.eval, emul(comp[i], eqs[i]));
  for (int l = 1; l <= f; ++l) {
    Point rest(c.svo.begin() + l, c.svo.end());
    std::vector<E> eqr = eq_table(rest, eone());
    size_t n = (size_t)1 << l;
    std::vector<E> out(n, ezero());
    for (size_t b = 0; b < eqr.size(); ++b)
      for (size_t a = 0; a < n; ++a) out[a] = eadd(out[a], emul(comp[b * n + a], eqr[b]));
    c.partial.push_back(std::move(out));
  }
  return c;
}

struct SvoCoeffs {  // svo.rs:132-138
  std::vector<std::vector<F>> eq0, eq2, points0, points2;
};

static std::vector<SvoCoeffs> svo_new(int l) {  // svo.rs:157-222
  auto extend = [](const std::vector<F>& prev, F zi) {
    std::vector<F> eq(prev.size() * 2);
    for (size_t j = 0; j < prev.size(); ++j) {
      eq[j] = fmul(prev[j], fsub(fone(), zi));
      eq[prev.size() + j] = fmul(prev[j], zi);
    }
    return eq;
  };
  std::vector<std::vector<F>> prev_eqs{{fone()}}, prev_pts{{}};
  std::vector<SvoCoeffs> out;
  F vals[3] = {F{0}, fone(), F::from_u32(2)};
  for (int r = 0; r < l; ++r) {
    size_t n = prev_eqs.size();
    std::vector<std::vector<F>> ne, np;
    for (int v = 0; v < 3; ++v)
      for (size_t i = 0; i < n; ++i) {
        ne.push_back(extend(prev_eqs[i], vals[v]));
        auto p = prev_pts[i];
        p.push_back(vals[v]);
        np.push_back(p);
      }
    SvoCoeffs c;
    c.eq0.assign(ne.begin(), ne.begin() + n);
    c.eq2.assign(ne.begin() + 2 * n, ne.end());
    c.points0.assign(np.begin(), np.begin() + n);
    c.points2.assign(np.begin() + 2 * n, np.end());
    out.push_back(std::move(c));
    prev_eqs.swap(ne);
    prev_pts.swap(np);
  }
  return out;
}

// accumulators[round][0|1] = acc0 / acc2 vectors (3^round entries)
using Accs = std::vector<std::array<std::vector<E>, 2>>;

static Accs svo_accumulators(const std::vector<SvoCoeffs>& svo, const std::vector<const SvoClaim*>& claims,
                             const std::vector<E>& alphas) {
  // svo.rs:224-278 (calculate_accumulators); all claims share one svo point.
  int k = (int)svo.size();
  Accs out(k);
  const Point& pt = claims[0]->svo;
  for (int ri = 0; ri < k; ++ri) {
    int l = ri + 1;
    std::vector<E> comb((size_t)1 << l, ezero());
    for (size_t c = 0; c < claims.size(); ++c)
      for (size_t a = 0; a < comb.size(); ++a)
        comb[a] = eadd(comb[a], emul(alphas[c], claims[c]->partial[ri][a]));
    Point active(pt.begin(), pt.begin() + l);
    for (int side = 0; side < 2; ++side) {
      const auto& pts = side == 0 ? svo[ri].points0 : svo[ri].points2;
      const auto& eqs = side == 0 ? svo[ri].eq0 : svo[ri].eq2;
      std::vector<E> acc(pts.size());
      for (size_t u = 0; u < pts.size(); ++u) {
        Point ue(pts[u].size());
        for (size_t i = 0; i < ue.size(); ++i) ue[i] = efrom(pts[u][i]);
        E dot = ezero();
        for (size_t a = 0; a < comb.size(); ++a) dot = eadd(dot, emul_base(comb[a], eqs[u][a]));
        acc[u] = emul(eval_eq_xy(ue, active), dot);
      }
      out[ri][side] = std::move(acc);
    }
  }
  return out;
}

static std::vector<E> lagrange_weights_012_multi(const Point& rs) {  // svo.rs:7-16
  std::vector<E> w{eone()};
  for (const E& r : rs) {
    E l[3];
    lagrange_weights_012(r, l);
    std::vector<E> nw;
    for (int v = 0; v < 3; ++v)
      for (const E& x : w) nw.push_back(emul(x, l[v]));
    w.swap(nw);
  }
  return w;
}

// ======================= sumcheck prover state =======================
struct Prover {
  E sum;
  Point rs;
  std::vector<E> poly, weights;
  E alpha;
};

template <class T>
static E round_step(T& tx, Prover& sc) {  // expr.rs:52-70
  E c0, c2;
  coeffs(sc.poly, sc.weights, c0, c2);
  tx.write_e(c0);
  tx.write_e(c2);
  E r = tx.draw_e();
  fix_lo_var(sc.poly, r);
  fix_lo_var(sc.weights, r);
  sc.sum = extrapolate_012(c0, esub(sc.sum, c0), c2, r);
  return r;
}

struct EqClaimE { Point point; E eval; };
struct PowClaimE { F var; E eval; };

static void combine_claims(Prover& sc, E alpha, const std::vector<EqClaimE>& eqc,
                           const std::vector<PowClaimE>& powc) {  // combine.rs:37-64
  std::vector<E> ev;
  for (auto& c : eqc) ev.push_back(c.eval);
  for (auto& c : powc) ev.push_back(c.eval);
  sc.sum = eadd(sc.sum, horner(ev, alpha));
  std::vector<Point> pts;
  for (auto& c : eqc) pts.push_back(c.point);
  if (!pts.empty()) combine_eqs(sc.weights, pts, alpha);
  std::vector<F> vars;
  for (auto& c : powc) vars.push_back(c.var);
  if (!vars.empty()) combine_pows(sc.weights, vars, alpha, eqc.size());
}

static Point sc_fold(Writer& w, Prover& sc, int d, const std::vector<EqClaimE>& eqc,
                     const std::vector<PowClaimE>& powc) {  // prover.rs:377-397
  sc.alpha = w.draw_e();
  combine_claims(sc, sc.alpha, eqc, powc);
  Point rs;
  for (int i = 0; i < d; ++i) rs.push_back(round_step(w, sc));
  sc.rs.insert(sc.rs.end(), rs.begin(), rs.end());
  return rs;
}

static Point sc_fold_with_code(Writer& w, Prover& sc, int d, const std::vector<EqClaimE>& eqc,
                               const std::vector<size_t>& indexes, const Tree& data) {
  // prover.rs:399-561
  sc.alpha = w.draw_e();
  int k_data = 0;
  while (((size_t)1 << k_data) < data.height) ++k_data;
  combine_claims(sc, sc.alpha, eqc, {});
  size_t n = indexes.size();
  std::vector<E> scales(n);
  {
    E a = epow(sc.alpha, eqc.size());
    for (size_t j = 0; j < n; ++j) {
      scales[j] = a;
      a = emul(a, sc.alpha);
    }
  }
  const size_t width = data.row_base;
  std::vector<E> eq = eq_table(sc.rs, eone());
  auto row_dot = [&](size_t index) {
    E acc = ezero();
    for (size_t i = 0; i < width; ++i) acc = eadd(acc, emul_base(eq[i], data.data[index * width + i]));
    return acc;
  };
  for (size_t j = 0; j < n; ++j) sc.sum = eadd(sc.sum, emul(scales[j], row_dot(indexes[j])));

  F omega = two_adic_generator(k_data);
  std::vector<F> ws(n), ts(n);
  size_t mid0 = ((size_t)1 << k_data) / 2;
  for (size_t j = 0; j < n; ++j) {
    ws[j] = fpow(omega, indexes[j]);
    ts[j] = finv(fpow(omega, indexes[j] & (mid0 - 1)));
  }
  std::map<size_t, E> evals;
  for (int round = 0; round < d; ++round) {
    size_t stride = (size_t)1 << (k_data - round - 1);
    size_t count = (size_t)1 << (round + 1);
    for (size_t index : indexes) {
      size_t base = index & (stride - 1);
      for (size_t t = 0; t < count; ++t) evals[base + t * stride] = ezero();
    }
  }
  for (auto& kv : evals) kv.second = row_dot(kv.first);

  const F two_inv = finv(F::from_u32(2));
  Point rs;
  // g: prover.rs:437-461
  std::function<E(int, size_t)> g = [&](int round, size_t index) -> E {
    if (round == 0) return evals.at(index);
    size_t l = (size_t)1 << (k_data - (round - 1));
    size_t mid = l / 2;
    size_t i_lo = index & (mid - 1), i_hi = i_lo + mid;
    E g_lo = g(round - 1, i_lo), g_hi = g(round - 1, i_hi);
    F om = two_adic_generator(k_data - (round - 1));
    E even = emul_base(eadd(g_lo, g_hi), two_inv);
    E odd = emul_base(emul_base(esub(g_lo, g_hi), two_inv), finv(fpow(om, i_lo)));
    return eadd(even, emul(rs[round - 1], esub(odd, even)));
  };
  for (int round = 0; round < d; ++round) {
    E c0, c2;
    coeffs(sc.poly, sc.weights, c0, c2);
    size_t l = (size_t)1 << (k_data - round), mid = l / 2;
    for (size_t j = 0; j < n; ++j) {
      size_t i_lo = indexes[j] & (mid - 1), i_hi = i_lo + mid;
      E g_lo = g(round, i_lo), g_hi = g(round, i_hi);
      E even = emul_base(eadd(g_lo, g_hi), two_inv);
      E odd = emul_base(emul_base(esub(g_lo, g_hi), two_inv), ts[j]);
      c0 = eadd(c0, emul(scales[j], even));
      E t = esub(edbl(odd), even);
      F wt = fsub(fdbl(ws[j]), fone());
      c2 = eadd(c2, emul_base(emul(scales[j], t), wt));
    }
    w.write_e(c0);
    w.write_e(c2);
    E r = w.draw_e();
    sc.sum = extrapolate_012(c0, esub(sc.sum, c0), c2, r);
    rs.push_back(r);
    fix_lo_var(sc.poly, r);
    fix_lo_var(sc.weights, r);
    for (size_t j = 0; j < n; ++j) {
      // s *= 1 + r * (w - 1)
      E upd = eadd(eone(), emul(r, esub(efrom(ws[j]), eone())));
      scales[j] = emul(scales[j], upd);
      ws[j] = fsq(ws[j]);
      ts[j] = fsq(ts[j]);
      if ((indexes[j] >> (k_data - round - 2)) & 1) ts[j] = fneg(ts[j]);
    }
  }
  combine_pows_scaled(sc.weights, ws, scales);
  sc.rs.insert(sc.rs.end(), rs.begin(), rs.end());
  return rs;
}

template <class T>
static std::vector<size_t> stir_indexes(T& tx, int bits, int n) {  // whir.rs:63-77
  std::vector<size_t> v;
  for (int i = 0; i < n; ++i) v.push_back((size_t)tx.draw_bits(bits));
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return v;
}

static Tree commit_rows(const std::vector<F>& flat, size_t n_rows_in, size_t row_base, size_t height) {
  // whir.rs:19-61: zero-pad to `height` rows, column DFT, Merkle commit.
  std::vector<F> mat(height * row_base, F{0});
  memcpy(mat.data(), flat.data(), n_rows_in * row_base * sizeof(F));
  dft_batch(mat.data(), height, row_base);
  return merkle_commit(std::move(mat), height, row_base);
}

std::vector<F> rand_poly(uint64_t seed, int k) {
  StdRng rng = StdRng::seed_from_u64(seed);
  std::vector<F> v((size_t)1 << k);
  for (auto& x : v) x = rng.next_f();
  return v;
}

// ======================= multi-polynomial stacking =======================
bool selector_less(int ka, size_t ia, int kb, size_t ib) {
  // derived Ord of Selector(Vec<bool>) (sumcheck/mod.rs:30-37): lexicographic over
  // the bits (index >> i) & 1, i = 0, 1, ..., with false < true; on a common
  // prefix the shorter selector sorts first.
  const int n = std::min(ka, kb);
  for (int i = 0; i < n; ++i) {
    const int a = (int)((ia >> i) & 1), b = (int)((ib >> i) & 1);
    if (a != b) return a < b;
  }
  return ka < kb;
}

StackLayout stack_layout(const std::vector<int>& k_polys, int k_svo) {
  // ProverStack::layout (prover.rs:92-124) = VerifierLayout::new (verifier.rs:152-174):
  // stable sort by k, then walk from the largest (for equal k, the last
  // registered first) assigning consecutive offsets.
  const size_t n = k_polys.size();
  if (n == 0) throw std::runtime_error("empty stack");
  std::vector<int> ids(n);
  for (size_t i = 0; i < n; ++i) ids[i] = (int)i;
  std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) { return k_polys[a] < k_polys[b]; });
  size_t total = 0;
  for (int kp : k_polys) total += (size_t)1 << kp;
  StackLayout L;
  while (((size_t)1 << L.k) < total) ++L.k;  // log2_ceil_usize
  if (L.k < k_svo) throw std::runtime_error("stack smaller than the SVO rounds");
  L.offset.resize(n);
  L.sel_k.resize(n);
  L.sel_index.resize(n);
  size_t off = 0;
  for (size_t j = n; j-- > 0;) {
    const int id = ids[j], kp = k_polys[id];
    L.sel_k[id] = L.k - kp;
    L.sel_index[id] = off >> kp;  // off is a multiple of 2^kp (larger blocks come first)
    L.offset[id] = off;
    off += (size_t)1 << kp;
  }
  L.order = ids;
  std::sort(L.order.begin(), L.order.end(), [&](int a, int b) {
    return selector_less(L.sel_k[a], L.sel_index[a], L.sel_k[b], L.sel_index[b]);
  });
  return L;
}

std::vector<F> stack_polys(const StackLayout& L, const std::vector<std::vector<F>>& polys) {
  // prover.rs:114-118: zero vector of 2^k, each poly copied at selector.index << k_poly
  std::vector<F> out((size_t)1 << L.k, F{0});
  for (size_t id = 0; id < polys.size(); ++id)
    memcpy(out.data() + L.offset[id], polys[id].data(), polys[id].size() * sizeof(F));
  return out;
}

std::vector<size_t> stack_claim_ranks(const StackLayout& L, const std::vector<StackClaim>& claims) {
  // claims_in_layout_order (prover.rs:155-166): layout (Selector) order, then each
  // poly's claims in the order they were made (poly_to_claim_map).
  std::vector<size_t> rank(claims.size());
  size_t r = 0;
  for (int id : L.order)
    for (size_t c = 0; c < claims.size(); ++c)
      if (claims[c].poly == id) rank[c] = r++;
  return rank;
}

static Point hypercube(size_t value, int k) {  // point.rs:85-96
  Point p(k);
  for (int i = 0; i < k; ++i) p[i] = ((value >> i) & 1) ? eone() : ezero();
  return p;
}

static int log2_exact(size_t n) {
  int k = 0;
  while (((size_t)1 << k) < n) ++k;
  if (((size_t)1 << k) != n) throw std::runtime_error("poly size is not a power of two");
  return k;
}

static void check_claims(const std::vector<int>& kp, const std::vector<StackClaim>& claims, int f) {
  // ProverStack::eval asserts point.k() == k_poly (prover.rs:73); SvoPoint::new
  // needs k_poly >= k_svo.
  for (size_t c = 0; c < claims.size(); ++c) {
    const int id = claims[c].poly;
    if (id < 0 || id >= (int)kp.size() || kp[id] < f) throw std::runtime_error("bad claim poly");
    if (!claims[c].new_point && (c == 0 || kp[claims[c - 1].poly] != kp[id]))
      throw std::runtime_error("shared point of a different size");
  }
}

// write_virtual_eval (prover.rs:231-277) for point `pt` over the stacked polys:
// each poly's claim at pt[0..k_poly) weighted by eq(selector, pt[k_poly..k)).
static E virtual_claim(const StackLayout& L, const std::vector<std::vector<F>>& polys, int f, const Point& pt,
                       const std::vector<SvoCoeffs>& svo, Accs& accs) {
  std::vector<SvoClaim> local;
  std::vector<E> wts;
  local.reserve(L.order.size());
  E eval = ezero();
  for (int id : L.order) {
    const int kp = (int)(L.k - L.sel_k[id]);
    Point lp(pt.begin(), pt.begin() + kp), rest(pt.begin() + kp, pt.end());
    const E wgt = eval_eq_xy(hypercube(L.sel_index[id], L.sel_k[id]), rest);
    local.push_back(svo_eval(f, lp, polys[id]));
    eval = eadd(eval, emul(wgt, local.back().eval));
    wts.push_back(wgt);
  }
  std::vector<const SvoClaim*> ptrs;
  for (auto& c : local) ptrs.push_back(&c);
  accs = svo_accumulators(svo, ptrs, wts);
  return eval;
}

int stack_virtual_check(const std::vector<std::vector<F>>& polys, int k_svo, E x) {
  std::vector<int> kp;
  for (auto& p : polys) kp.push_back(log2_exact(p.size()));
  const StackLayout L = stack_layout(kp, k_svo);
  const Point pt = expand(L.k, x);
  auto svo = svo_new(k_svo);
  Accs accs;
  const E eval = virtual_claim(L, polys, k_svo, pt, svo, accs);
  const SvoClaim whole = svo_eval(k_svo, pt, stack_polys(L, polys));
  const Accs whole_accs = svo_accumulators(svo, {&whole}, {eone()});
  return eval == whole.eval && eval == eval_base(stack_polys(L, polys), pt) && accs == whole_accs;
}

ProveResult whir_prove(const Params& p, const std::vector<F>& poly, int n_points) {
  // one registered poly, each claim at a fresh point (test.rs:218-227)
  return whir_prove_stacked(p, {poly}, std::vector<StackClaim>((size_t)n_points, StackClaim{0, 1}));
}

ProveResult whir_prove_stacked(const Params& p, const std::vector<std::vector<F>>& polys,
                               const std::vector<StackClaim>& claims, const Session* session, Challenger* after) {
  const Hashers& H = Hashers::get();
  Writer w(&H.p16);
  if (session)
    for (F x : session->prefix) w.write_f(x);
  const int f = p.folding;
  std::vector<int> kp;
  for (auto& q : polys) kp.push_back(log2_exact(q.size()));
  const StackLayout L = stack_layout(kp, f);
  if (L.k != p.k) throw std::runtime_error("stacked k differs from the WHIR k");  // whir.rs:185
  check_claims(kp, claims, f);
  const int k = p.k;
  const std::vector<F> poly = stack_polys(L, polys);

  // ---- commit the stacked vector (whir.rs:127-138, 19-36) ----
  Tree base = commit_rows(poly, (size_t)1 << (k - f), (size_t)1 << f, (size_t)1 << (k + p.rate - f));
  w.write_digest(base.root().d);

  // ---- evaluation claims (ProverStack::eval, prover.rs:64-90; draws as in
  // test.rs:223-227 and sumcheck/test.rs:113-121) ----
  std::vector<SvoClaim> cl;
  Point pt;
  size_t gi = 0;
  for (const StackClaim& c : claims) {
    if (c.new_point == 1) pt = expand(kp[c.poly], w.draw_e());
    if (c.new_point == 2) {
      if (!session || gi >= session->given.size() || (int)session->given[gi].size() != kp[c.poly])
        throw std::runtime_error("missing or mis-sized given point");
      pt = session->given[gi++];
    }
    cl.push_back(svo_eval(f, pt, polys[c.poly]));
    w.write_e(cl.back().eval);
  }

  // ---- open: OOD virtual claims (whir.rs:187-192; prover.rs:231-277) ----
  auto svo = svo_new(f);
  std::vector<SvoClaim> vclaims;
  std::vector<Accs> vaccs;
  int n_ood0 = n_ood_queries(p.soundness, p.security_level, k, p.rate, EXT_BITS);
  for (int i = 0; i < n_ood0; ++i) {
    const Point vp = expand(k, w.draw_e());
    SvoClaim vc;
    vc.svo.assign(vp.begin(), vp.begin() + f);
    vc.rest.assign(vp.begin() + f, vp.end());
    vaccs.emplace_back();
    vc.eval = virtual_claim(L, polys, f, vp, svo, vaccs.back());
    vclaims.push_back(std::move(vc));
    w.write_e(vclaims.back().eval);
  }

  // ---- new_prover: SVO rounds (prover.rs:279-353) ----
  Prover sc;
  E alpha = w.draw_e();
  const std::vector<size_t> rank = stack_claim_ranks(L, claims);
  std::vector<E> alphas(cl.size());
  for (size_t c = 0; c < cl.size(); ++c) alphas[c] = epow(alpha, rank[c]);
  {
    std::vector<E> ev(cl.size());  // claims in layout order, then the virtual ones (prover.rs:168-174)
    for (size_t c = 0; c < cl.size(); ++c) ev[rank[c]] = cl[c].eval;
    for (auto& c : vclaims) ev.push_back(c.eval);
    sc.sum = horner(ev, alpha);
  }
  std::vector<Accs> caccs;
  for (size_t i = 0; i < cl.size(); ++i) caccs.push_back(svo_accumulators(svo, {&cl[i]}, {alphas[i]}));
  Point rs;
  for (int ri = 0; ri < f; ++ri) {
    std::vector<E> wts = lagrange_weights_012_multi(rs);
    std::vector<E> acc0(wts.size(), ezero()), acc2(wts.size(), ezero());
    for (auto& a : caccs)
      for (size_t u = 0; u < wts.size(); ++u) {
        acc0[u] = eadd(acc0[u], a[ri][0][u]);
        acc2[u] = eadd(acc2[u], a[ri][1][u]);
      }
    E av = epow(alpha, cl.size());
    for (auto& a : vaccs) {
      for (size_t u = 0; u < wts.size(); ++u) {
        acc0[u] = eadd(acc0[u], emul(av, a[ri][0][u]));
        acc2[u] = eadd(acc2[u], emul(av, a[ri][1][u]));
      }
      av = emul(av, alpha);
    }
    E c0 = ezero(), c2 = ezero();
    for (size_t u = 0; u < wts.size(); ++u) {
      c0 = eadd(c0, emul(acc0[u], wts[u]));
      c2 = eadd(c2, emul(acc2[u], wts[u]));
    }
    w.write_e(c0);
    w.write_e(c2);
    E r = w.draw_e();
    sc.sum = extrapolate_012(c0, esub(sc.sum, c0), c2, r);
    rs.push_back(r);
  }
  // ---- expression (prover.rs:184-229) ----
  // compress_stacked: each poly compressed into its block at selector.index << (k_poly - f)
  sc.poly.assign((size_t)1 << (k - f), ezero());
  for (int id : L.order) {
    const std::vector<E> blk = compress_lo(polys[id], rs);
    std::copy(blk.begin(), blk.end(), sc.poly.begin() + (L.sel_index[id] << (kp[id] - f)));
  }
  sc.weights.assign(sc.poly.size(), ezero());
  {
    auto add_eq = [&](const SvoClaim& c, E scale_a, size_t off) {
      const E scale = emul(scale_a, eval_eq_xy(c.svo, rs));
      std::vector<E> eq = eq_table(c.rest, scale);
#pragma omp parallel for
      for (long long x = 0; x < (long long)eq.size(); ++x) sc.weights[off + x] = eadd(sc.weights[off + x], eq[x]);
    };
    // claims in layout order, each into its poly's block (prover.rs:190-199)
    E ai = eone();
    for (int id : L.order)
      for (size_t c = 0; c < cl.size(); ++c)
        if (claims[c].poly == id) {
          add_eq(cl[c], ai, L.sel_index[id] << (kp[id] - f));
          ai = emul(ai, alpha);
        }
    for (auto& c : vclaims) {  // virtual claims over the whole vector (prover.rs:201-207)
      add_eq(c, ai, 0);
      ai = emul(ai, alpha);
    }
  }
  sc.rs = rs;
  sc.alpha = alpha;

  // ---- WHIR rounds (whir.rs:197-312) ----
  int n_rounds, final_rounds;
  number_of_rounds(f, k, n_rounds, final_rounds);
  Tree round_tree;
  bool have_round_tree = false;
  Point round_point;
  for (int round = 0; round < n_rounds; ++round) {
    int kfd, this_rate, prev_rate;
    round_params(p, round, kfd, this_rate, prev_rate);
    int kc = (int)std::log2((double)sc.poly.size());
    // commit_ext: rows of 2^f ext, flattened coefficient-major per element
    std::vector<F> flat(sc.poly.size() * 4);
    for (size_t i = 0; i < sc.poly.size(); ++i)
      for (int c = 0; c < 4; ++c) flat[i * 4 + c] = sc.poly[i].c[c];
    Tree nt = commit_rows(flat, sc.poly.size() >> f, (size_t)4 << f, (size_t)1 << (kc + this_rate - f));
    w.write_digest(nt.root().d);
    // OOD claims in round
    std::vector<EqClaimE> ood;
    int n_ood = n_ood_queries(p.soundness, p.security_level, kc, this_rate, EXT_BITS);
    for (int i = 0; i < n_ood; ++i) {
      Point pt = expand(kc, w.draw_e());
      E ev = eval_ext(sc.poly, pt);
      w.write_e(ev);
      ood.push_back({pt, ev});
    }
    int n_stir = n_stir_queries(p.soundness, p.security_level, prev_rate);
    std::vector<size_t> idx = stir_indexes(w, kfd, n_stir);
    F omega = two_adic_generator(kfd);
    if (round == 0) {
      for (size_t index : idx) merkle_query(w, base, index);
      round_point = sc_fold_with_code(w, sc, f, ood, idx, base);
    } else {
      std::vector<PowClaimE> stir;
      for (size_t index : idx) {
        merkle_query(w, round_tree, index);
        std::vector<E> local((size_t)1 << f);
        for (size_t e = 0; e < local.size(); ++e)
          for (int c = 0; c < 4; ++c) local[e].c[c] = round_tree.data[index * round_tree.row_base + e * 4 + c];
        stir.push_back({fpow(omega, index), eval_ext(local, round_point)});
      }
      round_point = sc_fold(w, sc, f, ood, stir);
    }
    round_tree = std::move(nt);
    have_round_tree = true;
  }
  // ---- final (whir.rs:314-355) ----
  for (const E& e : sc.poly) w.write_e(e);
  {
    int kfd, this_rate, prev_rate;
    round_params(p, n_rounds, kfd, this_rate, prev_rate);
    int n_stir = n_stir_queries(p.soundness, p.security_level, prev_rate);
    std::vector<size_t> idx = stir_indexes(w, kfd, n_stir);
    for (size_t index : idx) merkle_query(w, have_round_tree ? round_tree : base, index);
  }
  ProveResult res;
  if (after) *after = w.ch;
  res.checkpoint = w.draw_f();  // test.rs:233
  res.root = base.root();
  res.proof = std::move(w.bytes);
  return res;
}

// ======================= verifier =======================
struct FoldingRounds {  // sumcheck/verifier.rs:22-68
  std::vector<Point> eq_points;
  std::vector<std::vector<F>> pow_points;
  Point rs;
  E alpha;
  E eval(const std::vector<E>& poly) const {
    std::vector<E> wts;
    size_t kr = rs.size();
    for (const Point& zs : eq_points) {
      Point zf(zs.begin(), zs.begin() + kr), zp(zs.begin() + kr, zs.end());
      wts.push_back(emul(eval_eq_xy(zf, rs), eval_ext(poly, zp)));
    }
    for (const auto& zs : pow_points) {
      std::vector<F> zf(zs.begin(), zs.begin() + kr);
      wts.push_back(emul(eval_pow_xy(zf, rs), eval_univariate_e(poly, efrom(zs[kr]))));
    }
    return horner(wts, alpha);
  }
};

#define TRY(x)                 \
  do {                         \
    int _s = (x);              \
    if (_s != OK) return _s;   \
  } while (0)

static int reduce(Reader& r, E& sum, E& out) {  // verifier.rs:8-20
  E v0, v2;
  TRY(r.read_e(v0));
  TRY(r.read_e(v2));
  E ch = r.draw_e();
  sum = extrapolate_012(v0, esub(sum, v0), v2, ch);
  out = ch;
  return OK;
}

int whir_verify(const Params& p, const uint8_t* proof, size_t len, int n_points, F& checkpoint) {
  // the test's claim reading (test.rs:237-250): one poly, a fresh point per claim
  return whir_verify_stacked(p, proof, len, {p.k}, std::vector<StackClaim>((size_t)n_points, StackClaim{0, 1}),
                             checkpoint);
}

int whir_verify_stacked(const Params& p, const uint8_t* proof, size_t len, const std::vector<int>& k_polys,
                        const std::vector<StackClaim>& claims, F& checkpoint, const Session* session) {
  // whir.rs:361-522 with VerifierStack::read_eval / VerifierLayout (verifier.rs:100-212).
  const Hashers& H = Hashers::get();
  Reader r(&H.p16, proof, len);
  if (session)
    for (F x : session->prefix) {
      F got;
      TRY(r.read_f(got));
      if (got.v != x.v) return ERR_VERIFY;
    }
  const int k = p.k, f = p.folding;
  const StackLayout L = stack_layout(k_polys, f);
  if (L.k != k) return ERR_VERIFY;
  check_claims(k_polys, claims, f);
  Digest comm;
  TRY(r.read_digest(comm.d));
  const std::vector<size_t> rank = stack_claim_ranks(L, claims);
  std::vector<Point> pts(claims.size());
  std::vector<E> evals(claims.size());
  Point pt;
  size_t gi = 0;
  for (size_t c = 0; c < claims.size(); ++c) {
    const int id = claims[c].poly;
    if (claims[c].new_point == 1) pt = expand(k_polys[id], r.draw_e());
    if (claims[c].new_point == 2) {
      if (!session || gi >= session->given.size() || (int)session->given[gi].size() != k_polys[id])
        return ERR_VERIFY;
      pt = session->given[gi++];
    }
    E ev;
    TRY(r.read_e(ev));
    // lift: point ++ selector point (mod.rs:55-57), stored in layout order (verifier.rs:204-212)
    Point lifted = pt;
    const Point sel = hypercube(L.sel_index[id], L.sel_k[id]);
    lifted.insert(lifted.end(), sel.begin(), sel.end());
    pts[rank[c]] = std::move(lifted);
    evals[rank[c]] = ev;
  }
  int n_ood0 = n_ood_queries(p.soundness, p.security_level, k, p.rate, EXT_BITS);
  for (int i = 0; i < n_ood0; ++i) {  // verifier.rs:176-189
    Point vp = expand(k, r.draw_e());
    E ev;
    TRY(r.read_e(ev));
    pts.push_back(vp);
    evals.push_back(ev);
  }
  // SumcheckVerifier::new (verifier.rs:227-255)
  E alpha = r.draw_e();
  E sum = horner(evals, alpha);
  Point round_rs;
  for (int i = 0; i < f; ++i) {
    E ch;
    TRY(reduce(r, sum, ch));
    round_rs.push_back(ch);
  }
  std::vector<FoldingRounds> multi{{pts, {}, round_rs, alpha}};
  int kcur = k - f;
  Point round_point = round_rs;
  Digest round_comm;
  bool have_round = false;
  int n_rounds, final_rounds;
  number_of_rounds(f, k, n_rounds, final_rounds);
  for (int round = 0; round < n_rounds; ++round) {
    int kfd, this_rate, prev_rate;
    round_params(p, round, kfd, this_rate, prev_rate);
    Digest nc;
    TRY(r.read_digest(nc.d));
    int n_ood = n_ood_queries(p.soundness, p.security_level, kcur, this_rate, EXT_BITS);
    std::vector<Point> eqp;
    std::vector<E> ev;
    for (int i = 0; i < n_ood; ++i) {
      E z = r.draw_e();
      E e;
      TRY(r.read_e(e));
      eqp.push_back(expand(kcur, z));
      ev.push_back(e);
    }
    int n_stir = n_stir_queries(p.soundness, p.security_level, prev_rate);
    std::vector<size_t> idx = stir_indexes(r, kfd, n_stir);
    F omega = two_adic_generator(kfd);
    std::vector<std::vector<F>> powp;
    for (size_t index : idx) {
      std::vector<F> row;
      E e;
      if (round == 0) {
        TRY(merkle_verify(r, comm, index, (size_t)1 << f, kfd, row));
        e = eval_base(row, round_point);
      } else {
        TRY(merkle_verify(r, round_comm, index, (size_t)4 << f, kfd, row));
        std::vector<E> local((size_t)1 << f);
        for (size_t i = 0; i < local.size(); ++i)
          for (int c = 0; c < 4; ++c) local[i].c[c] = row[i * 4 + c];
        e = eval_ext(local, round_point);
      }
      F var = fpow(omega, index);
      std::vector<F> pp(kcur);
      for (int i = 0; i < kcur; ++i) {
        pp[i] = var;
        var = fsq(var);
      }
      powp.push_back(pp);
      ev.push_back(e);
    }
    round_comm = nc;
    have_round = true;
    // SumcheckVerifier::fold (verifier.rs:262-306)
    E a = r.draw_e();
    sum = eadd(sum, horner(ev, a));
    Point rr;
    for (int i = 0; i < f; ++i) {
      E ch;
      TRY(reduce(r, sum, ch));
      rr.push_back(ch);
    }
    for (auto& m : multi) m.rs.insert(m.rs.end(), rr.begin(), rr.end());
    multi.push_back({eqp, powp, rr, a});
    kcur -= f;
    round_point = rr;
  }
  if (kcur != final_rounds) return ERR_VERIFY;
  std::vector<E> final_poly((size_t)1 << kcur);
  for (auto& e : final_poly) TRY(r.read_e(e));
  {
    int kfd, this_rate, prev_rate;
    round_params(p, n_rounds, kfd, this_rate, prev_rate);
    int n_stir = n_stir_queries(p.soundness, p.security_level, prev_rate);
    std::vector<size_t> idx = stir_indexes(r, kfd, n_stir);
    F omega = two_adic_generator(kfd);
    for (size_t index : idx) {
      std::vector<F> row;
      E e;
      if (have_round) {
        TRY(merkle_verify(r, round_comm, index, (size_t)4 << f, kfd, row));
        std::vector<E> local((size_t)1 << f);
        for (size_t i = 0; i < local.size(); ++i)
          for (int c = 0; c < 4; ++c) local[i].c[c] = row[i * 4 + c];
        e = eval_ext(local, round_point);
      } else {
        TRY(merkle_verify(r, comm, index, (size_t)1 << f, kfd, row));
        e = eval_base(row, round_point);
      }
      if (e != eval_univariate_e(final_poly, efrom(fpow(omega, index)))) return ERR_VERIFY;
    }
  }
  // finalize (verifier.rs:308-323)
  E total = ezero();
  for (auto& m : multi) total = eadd(total, m.eval(final_poly));
  if (total != sum) return ERR_VERIFY;
  checkpoint = r.draw_f();
  return OK;
}

}  // namespace oracle
