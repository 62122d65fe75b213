// ORACLE — test infrastructure only (see field.h).
// Field extension arithmetic, ChaCha/StdRng, Poseidon2.
#include <cstring>
#include <mutex>
#include "field.h"
#include "rng.h"
#include "poseidon2.h"

namespace oracle {

F two_adic_generator(unsigned bits) {
  // Plonky3 koala-bear: GENERATOR = 3, two-adicity 24.
  return fpow(F::from_u32(GENERATOR_CANON), (uint64_t)(P - 1) >> bits);
}

E emul(const E& a, const E& b) {
  // (sum a_i x^i)(sum b_j x^j) mod x^4 - W
  const F w = F::from_u32(W_CANON);
  F t[7];
  for (int i = 0; i < 7; ++i) t[i] = F{0};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) t[i + j] = fadd(t[i + j], fmul(a.c[i], b.c[j]));
  E r;
  r.c[0] = fadd(t[0], fmul(w, t[4]));
  r.c[1] = fadd(t[1], fmul(w, t[5]));
  r.c[2] = fadd(t[2], fmul(w, t[6]));
  r.c[3] = t[3];
  return r;
}

E epow(E a, uint64_t e) {
  E r = eone();
  while (e) {
    if (e & 1) r = emul(r, a);
    a = emul(a, a);
    e >>= 1;
  }
  return r;
}

// ---------------- ChaCha / StdRng ----------------

static inline uint32_t rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }

void chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream, int rounds,
                  uint32_t out[16]) {
  uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                    key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                    (uint32_t)counter, (uint32_t)(counter >> 32),
                    (uint32_t)stream, (uint32_t)(stream >> 32)};
  uint32_t x[16];
  memcpy(x, s, sizeof(x));
#define QR(a, b, c, d)              \
  x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16); \
  x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 12); \
  x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8);  \
  x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 7);
  for (int r = 0; r < rounds; r += 2) {
    QR(0, 4, 8, 12) QR(1, 5, 9, 13) QR(2, 6, 10, 14) QR(3, 7, 11, 15)
    QR(0, 5, 10, 15) QR(1, 6, 11, 12) QR(2, 7, 8, 13) QR(3, 4, 9, 14)
  }
#undef QR
  for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
}

StdRng StdRng::seed_from_u64(uint64_t state) {
  // rand_core SeedableRng::seed_from_u64: PCG32 fills the seed 4 bytes at a time.
  const uint64_t MUL = 6364136223846793005ull;
  const uint64_t INC = 11634580027462260723ull;
  StdRng r;
  for (int i = 0; i < 8; ++i) {
    state = state * MUL + INC;
    uint64_t s = state;
    uint32_t xorshifted = (uint32_t)(((s >> 18) ^ s) >> 27);
    uint32_t rot = (uint32_t)(s >> 59);
    uint32_t x = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
    r.key[i] = x;  // little-endian bytes -> u32 word
  }
  r.counter = 0;
  r.idx = 16;
  return r;
}

uint32_t StdRng::next_u32() {
  if (idx >= 16) {
    chacha_block(key, counter, 0, rounds, buf);
    counter += 1;
    idx = 0;
  }
  return buf[idx++];
}

F StdRng::next_f() {
  for (;;) {
    uint32_t v = next_u32() >> 1;
    if (v < P) return F::from_monty(v);
  }
}

E StdRng::next_e() {
  E e;
  for (int i = 0; i < 4; ++i) e.c[i] = next_f();
  return e;
}

// ---------------- Poseidon2 ----------------

// Canonical values of the internal diagonal V (Plonky3 koala-bear poseidon2.rs).
// Each entry is sign * num / 2^sh.
struct DiagEntry { int sign; uint32_t num; int sh; };
#if defined(ORACLE_BABYBEAR)
// Plonky3 baby-bear poseidon2.rs (recalled): parity unpinned.
static const DiagEntry V16[16] = {
    {-1, 2, 0}, {1, 1, 0}, {1, 2, 0}, {1, 1, 1}, {1, 3, 0}, {1, 4, 0}, {-1, 1, 1}, {-1, 3, 0},
    {-1, 4, 0}, {1, 1, 8}, {1, 1, 2}, {1, 1, 3}, {1, 1, 27}, {-1, 1, 8}, {-1, 1, 4}, {-1, 1, 27}};
static const DiagEntry V24[24] = {
    {-1, 2, 0}, {1, 1, 0}, {1, 2, 0}, {1, 1, 1}, {1, 3, 0}, {1, 4, 0}, {-1, 1, 1}, {-1, 3, 0},
    {-1, 4, 0}, {1, 1, 8}, {1, 1, 2}, {1, 1, 3}, {1, 1, 4}, {1, 1, 7}, {1, 1, 9}, {1, 1, 27},
    {-1, 1, 8}, {-1, 1, 2}, {-1, 1, 3}, {-1, 1, 4}, {-1, 1, 5}, {-1, 1, 6}, {-1, 1, 7}, {-1, 1, 27}};
constexpr int ROUNDS_P16 = 13, ROUNDS_P24 = 21;
#else
static const DiagEntry V16[16] = {
    {-1, 2, 0}, {1, 1, 0}, {1, 2, 0}, {1, 1, 1}, {1, 3, 0}, {1, 4, 0}, {-1, 1, 1}, {-1, 3, 0},
    {-1, 4, 0}, {1, 1, 8}, {1, 1, 3}, {1, 1, 24}, {-1, 1, 8}, {-1, 1, 3}, {-1, 1, 4}, {-1, 1, 24}};
static const DiagEntry V24[24] = {
    {-1, 2, 0}, {1, 1, 0}, {1, 2, 0}, {1, 1, 1}, {1, 3, 0}, {1, 4, 0}, {-1, 1, 1}, {-1, 3, 0},
    {-1, 4, 0}, {1, 1, 8}, {1, 1, 2}, {1, 1, 3}, {1, 1, 4}, {1, 1, 5}, {1, 1, 6}, {1, 1, 24},
    {-1, 1, 8}, {-1, 1, 3}, {-1, 1, 4}, {-1, 1, 5}, {-1, 1, 6}, {-1, 1, 7}, {-1, 1, 9}, {-1, 1, 24}};
constexpr int ROUNDS_P16 = 20, ROUNDS_P24 = 23;
#endif

void internal_diag_canonical(int width, std::vector<F>& out) {
  const DiagEntry* v = width == 16 ? V16 : V24;
  out.resize(width);
  F inv2 = finv(F::from_u32(2));
  for (int i = 0; i < width; ++i) {
    F x = fmul(F::from_u32(v[i].num), fpow(inv2, v[i].sh));
    out[i] = v[i].sign < 0 ? fneg(x) : x;
  }
}

Poseidon2 Poseidon2::make(int width, uint64_t seed) {
  Poseidon2 p;
  p.width = width;
  p.rounds_p = width == 16 ? ROUNDS_P16 : ROUNDS_P24;
  StdRng rng = StdRng::seed_from_u64(seed);
  p.ext_initial.resize(4 * width);
  p.ext_terminal.resize(4 * width);
  p.internal.resize(p.rounds_p);
  for (auto& x : p.ext_initial) x = rng.next_f();
  for (auto& x : p.ext_terminal) x = rng.next_f();
  for (auto& x : p.internal) x = rng.next_f();
  internal_diag_canonical(width, p.diag);
  return p;
}

void Poseidon2::mds_light(F* s) const {
  for (int c = 0; c < width; c += 4) {
    F x0 = s[c], x1 = s[c + 1], x2 = s[c + 2], x3 = s[c + 3];
    F sum = fadd(fadd(x0, x1), fadd(x2, x3));
    s[c + 0] = fadd(fadd(sum, x0), fdbl(x1));
    s[c + 1] = fadd(fadd(sum, x1), fdbl(x2));
    s[c + 2] = fadd(fadd(sum, x2), fdbl(x3));
    s[c + 3] = fadd(fadd(sum, x3), fdbl(x0));
  }
  F sums[4] = {F{0}, F{0}, F{0}, F{0}};
  for (int i = 0; i < width; ++i) sums[i % 4] = fadd(sums[i % 4], s[i]);
  for (int i = 0; i < width; ++i) s[i] = fadd(s[i], sums[i % 4]);
}

void Poseidon2::internal_layer(F* s) const {
  F sum = F{0};
  for (int i = 0; i < width; ++i) sum = fadd(sum, s[i]);
  for (int i = 0; i < width; ++i) s[i] = fadd(fmul(s[i], diag[i]), sum);
}

// The S-box x^d, d the smallest exponent with gcd(d, p - 1) = 1: 3 for
// KoalaBear, 7 for BabyBear.
#if defined(ORACLE_BABYBEAR)
static inline F sbox(F x) {
  const F x2 = fsq(x), x3 = fmul(x2, x);
  return fmul(x3, fsq(x2));
}
#else
static inline F sbox(F x) { return fmul(fsq(x), x); }
#endif

void Poseidon2::permute(F* s) const {
  mds_light(s);
  for (int r = 0; r < 4; ++r) {
    for (int i = 0; i < width; ++i) s[i] = sbox(fadd(s[i], ext_initial[r * width + i]));
    mds_light(s);
  }
  for (int r = 0; r < rounds_p; ++r) {
    s[0] = sbox(fadd(s[0], internal[r]));
    internal_layer(s);
  }
  for (int r = 0; r < 4; ++r) {
    for (int i = 0; i < width; ++i) s[i] = sbox(fadd(s[i], ext_terminal[r * width + i]));
    mds_light(s);
  }
}

const Hashers& Hashers::get() {
  static Hashers h = []() {
    Hashers x;
    x.p16 = Poseidon2::make(16, 1000);
    x.p24 = Poseidon2::make(24, 1000);
    return x;
  }();
  return h;
}

void sponge24(const Poseidon2& p24, const F* in, size_t n, F out[8]) {
  // p3-symmetric PaddingFreeSponge::hash_iter: overwrite the first RATE=16 lanes
  // with the next chunk and permute; a final partial chunk is permuted too; an
  // empty final chunk is not.
  F st[24];
  for (int i = 0; i < 24; ++i) st[i] = F{0};
  size_t pos = 0;
  for (;;) {
    size_t take = n - pos < 16 ? n - pos : 16;
    for (size_t i = 0; i < take; ++i) st[i] = in[pos + i];
    pos += take;
    if (take == 0) break;
    p24.permute(st);
    if (take < 16) break;
  }
  for (int i = 0; i < 8; ++i) out[i] = st[i];
}

void compress16(const Poseidon2& p16, const F l[8], const F r[8], F out[8]) {
  F st[16];
  for (int i = 0; i < 8; ++i) {
    st[i] = l[i];
    st[8 + i] = r[i];
  }
  p16.permute(st);
  for (int i = 0; i < 8; ++i) out[i] = st[i];
}

}  // namespace oracle
