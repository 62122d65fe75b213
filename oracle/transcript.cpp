// ORACLE — test infrastructure only (see field.h).
#include "transcript.h"

namespace oracle {

void Challenger::duplex() {
  // Plonky3 DuplexChallenger::duplexing (overwrite mode).
  for (size_t i = 0; i < in.size(); ++i) state[i] = in[i];
  in.clear();
  perm->permute(state);
  out.assign(state, state + 8);
}

void Challenger::observe(F v) {
  out.clear();
  in.push_back(v);
  if (in.size() == 8) duplex();
}

F Challenger::sample() {
  if (!in.empty() || out.empty()) duplex();
  F v = out.back();
  out.pop_back();
  return v;
}

E Challenger::sample_ext() {
  E e;
  for (int i = 0; i < 4; ++i) e.c[i] = sample();
  return e;
}

uint64_t Challenger::sample_bits(unsigned bits) {
  F v = sample();
  return (uint64_t)v.canonical() & ((1ull << bits) - 1);
}

static void push_u32(std::vector<uint8_t>& b, uint32_t v) {
  uint8_t t[4];
  put_u32(t, v);
  b.insert(b.end(), t, t + 4);
}

void Writer::hint_f(F x) { push_u32(bytes, x.v); }
void Writer::write_f(F x) {
  hint_f(x);
  ch.observe(x);
}
void Writer::hint_e(const E& x) {
  for (int i = 0; i < 4; ++i) push_u32(bytes, x.c[i].v);
}
void Writer::write_e(const E& x) {
  hint_e(x);
  for (int i = 0; i < 4; ++i) ch.observe(x.c[i]);
}
void Writer::hint_digest(const F d[8]) {
  for (int i = 0; i < 8; ++i) push_u32(bytes, d[i].v);
}
void Writer::write_digest(const F d[8]) {
  hint_digest(d);
  for (int i = 0; i < 8; ++i) ch.observe(d[i]);
}

int Reader::hint_f(F& x) {
  if (pos + 4 > len) return ERR_TRANSCRIPT;
  uint32_t v = get_u32(data + pos);
  if (v >= P) return ERR_TRANSCRIPT;  // serde rejects non-canonical Montgomery words
  pos += 4;
  x = F::from_monty(v);
  return OK;
}
int Reader::read_f(F& x) {
  int s = hint_f(x);
  if (s == OK) ch.observe(x);
  return s;
}
int Reader::hint_e(E& x) {
  for (int i = 0; i < 4; ++i) {
    int s = hint_f(x.c[i]);
    if (s != OK) return s;
  }
  return OK;
}
int Reader::read_e(E& x) {
  int s = hint_e(x);
  if (s != OK) return s;
  for (int i = 0; i < 4; ++i) ch.observe(x.c[i]);
  return OK;
}
int Reader::hint_digest(F d[8]) {
  for (int i = 0; i < 8; ++i) {
    int s = hint_f(d[i]);
    if (s != OK) return s;
  }
  return OK;
}
int Reader::read_digest(F d[8]) {
  int s = hint_digest(d);
  if (s != OK) return s;
  for (int i = 0; i < 8; ++i) ch.observe(d[i]);
  return OK;
}

}  // namespace oracle
