// Device-resident Fiat–Shamir transcript: PoseidonWriter over
// DuplexChallenger<KoalaBear, Poseidon2<16>, 16, 8>
// (crates/transcript/src/poseidon.rs:13-87; crates/whir/src/test.rs:190).
//   write(x)      = append the Montgomery words to the proof log, then observe;
//   write_hint(x) = append only;
//   observe(v)    : clear output; push v; duplex when 8 inputs are buffered;
//   sample()      : duplex if input pending or output empty; pop from the end;
//   duplex        : state[0..len) = input; P16(state); output = state[0..8);
//   sample_bits(b): canonical(sample()) & (2^b - 1).
// One wavefront runs the transcript: every lane holds the (uniform) buffers and
// lane i owns state word i & 15, so P16 runs 16-wide with cross-lane shuffles.
#pragma once
#include "field.h"
#include "poseidon2.h"

namespace kb {

struct DevTx {
  uint32_t state[16];
  uint32_t in_buf[8];
  uint32_t out_buf[8];
  uint32_t in_len, out_len;
  uint32_t log_words, log_cap;
  uint32_t overflow;
  uint32_t pad[3];
};

__constant__ uint32_t c_diag16[16];  // Montgomery form of V16 (cooperative P16)

__device__ __forceinline__ uint32_t shfl16(uint32_t v, int src) { return __shfl((int)v, src, 16); }
__device__ __forceinline__ uint32_t shfl16_xor(uint32_t v, int m) { return __shfl_xor((int)v, m, 16); }

// Cooperative Poseidon2-16: lane i (mod 16) holds state word i.
__device__ __forceinline__ void mds_light16_coop(uint32_t& x, int i) {
  const int a = i & 3;
  const uint32_t xn = shfl16(x, (i & 12) | ((a + 1) & 3));
  uint32_t t = add(x, shfl16_xor(x, 1));
  const uint32_t s4 = add(t, shfl16_xor(t, 2));
  const uint32_t y = add(add(s4, x), dbl(xn));  // s4 + x_a + 2 x_{a+1}
  uint32_t u = add(y, shfl16_xor(y, 4));
  const uint32_t cs = add(u, shfl16_xor(u, 8));
  x = add(y, cs);
}

__device__ __forceinline__ void permute16_coop(uint32_t& s) {
  const int i = threadIdx.x & 15;
  mds_light16_coop(s, i);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    s = cube(add(s, c_p2.ext16_init[r][i]));
    mds_light16_coop(s, i);
  }
  const uint32_t d = c_diag16[i];
#pragma unroll
  for (int r = 0; r < 20; ++r) {
    const uint32_t t = cube(add(s, c_p2.int16[r]));
    s = i == 0 ? t : s;
    uint32_t sum = add(s, shfl16_xor(s, 1));
    sum = add(sum, shfl16_xor(sum, 2));
    sum = add(sum, shfl16_xor(sum, 4));
    sum = add(sum, shfl16_xor(sum, 8));
    s = add(mul(s, d), sum);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    s = cube(add(s, c_p2.ext16_term[r][i]));
    mds_light16_coop(s, i);
  }
}

// Select element `idx` (uniform, < 8) of a register array without dynamic indexing.
__device__ __forceinline__ uint32_t sel8(const uint32_t (&a)[8], uint32_t idx) {
  uint32_t v = a[0];
#pragma unroll
  for (int j = 1; j < 8; ++j) v = idx == (uint32_t)j ? a[j] : v;
  return v;
}

// Must be used by all 64 lanes of exactly one wavefront, in uniform control flow.
struct WaveTx {
  uint32_t s;
  uint32_t in_buf[8], out_buf[8];
  uint32_t in_len, out_len;
  uint32_t log_words, log_cap, overflow;
  uint32_t* log;

  __device__ void load(const DevTx* t, uint32_t* log_ptr) {
    const int lane = threadIdx.x & 63;
    s = t->state[lane & 15];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      in_buf[j] = t->in_buf[j];
      out_buf[j] = t->out_buf[j];
    }
    in_len = t->in_len;
    out_len = t->out_len;
    log_words = t->log_words;
    log_cap = t->log_cap;
    overflow = t->overflow;
    log = log_ptr;
  }
  __device__ void store(DevTx* t) {
    const int lane = threadIdx.x & 63;
    if (lane < 16) t->state[lane] = s;
    if (lane == 0) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        t->in_buf[j] = in_buf[j];
        t->out_buf[j] = out_buf[j];
      }
      t->in_len = in_len;
      t->out_len = out_len;
      t->log_words = log_words;
      t->log_cap = log_cap;
      t->overflow = overflow;
    }
  }
  __device__ void duplex() {
    const int i = threadIdx.x & 15;
    const uint32_t inv = sel8(in_buf, (uint32_t)(i & 7));
    if ((uint32_t)i < in_len) s = inv;
    in_len = 0;
    permute16_coop(s);
#pragma unroll
    for (int j = 0; j < 8; ++j) out_buf[j] = shfl16(s, j);
    out_len = 8;
  }
  __device__ void observe(uint32_t v) {
    out_len = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) in_buf[j] = in_len == (uint32_t)j ? v : in_buf[j];
    in_len += 1;
    if (in_len == 8) duplex();
  }
  __device__ uint32_t sample() {
    if (in_len != 0 || out_len == 0) duplex();
    out_len -= 1;
    return sel8(out_buf, out_len);
  }
  __device__ Ext sample_ext() {
    Ext e;
    e.c[0] = sample();
    e.c[1] = sample();
    e.c[2] = sample();
    e.c[3] = sample();
    return e;
  }
  __device__ uint32_t sample_bits(uint32_t bits) { return canonical(sample()) & ((1u << bits) - 1u); }
  __device__ void hint_word(uint32_t v) {
    if ((threadIdx.x & 63) == 0) {
      if (log_words < log_cap) log[log_words] = v;
    }
    if (log_words >= log_cap) overflow = 1;
    log_words += 1;
  }
  __device__ void write_f(uint32_t v) {
    hint_word(v);
    observe(v);
  }
  __device__ void write_e(const Ext& e) {
#pragma unroll
    for (int c = 0; c < 4; ++c) hint_word(e.c[c]);
#pragma unroll
    for (int c = 0; c < 4; ++c) observe(e.c[c]);
  }
  __device__ void write_digest(const uint32_t* d) {
    for (int j = 0; j < 8; ++j) hint_word(d[j]);
    for (int j = 0; j < 8; ++j) observe(d[j]);
  }
};

}  // namespace kb
