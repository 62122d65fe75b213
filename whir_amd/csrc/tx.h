// Device-resident Fiat–Shamir transcript: PoseidonWriter over
// DuplexChallenger<KoalaBear, Poseidon2<16>, 16, 8>
// (crates/transcript/src/poseidon.rs:13-87; crates/whir/src/test.rs:190).
//   write(x)      = append the Montgomery words to the proof log, then observe;
//   write_hint(x) = append only;
//   observe(v)    : clear output; push v; duplex when 8 inputs are buffered;
//   sample()      : duplex if input pending or output empty; pop from the end;
//   duplex        : state[0..len) = input; P16(state); output = state[0..8);
//   sample_bits(b): canonical(sample()) & (2^b - 1).
// One wavefront runs the transcript: lane i < 16 owns state word i; P16 runs
// on the whole wavefront (permute16_wave; BabyBear: 16-wide rows with DPP
// cross-lane moves).
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

// Cross-lane moves inside each 16-lane row via DPP (VALU latency, no LDS):
//   quad_perm(p0,p1,p2,p3): lane i of each quad reads lane p_i of that quad;
//   row_ror:n             : lane i of each 16-lane row reads lane (i - n) mod 16.
// Every lane's source is a valid lane under these controls, so bound_ctrl (0
// for an invalid source) changes no result; it lets the compiler drop the
// "old" operand, which would otherwise be materialised before each move
// (cooperative P16: 971 -> 904 VALU instructions, 1.72 -> 1.56 us).
template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
}
constexpr int DPP_QUAD_NEXT = 1 | (2 << 2) | (3 << 4) | (0 << 6);  // lane a <- a+1 (mod 4)
constexpr int DPP_QUAD_XOR1 = 1 | (0 << 2) | (3 << 4) | (2 << 6);
constexpr int DPP_QUAD_XOR2 = 2 | (3 << 2) | (0 << 4) | (1 << 6);
constexpr int DPP_ROR(int n) { return 0x120 + n; }

#if defined(WHIR_FIELD_BABYBEAR)
// Cooperative Poseidon2-16 over BabyBear (plain Montgomery arithmetic on
// canonical words): lane i (mod 16) holds state word i; the MDS-light chunk
// and column sums and the internal layer's row sum are DPP moves and modular
// additions inside each 16-lane row.
struct CoopP16 {
  uint32_t ei[4], et[4];  // this lane's external round constants (initial, terminal)
  uint32_t d;             // V_lane (Montgomery)

  __device__ __forceinline__ void load() {
    const Poseidon2Plain<16>& C = c_p2b.w16;
    const uint32_t i = threadIdx.x & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      ei[r] = C.ext_init[r][i];
      et[r] = C.ext_term[r][i];
    }
    d = C.diag[i];
  }
};

__device__ __forceinline__ uint32_t mds16_coop_plain(uint32_t x) {
  uint32_t t = add(x, dpp<DPP_QUAD_XOR1>(x));
  t = add(t, dpp<DPP_QUAD_XOR2>(t));  // the chunk's sum
  const uint32_t y = add(add(t, x), dbl(dpp<DPP_QUAD_NEXT>(x)));
  uint32_t u = add(y, dpp<DPP_ROR(4)>(y));
  u = add(u, dpp<DPP_ROR(8)>(u));  // the chunks' words at this position
  return add(y, u);
}

// In and out: the lane's state word in [0, p).
__device__ __forceinline__ void permute16_coop(uint32_t& s, const CoopP16& K) {
  const Poseidon2Plain<16>& C = c_p2b.w16;
  const bool lane0 = (threadIdx.x & 15) == 0;
  uint32_t x = mds16_coop_plain(s);
#pragma unroll
  for (int r = 0; r < 4; ++r) x = mds16_coop_plain(sbox7(add(x, K.ei[r])));
#pragma unroll 1
  for (int r = 0; r < Poseidon2Plain<16>::RP; ++r) {
    const uint32_t c = sbox7(add(x, C.internal[r]));
    x = lane0 ? c : x;
    uint32_t sum = add(x, dpp<DPP_ROR(8)>(x));
    sum = add(sum, dpp<DPP_ROR(4)>(sum));
    sum = add(sum, dpp<DPP_ROR(2)>(sum));
    sum = add(sum, dpp<DPP_ROR(1)>(sum));
    x = add(mul(x, K.d), sum);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) x = mds16_coop_plain(sbox7(add(x, K.et[r])));
  s = x;
}

using CoopLane = CoopP16;

// The transcript's permutation, out of line: the row form here (every row a
// copy).
__device__ __noinline__ uint32_t permute16_wave_shared(uint32_t s, CoopLane L) {
  permute16_coop(s, L);
  return s;
}

#else
// Cooperative Poseidon2-16: lane i (mod 16) holds state word i, with the
// lazy reduction and scaled state of permute_lazy (poseidon2.h). The per-lane
// constants are loaded into registers once per kernel (CoopP16::load); the
// wave-uniform ones (internal round constants, kappa, the final scale) are
// scalar operands read from constant memory.
struct CoopP16 {
  uint32_t la[9];  // lin_add[k][lane]
  uint32_t d;      // V_lane as an integer; 2 on lane 0, applied to 2p - x0: V_0 = -2, and the small
                   // product keeps lane 0's REDC near-canonical for the next S-box's square
  uint32_t m0;     // 1 on lane 0 of the row, else 0: lane 0's internal round constant enters as m0 * int_add

  __device__ __forceinline__ void load() {
    const Poseidon2Lazy<16>& C = c_p2l.w16;
    const uint32_t i = threadIdx.x & 15;
#pragma unroll
    for (int k = 0; k < 9; ++k) la[k] = C.lin_add[k][i];
    uint32_t dv = 2;
#pragma unroll
    for (int j = 1; j < 16; ++j) dv = i == (uint32_t)j ? Poseidon2Lazy<16>::diag(j) : dv;
    d = dv;
    m0 = i == 0 ? 1u : 0u;
  }
};
using CoopLane = CoopP16;

// The cross-lane sums below are exact 64-bit integers: a DPP add of the low
// words whose carry (vcc) enters the DPP add of the high words. Each asm
// block starts with two wait states, which a DPP read of a VGPR that the
// previous VALU instruction wrote needs; inside, one independent instruction
// and one s_nop separate a write from its DPP read.
#define KB_DPP_CTL " row_mask:0xf bank_mask:0xf bound_ctrl:1\n"

// Sum of x (< 2^32) over the 16 lanes of each row, on every lane.
__device__ __forceinline__ uint64_t row_sum64(uint32_t x) {
  uint32_t lo, hi;
  asm("s_nop 1\n"
      "v_add_co_u32_dpp %0, vcc, %2, %2 row_ror:8" KB_DPP_CTL
      "v_addc_co_u32_e64 %1, vcc, 0, 0, vcc\n"
      "s_nop 0\n"
      "v_add_co_u32_dpp %0, vcc, %0, %0 row_ror:4" KB_DPP_CTL
      "v_addc_co_u32_dpp %1, vcc, %1, %1, vcc row_ror:4" KB_DPP_CTL
      "s_nop 0\n"
      "v_add_co_u32_dpp %0, vcc, %0, %0 row_ror:2" KB_DPP_CTL
      "v_addc_co_u32_dpp %1, vcc, %1, %1, vcc row_ror:2" KB_DPP_CTL
      "s_nop 0\n"
      "v_add_co_u32_dpp %0, vcc, %0, %0 row_ror:1" KB_DPP_CTL
      "v_addc_co_u32_dpp %1, vcc, %1, %1, vcc row_ror:1" KB_DPP_CTL
      : "=&v"(lo), "=&v"(hi)
      : "v"(x)
      : "vcc");
  return ((uint64_t)hi << 32) | lo;
}

// REDC(MDS-light(x) + add) across the 16 lanes of a row; x < 2p, result
// < p + 2^7. Per chunk (quad) row a of M4 is S4 + x_a + 2 x_{a+1} (S4 the
// quad's sum); then every lane adds the sum over the four chunks at its
// position (rotations by 4 and 8).
__device__ __forceinline__ uint32_t mds16_coop_lazy(uint32_t x, uint32_t add) {
  uint32_t slo, shi, xn;
  asm("s_nop 1\n"
      "v_mov_b32_dpp %2, %3 quad_perm:[1,2,3,0]" KB_DPP_CTL
      "v_add_co_u32_dpp %0, vcc, %3, %3 quad_perm:[1,0,3,2]" KB_DPP_CTL
      "v_addc_co_u32_e64 %1, vcc, 0, 0, vcc\n"
      "s_nop 0\n"
      "v_add_co_u32_dpp %0, vcc, %0, %0 quad_perm:[2,3,0,1]" KB_DPP_CTL
      "v_addc_co_u32_dpp %1, vcc, %1, %1, vcc quad_perm:[2,3,0,1]" KB_DPP_CTL
      : "=&v"(slo), "=&v"(shi), "=&v"(xn)
      : "v"(x)
      : "vcc");
  const uint64_t y = mad_k<2>(xn, mad_k<1>(x, ((uint64_t)shi << 32) | slo));  // S4 + x_a + 2 x_{a+1}
  uint32_t ulo, uhi;
  asm("s_nop 1\n"
      "v_add_co_u32_dpp %0, vcc, %2, %2 row_ror:4" KB_DPP_CTL
      "v_addc_co_u32_dpp %1, vcc, %3, %3, vcc row_ror:4" KB_DPP_CTL
      "s_nop 0\n"
      "v_add_co_u32_dpp %0, vcc, %0, %0 row_ror:8" KB_DPP_CTL
      "v_addc_co_u32_dpp %1, vcc, %1, %1, vcc row_ror:8" KB_DPP_CTL
      : "=&v"(ulo), "=&v"(uhi)
      : "v"((uint32_t)y), "v"((uint32_t)(y >> 32))
      : "vcc");
  return redc(mad_k<1>(add, y + (((uint64_t)uhi << 32) | ulo)));  // y + column sums + add
}

// REDC(a k) for a constant k with kp = k * (-p^-1) mod 2^32: the Montgomery
// quotient a * kp does not wait for the product, so the chain is two deep.
__device__ __forceinline__ uint32_t mulr_k(uint32_t a, uint32_t k, uint32_t kp) {
  const uint64_t t = (uint64_t)a * k;
  return (uint32_t)(mad64(a * kp, P, t) >> 32);
}

// In: s = the lane's state word in [0, p). Out: the permuted word in [0, p).
// Internal rounds: lane 0's S-box (times kappa_r), then the row sum, then
// x_i <- REDC(sum + V_i x_i + c) as in internal_lazy (c on lane 0 only).
// Measured on one wavefront (100 chained calls): 1.84 us per permutation
// against 1.96 us with 16-bit limb sums. One wavefront pays about one issue
// slot (~4 cycles) per instruction, s_nop included, and ~7 cycles per
// instruction that waits on the one before; variants that took lanes 1..15's
// sum off lane 0's S-box chain (row_newbcast of its result) or filled the DPP
// wait states with the lane-0 select measured slower (2.10, 1.92 us).
__device__ __forceinline__ void permute16_coop(uint32_t& s, const CoopP16& K) {
  const Poseidon2Lazy<16>& C = c_p2l.w16;
  // the wave-uniform constants in one run of scalar loads
  uint32_t kappa[20], iadd[20];
#pragma unroll
  for (int r = 0; r < 20; ++r) {
    kappa[r] = C.kappa[r];
    iadd[r] = C.int_add[r];
  }
  const uint32_t fix = C.fix;
  const bool lane0 = K.m0 != 0;
  uint32_t x = mds16_coop_lazy(s, K.la[0]);
#pragma unroll
  for (int r = 1; r < 5; ++r) x = mds16_coop_lazy(mulr(mulr(x, x), x), K.la[r]);
#pragma unroll
  for (int r = 0; r < 20; ++r) {
    const uint32_t c = mulr_k(mulr(mulr(x, x), x), kappa[r], kappa[r] * NPINV);
    x = lane0 ? c : x;
    const uint64_t sum = row_sum64(x);
    const uint32_t y = lane0 ? 2u * P - x : x;
    const uint64_t s2 = r < 19 ? mad_u(K.m0, iadd[r + 1], sum) : mad_k<1>(K.la[5], sum);
    const uint64_t t = mad64(y, K.d, s2);  // < 2 p^2 + 2^37
    x = r < 19 ? redc(t) : canon2(redc(t));
  }
#pragma unroll
  for (int r = 6; r < 9; ++r) x = mds16_coop_lazy(mulr(mulr(x, x), x), K.la[r]);
  x = mds16_coop_lazy(mulr(mulr(x, x), x), 0u);
  s = canon2(mulr(x, fix));
}

// T with its high word below 2p -> the same value mod p with the high word
// below p (so that T < p R, the bound of redc).
__device__ __forceinline__ uint64_t hi_reduce(uint64_t t) {
  const uint32_t h = (uint32_t)(t >> 32), h2 = h - P;
  return ((uint64_t)(h < h2 ? h : h2) << 32) | (uint32_t)t;
}

// Poseidon2-16 on one wavefront: lane i < 16 holds word i in and out (lanes
// 16..63 carry no state: they end with unspecified values). External rounds
// as in permute16_coop; the 20 internal rounds as the recurrence of
// Poseidon2Wave16: every lane keeps one 64-bit accumulator (lane 16 + r: the
// S-box input of round r; lane i < 16: output word i), seeded from the 15
// input words 1..15 (read lane by lane as scalars). Round r cubes every
// lane's reduced accumulator, reads the cube of lane 16 + r as a scalar and
// adds coefficient x cube to every accumulator in one instruction, so a
// round's dependent chain is one multiply-add, two reductions, the S-box and
// a lane read, with no cross-lane sum. Measured on one wavefront (100 chained
// calls): 1.41 us per permutation against 1.94 us for permute16_coop (with
// the S-box on the scalar read, computed on the scalar unit: 1.74 us).
__device__ __forceinline__ void permute16_wave(uint32_t& s, const CoopP16& K) {
  const Poseidon2Lazy<16>& C = c_p2l.w16;
  const uint4* cp = reinterpret_cast<const uint4*>(c_p2w.col[threadIdx.x & 63]);
  uint32_t cw[36];
#pragma unroll
  for (int i = 0; i < 9; ++i) {  // issued first: the external rounds cover their latency
    const uint4 v = cp[i];
    cw[4 * i] = v.x;
    cw[4 * i + 1] = v.y;
    cw[4 * i + 2] = v.z;
    cw[4 * i + 3] = v.w;
  }
  const uint32_t fix = C.fix;
  uint32_t x = mds16_coop_lazy(s, K.la[0]);
#pragma unroll
  for (int r = 1; r < 5; ++r) x = mds16_coop_lazy(mulr(mulr(x, x), x), K.la[r]);
  // accumulators from the input words 1..15 (< p + 2^7): each term adds less
  // than p/2 to the high word, so it is reduced after terms 4, 6, ..., 14, 15
  uint64_t acc = cw[35];
#pragma unroll
  for (int k = 1; k < 16; ++k) {
    acc = mad64(cw[19 + k], (uint32_t)__builtin_amdgcn_readlane((int)x, k), acc);
    if (k >= 4 && (k % 2 == 0 || k == 15)) acc = hi_reduce(acc);
  }
  uint32_t xr = x;  // round 0's S-box input is lane 0's word
#pragma unroll
  for (int r = 0; r < 20; ++r) {
    // cube < 1.75 p (inputs < p + 2^7); the high word stays below 2p before the reduction
    const uint32_t y = (uint32_t)__builtin_amdgcn_readlane((int)mulr(mulr(xr, xr), xr), r == 0 ? 0 : 16 + r);
    acc = hi_reduce(mad64(cw[r], y, acc));
    xr = canon2(redc(acc));
  }
  x = xr;
#pragma unroll
  for (int r = 6; r < 9; ++r) x = mds16_coop_lazy(mulr(mulr(x, x), x), K.la[r]);
  x = mds16_coop_lazy(mulr(mulr(x, x), x), 0u);
  s = canon2(mulr(x, fix));
}

// The transcript's permutation, out of line. A transcript kernel is one
// workgroup running straight-line code once, so its time goes to fetching
// instructions into a cold instruction cache as much as to executing them; with
// the permutation (~600 instructions) inlined at every observe/sample site, each
// site is a copy of its own and each executed copy is fetched cold (measured on
// MI355X: a claim's draw, eq tables and value took ~14 us in one workgroup, 3-5x
// their instruction count). One shared copy is fetched once per kernel (and
// stays in L2 between kernels). Its arguments are the lane's state word and
// lane constants; the wave-uniform constants and the coefficient columns come
// from constant memory.
__device__ __noinline__ uint32_t permute16_wave_shared(uint32_t s, CoopLane L) {
  permute16_wave(s, L);
  return s;
}

#endif  // WHIR_FIELD_BABYBEAR

// One wavefront runs the transcript in uniform control flow. The sponge state
// and both rate buffers are lane-distributed so nothing is indexed dynamically:
//   s   : lane i < 16 holds state[i] (KoalaBear: lanes 16..63 hold no state
//         after a duplex, see permute16_wave; BabyBear: every row a copy);
//   pin : lane i (< 8) holds the pending input word i (written into the state at
//         the next duplex);
//   the output buffer is state[0..8) right after a duplex (the state does not
//   change until the next duplex), so pop() reads lane out_len - 1 of s.
struct WaveTx {
  uint32_t s, pin;
  uint32_t in_len, out_len;
  uint32_t log_words, log_cap, overflow;
  uint32_t* log;
  CoopLane k;

  // A fresh transcript: zero sponge state, empty buffers, an empty log of
  // `cap` words.
  // pre: lane constants the kernel has loaded already (else loaded here)
  __device__ __forceinline__ void fresh(uint32_t cap, uint32_t* log_ptr, const CoopLane* pre = nullptr) {
    s = 0u;
    pin = 0u;
    in_len = out_len = 0;
    log_words = 0;
    log_cap = cap;
    overflow = 0;
    log = log_ptr;
    if (pre) k = *pre;
    else k.load();
  }
  __device__ __forceinline__ void load(const DevTx* t, uint32_t* log_ptr, const CoopLane* pre = nullptr) {
    const int lane = threadIdx.x & 63;
    s = t->state[lane & 15];
    pin = t->in_buf[lane & 7];
    in_len = t->in_len;
    out_len = t->out_len;
    log_words = t->log_words;
    log_cap = t->log_cap;
    overflow = t->overflow;
    log = log_ptr;
    if (pre) k = *pre;
    else k.load();
  }
  __device__ __forceinline__ void store(DevTx* t) {
    const int lane = threadIdx.x & 63;
    if (lane < 16) t->state[lane] = s;
    if (lane < 8) {
      t->in_buf[lane] = pin;
      t->out_buf[lane] = s;
    }
    if (lane == 0) {
      t->in_len = in_len;
      t->out_len = out_len;
      t->log_words = log_words;
      t->log_cap = log_cap;
      t->overflow = overflow;
    }
  }
  __device__ __forceinline__ void duplex() {
    const uint32_t i = threadIdx.x & 15;
    if (i < in_len) s = pin;
    in_len = 0;
    s = permute16_wave_shared(s, k);
    out_len = 8;
  }
  __device__ __forceinline__ void observe(uint32_t v) {
    out_len = 0;
    if ((threadIdx.x & 15) == in_len) pin = v;
    in_len += 1;
    if (in_len == 8) duplex();
  }
  __device__ __forceinline__ uint32_t sample() {
    if (in_len != 0 || out_len == 0) duplex();
    out_len -= 1;
    return (uint32_t)__builtin_amdgcn_readlane((int)s, (int)out_len);
  }
  __device__ __forceinline__ Ext sample_ext() {
    Ext e;
    e.c[0] = sample();
    e.c[1] = sample();
    e.c[2] = sample();
    e.c[3] = sample();
    return e;
  }
  __device__ __forceinline__ uint32_t sample_bits(uint32_t bits) {
    return canonical(sample()) & ((1u << bits) - 1u);
  }
  // out[0..n) = n successive sample_bits(bits). After a duplex the next pops
  // are state[out_len - 1], state[out_len - 2], ..., so every draw that one
  // duplex serves is taken at once by the lane holding its word.
  __device__ __forceinline__ void sample_bits_many(uint32_t n, uint32_t bits, uint32_t* out) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t mask = (1u << bits) - 1u;
    if (n == 0) return;
    if (in_len != 0) duplex();
    for (uint32_t i = 0; i < n;) {
      if (out_len == 0) duplex();
      const uint32_t take = out_len < n - i ? out_len : n - i;
      if (lane < out_len && lane + take >= out_len) out[i + (out_len - 1 - lane)] = canonical(s) & mask;
      i += take;
      out_len -= take;
    }
  }
  __device__ __forceinline__ void hint_word(uint32_t v) {
    if ((threadIdx.x & 63) == 0) {
      if (log_words < log_cap) log[log_words] = v;
    }
    if (log_words >= log_cap) overflow = 1;
    log_words += 1;
  }
  __device__ __forceinline__ void write_f(uint32_t v) {
    hint_word(v);
    observe(v);
  }
  __device__ __forceinline__ void write_e(const Ext& e) {
    hint_word(e.c[0]);
    hint_word(e.c[1]);
    hint_word(e.c[2]);
    hint_word(e.c[3]);
    observe(e.c[0]);
    observe(e.c[1]);
    observe(e.c[2]);
    observe(e.c[3]);
  }
};

}  // namespace kb
