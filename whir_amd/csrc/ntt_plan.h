// Host-side planning of the coset-form NTT (ntt.h): how a transform splits
// into passes, and each pass's kernel arguments and launch shape. Shared by the
// library's launches and by host emulations of the pass kernels.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstring>
#include "ntt.h"

namespace kb {

constexpr int NTT3_LT = 6;  // log2 threads per line of an ntt3_kernel pass

// Plan of one coset-form transform: N = 2^log_nn output rows of C = 2^log_c
// columns from n_in input rows (n = 2^log_n >= n_in rows of the work array,
// B = 2^log_b = (N / n) C columns), P passes of L_p = 2^logl[p] rows (<= 256,
// or 512 / 1024 for an ntt3_kernel pass).
struct NttPlan {
  int log_nn = 0, log_c = 0, log_n = 0, log_b = 0, lv = 0, P = 0;
  int logl[3] = {0, 0, 0};
  bool tri[3] = {false, false, false};  // pass p is an ntt3_kernel pass (logl[p] = 9 or 10)
  bool wide[3] = {false, false, false}; // pass p is an ntt_wide_kernel pass (logl[p] = 9 or 10)
  bool reg_first = false;  // pass 0 is ntt_kernel<logl[0], 0, V> (no LDS)
};

inline bool ntt_plan(uint64_t n_in, int log_nn, int log_c, NttPlan& pl, int max_logl, int first_log = 0,
                     int max_lv = 2, int two_pass = 0) {
  if (log_nn < 1 || log_nn > 24 || log_c < 0 || log_c > 10 || log_nn + log_c > 31) return false;
  if (n_in > (1ull << log_nn)) return false;
  int log_n = 1;
  while ((1ull << log_n) < n_in) ++log_n;
  if (log_n > log_nn) return false;
  pl.log_nn = log_nn;
  pl.log_c = log_c;
  pl.log_n = log_n;
  pl.log_b = log_c + (log_nn - log_n);
  pl.lv = std::min(std::min(2, max_lv), log_c);
  pl.reg_first = false;
  pl.tri[0] = pl.tri[1] = pl.tri[2] = false;
  pl.wide[0] = pl.wide[1] = pl.wide[2] = false;
  if (two_pass == 2 && log_n >= 18 && log_n <= 20 && log_c >= 4) {
    // two passes of 9-10 bits, each two register stages of up to 32 points,
    // two words per lane and line (ntt_wide_kernel)
    pl.P = 2;
    pl.lv = 1;
    pl.logl[0] = (log_n + 1) / 2;
    pl.logl[1] = log_n / 2;
    pl.wide[0] = pl.wide[1] = true;
    return true;
  }
  if (two_pass == 1 && log_n >= 17 && log_n <= 20) {
    pl.P = 2;
    pl.logl[0] = (log_n + 1) / 2;
    pl.logl[1] = log_n / 2;
    pl.tri[0] = pl.logl[0] >= 9;
    pl.tri[1] = pl.logl[1] >= 9;
    return true;
  }
  if (first_log > 0 && log_n > 2 * std::min(max_logl, 8) && log_n - first_log <= 16) {
    // a transform that needs three passes anyway: a register-only first pass,
    // then two even passes
    pl.reg_first = true;
    pl.P = 3;
    pl.logl[0] = first_log;
    pl.logl[1] = (log_n - first_log + 1) / 2;
    pl.logl[2] = log_n - first_log - pl.logl[1];
    return true;
  }
  max_logl = std::max(max_logl, (log_n + 2) / 3);  // at most three passes
  pl.P = (log_n + max_logl - 1) / max_logl;
  if (pl.P > 3) return false;
  int rem = log_n;
  for (int p = 0; p < pl.P; ++p) {
    pl.logl[p] = (rem + (pl.P - p) - 1) / (pl.P - p);
    rem -= pl.logl[p];
  }
  return true;
}

enum { NTT_PASS_LDS = 0, NTT_PASS_REG = 1, NTT_PASS_TRI = 2, NTT_PASS_WIDE = 3 };
constexpr int NTT_WIDE_LR1 = 5;  // first register stage of an ntt_wide_kernel pass
constexpr int NTT_WIDE_LT = 8;   // log2 threads of an ntt_wide_kernel workgroup

struct NttPass {
  NttArgs a;
  unsigned grid = 0, threads = 0;
  size_t lds = 0;
  int kind = NTT_PASS_LDS;  // ntt_kernel<l - l/2, l/2>, ntt_kernel<l, 0> or ntt3_kernel
};

// Arguments and launch shape of pass p of plan pl; log_threads caps the
// workgroup size at 2^log_threads threads (more when one line needs more).
inline void ntt_pass_setup(const NttPlan& pl, int p, uint64_t n_in, int log_threads, const uint32_t* in,
                           uint32_t* out, uint32_t* tmp, const uint32_t* tw_lo, const uint32_t* tw_hi, NttPass& ps) {
  int used = 0;
  for (int q = 0; q <= p; ++q) used += pl.logl[q];
  const int l = pl.logl[p];
  ps.kind = (pl.reg_first && p == 0) ? NTT_PASS_REG
            : pl.tri[p]                ? NTT_PASS_TRI
            : pl.wide[p]               ? NTT_PASS_WIDE
                                       : NTT_PASS_LDS;
  const int lr2 = ps.kind == NTT_PASS_REG    ? 0
                  : ps.kind == NTT_PASS_TRI  ? NTT3_LT
                  : ps.kind == NTT_PASS_WIDE ? l - NTT_WIDE_LR1
                                             : l / 2;  // log2 threads per line
  NttArgs& a = ps.a;
  memset(&a, 0, sizeof(a));
  a.tw_lo = tw_lo;
  a.tw_hi = tw_hi;
  a.n_in = (uint32_t)n_in;
  a.log_c = (uint32_t)pl.log_c;
  a.log_nn = (uint32_t)pl.log_nn;
  a.log_b = (uint32_t)pl.log_b;
  a.log_n = (uint32_t)pl.log_n;
  a.log_m = (uint32_t)(pl.log_n - used);
  const int log_lines = pl.log_n + pl.log_b - pl.lv - l;
  const int lt = ps.kind == NTT_PASS_TRI    ? log_threads
                 : ps.kind == NTT_PASS_WIDE ? NTT_WIDE_LT
                                            : std::min(log_threads, 8);  // ntt_kernel: <= 256 threads
  a.log_qv = (uint32_t)std::max(0, std::min(lt - lr2, log_lines));
  a.first = p == 0;
  a.twiddle_out = p + 1 < pl.P;
  a.natural = pl.P > 1 && p + 1 == pl.P;
  a.log_l1 = (uint32_t)pl.logl[0];
  a.log_l2 = (uint32_t)pl.logl[1];
  a.n_pass = (uint32_t)pl.P;
  a.src = p == 0 ? in : tmp;
  a.dst = p + 1 == pl.P ? out : tmp;
  ps.grid = 1u << (log_lines - (int)a.log_qv);
  ps.threads = 1u << (lr2 + (int)a.log_qv);
  ps.lds = ps.kind == NTT_PASS_REG ? 0 : ((size_t)1 << (l + (int)a.log_qv + pl.lv)) * 4;
}

}  // namespace kb
