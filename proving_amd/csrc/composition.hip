// Composition polynomial of the wide-Fibonacci AIR on the evaluation domain.
//
// ComponentProver::evaluate_constraint_quotients_on_domain
// (crates/constraint-framework/src/prover/component_prover.rs:118-193) for
// WideFibonacciEval (crates/examples/src/wide_fibonacci/mod.rs:77-98): row r of
// the evaluation domain (bit-reversed order, the first 2^eval_log rows of the
// committed evaluations) reads the N trace columns c_0..c_{N-1} and adds
//   sum_j  w_j * (c_{j+2} - (c_j^2 + c_{j+1}^2))  *  denom_inv[r >> trace_log]
// to the QM31 accumulator, where w_j is the accumulator's random-coefficient
// power of constraint j (alpha^(M-1-j) for the only component, M = N - 2) and
// denom_inv[i] = 1 / coset_vanishing(trace coset, eval_domain.at(i)), bit-reversed
// (get_constraint_quotient_inputs, component_prover.rs:77-110).
//
// One lane per row; every load of a column is coalesced across the wave.  The
// weights are uniform across lanes (scalar loads).  The QM31 sum is kept as four
// unreduced 64-bit accumulators, folded every three constraints, so a
// constraint costs two squarings, one subtraction and four 32x32->64 multiplies.
#include "common.h"

namespace {

constexpr uint32_t kThreads = 256;

__device__ __forceinline__ uint64_t fold64(uint64_t v) {  // v mod P, up to 2^33 + small
  return (v & sh::P) + (v >> 31);
}

__global__ __launch_bounds__(kThreads) void k_wide_fib_composition(const uint32_t* const* __restrict__ cols,
                                                                   uint32_t n_cols, uint32_t eval_log,
                                                                   uint32_t trace_log,
                                                                   const uint32_t* __restrict__ w,
                                                                   const uint32_t* __restrict__ denom_inv,
                                                                   uint32_t* const* __restrict__ acc) {
  const uint32_t r = blockIdx.x * kThreads + threadIdx.x;
  if (r >= (1u << eval_log)) return;
  // the squares of the two previous columns: each column is squared once, for the next two
  // constraints
  const uint32_t a = cols[0][r], b = cols[1][r];
  uint32_t sa = sh::mmul(a, a), sb = sh::mmul(b, b);
  uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  const uint32_t m = n_cols - 2;
  auto step = [&](uint32_t c, uint32_t jj) {
    const uint32_t k = sh::msub(c, sh::madd(sa, sb));
    const uint4 wj = *(const uint4*)&w[4 * jj];
    s0 += (uint64_t)k * wj.x;
    s1 += (uint64_t)k * wj.y;
    s2 += (uint64_t)k * wj.z;
    s3 += (uint64_t)k * wj.w;
    sa = sb;
    sb = sh::mmul(c, c);
  };
  // three products of < 2^62 each on top of a folded value (< 2^34) stay below 2^64
  auto fold_all = [&]() {
    s0 = fold64(s0);
    s1 = fold64(s1);
    s2 = fold64(s2);
    s3 = fold64(s3);
  };
  uint32_t j = 0;
  // six columns per step, their loads issued together
  for (; j + 6 <= m; j += 6) {
    uint32_t c[6];
#pragma unroll
    for (int u = 0; u < 6; u++) c[u] = cols[j + 2 + u][r];
#pragma unroll
    for (int u = 0; u < 6; u++) {
      step(c[u], j + u);
      if (u == 2 || u == 5) fold_all();
    }
  }
  for (; j < m; j++) {
    step(cols[j + 2][r], j);
    fold_all();
  }
  const uint32_t d = denom_inv[r >> trace_log];
  auto finish = [&](uint64_t s, uint32_t* out) {
    const uint32_t v = sh::reduce_once((uint32_t)fold64(s));  // s < 2^34: one more fold is < 2P
    out[r] = sh::madd(out[r], sh::mmul(v, d));
  };
  finish(s0, acc[0]);
  finish(s1, acc[1]);
  finish(s2, acc[2]);
  finish(s3, acc[3]);
}

// generate_trace (crates/examples/src/wide_fibonacci/mod.rs:27-48) from its two input
// columns: one lane per row runs the recurrence c_k = c_{k-2}^2 + c_{k-1}^2 down the row.
__global__ __launch_bounds__(kThreads) void k_wide_fib_trace(const uint32_t* __restrict__ c0,
                                                             const uint32_t* __restrict__ c1, uint32_t n_rows,
                                                             uint32_t n_cols, uint32_t* const* __restrict__ out) {
  const uint32_t r = blockIdx.x * kThreads + threadIdx.x;
  if (r >= n_rows) return;
  const uint32_t a = c0[r], b = c1[r];
  out[0][r] = a;
  out[1][r] = b;
  uint32_t sa = sh::mmul(a, a), sb = sh::mmul(b, b);  // each column squared once
  for (uint32_t k = 2; k < n_cols; k++) {
    const uint32_t c = sh::madd(sa, sb);
    out[k][r] = c;
    sa = sb;
    sb = sh::mmul(c, c);
  }
}

}  // namespace

extern "C" int sh_wide_fibonacci_trace(sh_ctx* ctx, const uint32_t* d_c0, const uint32_t* d_c1, uint32_t log_size,
                                       uint64_t n_cols, uint32_t* const* d_cols) {
  if (!ctx || !d_c0 || !d_c1 || !d_cols) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, n_cols >= 2 && n_cols <= 4096 && log_size <= 30, "invalid trace shape (%llu columns, log %u)",
             (unsigned long long)n_cols, log_size);
  void* d_ptrs;
  SH_TRY(sh_stage(ctx, d_cols, n_cols * sizeof(void*), &d_ptrs));
  SpanGuard span(ctx, "trace_generation");
  const uint32_t rows = 1u << log_size;
  k_wide_fib_trace<<<(rows + kThreads - 1) / kThreads, kThreads, 0, ctx->stream>>>(d_c0, d_c1, rows, (uint32_t)n_cols,
                                                                                  (uint32_t* const*)d_ptrs);
  return sh_launch_check(ctx, "wide_fib_trace");
}

// coset_vanishing (crates/stwo/src/core/constraints.rs:12-35) of a coset given by
// index (initial, log) at the M31 point of index p: the x coordinate of
// (p - initial + step/2) doubled log - 1 times.
static uint32_t coset_vanishing_idx(uint32_t coset_initial, uint32_t coset_log, uint32_t p) {
  const uint32_t half_step = sh::subgroup_gen(coset_log) >> 1;
  uint64_t q = (uint64_t)p + (1ull << 31) - coset_initial + half_step;
  q <<= (coset_log - 1);
  return sh::index_to_point_host(sh::idx_reduce(q)).x;
}

// powers: host words (staged here) or, when powers_on_device, a device array the stream has
// already produced (4 words per constraint).
static int composition(sh_ctx* ctx, const uint32_t* const* d_cols, uint64_t n_cols, uint32_t trace_log,
                       uint32_t eval_log, uint32_t log_expansion, const uint32_t* powers, bool powers_on_device,
                       uint32_t* const d_accum[4]) {
  if (!ctx || !d_cols || !powers || !d_accum) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, n_cols >= 3 && n_cols <= 4096, "wide Fibonacci needs 3..4096 columns (got %llu)",
             (unsigned long long)n_cols);
  SH_REQUIRE(ctx, trace_log >= 1 && eval_log >= trace_log && eval_log <= 30 && eval_log - trace_log <= 8 &&
                      log_expansion <= 4,
             "invalid composition domain (trace %u, eval %u, expansion %u)", trace_log, eval_log, log_expansion);
  // evaluation domain: CanonicCoset(eval_log + log_expansion).circle_domain().split(log_expansion).0,
  // i.e. the circle domain whose half coset starts at G_(eval_log + log_expansion + 1) and has
  // log size eval_log - 1 (crates/constraint-framework/src/prover/component_prover.rs:234-238)
  const uint32_t half_init = sh::subgroup_gen(eval_log + log_expansion + 1);
  const uint32_t half_step = sh::subgroup_gen(eval_log - 1);
  const uint32_t n_denom = 1u << (eval_log - trace_log);
  std::vector<uint32_t> denom(n_denom);
  // trace coset: CanonicCoset(trace_log).coset() = odds of G_trace_log
  const uint32_t trace_init = sh::subgroup_gen(trace_log + 1);
  for (uint32_t i = 0; i < n_denom; i++) {
    // CircleDomain::at (crates/stwo/src/core/poly/circle/domain.rs): half coset, then its negation
    const uint32_t hs = 1u << (eval_log - 1);
    uint32_t p;
    if (i < hs) {
      p = sh::idx_reduce((uint64_t)half_init + (uint64_t)half_step * i);
    } else {
      p = sh::idx_reduce((1ull << 31) - sh::idx_reduce((uint64_t)half_init + (uint64_t)half_step * (i - hs)));
    }
    const uint32_t v = coset_vanishing_idx(trace_init, trace_log, p);
    SH_REQUIRE(ctx, v != 0, "the evaluation domain meets the trace domain");
    denom[sh::bitrev(i, eval_log - trace_log)] = sh::minv(v);
  }
  // staged tables: column pointers, weights (4 words per constraint), denominators
  void *d_ptrs, *d_w, *d_den, *d_acc;
  SH_TRY(sh_stage(ctx, d_cols, n_cols * sizeof(void*), &d_ptrs));
  if (powers_on_device)
    d_w = (void*)powers;
  else
    SH_TRY(sh_stage(ctx, powers, 16 * (n_cols - 2), &d_w));
  SH_TRY(sh_stage(ctx, denom.data(), 4ull * n_denom, &d_den));
  SH_TRY(sh_stage(ctx, d_accum, 4 * sizeof(void*), &d_acc));
  SpanGuard span(ctx, "composition");
  const uint32_t blocks = (uint32_t)(((1ull << eval_log) + kThreads - 1) / kThreads);
  k_wide_fib_composition<<<blocks, kThreads, 0, ctx->stream>>>(
      (const uint32_t* const*)d_ptrs, (uint32_t)n_cols, eval_log, trace_log, (const uint32_t*)d_w,
      (const uint32_t*)d_den, (uint32_t* const*)d_acc);
  return sh_launch_check(ctx, "wide_fib_composition");
}

extern "C" int sh_wide_fibonacci_composition(sh_ctx* ctx, const uint32_t* const* d_cols, uint64_t n_cols,
                                             uint32_t trace_log, uint32_t eval_log, uint32_t log_expansion,
                                             const uint32_t* powers, uint32_t* const d_accum[4]) {
  return composition(ctx, d_cols, n_cols, trace_log, eval_log, log_expansion, powers, false, d_accum);
}

extern "C" int sh_wide_fibonacci_composition_dev(sh_ctx* ctx, const uint32_t* const* d_cols, uint64_t n_cols,
                                                 uint32_t trace_log, uint32_t eval_log, uint32_t log_expansion,
                                                 const uint32_t* d_powers, uint32_t* const d_accum[4]) {
  return composition(ctx, d_cols, n_cols, trace_log, eval_log, log_expansion, d_powers, true, d_accum);
}

// ---------------------------------------------------------------- OODS sanity check
// The check closing prove_ex (crates/stwo/src/prover/mod.rs:118-132) for the wide-Fibonacci AIR,
// on the device: the composition value at the OODS point z rebuilt from its 8 sampled
// coordinates (StarkProof::extract_composition_oods_eval, crates/stwo/src/core/proof.rs:27-53:
// left + x * right, x = z.x doubled log_n - 1 times, each half from_partial_evals) against the
// constraints on the sampled trace values (Horner in alpha, i.e. constraint j weighted by
// alpha^(m-1-j), times 1 / coset_vanishing(CanonicCoset(log_n), z),
// crates/constraint-framework/src/component.rs:220-243).  values: n_cols trace values then the 8
// composition values (4 words each); powers_rev[j] = alpha^(m-1-j).  out: [mismatch flag, lhs (4
// words), rhs (4 words)].  One workgroup.
namespace {

__device__ sh::qm31 from_partial_evals(const uint32_t* e) {
  const sh::qm31 bi = {{0, 1}, {0, 0}}, bu = {{0, 0}, {1, 0}}, biu = {{0, 0}, {0, 1}};
  sh::qm31 r = sh::qload(e);
  r = sh::qadd(r, sh::qmul(sh::qload(e + 4), bi));
  r = sh::qadd(r, sh::qmul(sh::qload(e + 8), bu));
  return sh::qadd(r, sh::qmul(sh::qload(e + 12), biu));
}

__device__ sh::qm31 double_x_times(sh::qm31 x, uint32_t n) {
  for (uint32_t k = 0; k < n; k++) {
    const sh::qm31 xx = sh::qmul(x, x);
    x = sh::qsub(sh::qadd(xx, xx), sh::qone());
  }
  return x;
}

__global__ __launch_bounds__(256) void k_wide_fib_oods_check(const uint32_t* __restrict__ values, uint32_t n_cols,
                                                             const uint32_t* __restrict__ powers_rev,
                                                             const uint32_t* __restrict__ point, uint32_t shift_x,
                                                             uint32_t shift_y, uint32_t log_n,
                                                             uint32_t* __restrict__ out) {
  __shared__ uint32_t red[256][4];
  const uint32_t t = threadIdx.x, m = n_cols - 2;
  sh::qm31 acc = sh::qzero();
  for (uint32_t j = t; j < m; j += 256) {
    const sh::qm31 a = sh::qload(values + 4 * j), b = sh::qload(values + 4 * (j + 1)),
                   c = sh::qload(values + 4 * (j + 2));
    const sh::qm31 cons = sh::qsub(c, sh::qadd(sh::qmul(a, a), sh::qmul(b, b)));
    acc = sh::qadd(acc, sh::qmul(sh::qload(powers_rev + 4 * j), cons));
  }
  sh::qstore(acc, red[t]);
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (t < w) sh::qstore(sh::qadd(sh::qload(red[t]), sh::qload(red[t + w])), red[t]);
    __syncthreads();
  }
  if (t != 0) return;
  const sh::qm31 zx = sh::qload(point), zy = sh::qload(point + 4);
  // coset_vanishing (crates/stwo/src/core/constraints.rs:12-35): x of z - initial + step / 2
  const sh::qm31 px = sh::qsub(sh::qmulm(zx, shift_x), sh::qmulm(zy, shift_y));
  const sh::qm31 rhs = sh::qmul(sh::qinv(double_x_times(px, log_n - 1)), sh::qload(red[0]));
  const uint32_t* comp = values + 4 * n_cols;
  const sh::qm31 lhs =
      sh::qadd(from_partial_evals(comp), sh::qmul(double_x_times(zx, log_n - 1), from_partial_evals(comp + 16)));
  out[0] = sh::qeq(lhs, rhs) ? 0u : 1u;
  sh::qstore(lhs, out + 1);
  sh::qstore(rhs, out + 5);
}

}  // namespace

extern "C" int sh_wide_fibonacci_oods_check(sh_ctx* ctx, const uint32_t* d_values, uint64_t n_cols,
                                            const uint32_t* d_powers_rev, const uint32_t* d_point, uint32_t log_n,
                                            uint32_t* d_out) {
  if (!ctx || !d_values || !d_point || !d_out || (n_cols > 2 && !d_powers_rev)) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, n_cols >= 2 && n_cols < (1ull << 24), "column count %llu out of range",
             (unsigned long long)n_cols);
  SH_REQUIRE(ctx, log_n >= 1 && log_n <= 30, "trace log size %u out of range", log_n);
  // CanonicCoset(log_n).coset: initial index 2^(31 - (log_n + 1)), step 2^(31 - log_n)
  const uint32_t init = 1u << (31 - (log_n + 1)), half_step = (1u << (31 - log_n)) >> 1;
  const sh::mpt shift = sh::index_to_point_host((uint32_t)((half_step - init) & 0x7fffffffu));
  SpanGuard span(ctx, "composition");
  k_wide_fib_oods_check<<<1, 256, 0, ctx->stream>>>(d_values, (uint32_t)n_cols, d_powers_rev, d_point, shift.x,
                                                     shift.y, log_n, d_out);
  return sh_launch_check(ctx, "wide_fib_oods_check");
}
