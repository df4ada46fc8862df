// Out-of-domain sampling: PolyOps::eval_at_point, barycentric_weights_into,
// barycentric_eval_at_point (crates/stwo/src/prover/poly/circle/ops.rs:44-80).
//
// eval_at_point (CpuBackend, crates/stwo/src/prover/backend/cpu/circle.rs:81-98):
// value = sum_i c_i * prod_{bit k of i} f_k with f_0 = p.y, f_1 = p.x,
// f_k = pi^(k-1)(p.x) (the fold of crates/stwo/src/core/poly/utils.rs:27-38).
// The device folds 4 index bits per pass: lane t combines 16 consecutive
// values with the next four factors, so a 2^n polynomial takes ceil(n/4) passes.
//
// Barycentric weights (cpu/circle.rs:100-131), for the bit-reversed domain point q_i:
//   W_i = V_n(p) / (-2 q_i.y * V'_n(q_i) * (h.y / (1 + h.x))),  h = p - q_i,
// with V'_n the coset-vanishing derivative (crates/stwo/src/core/constraints.rs:37-55).
// q_i has M31 coordinates, so V'_n(q_i) is evaluated in M31.
#include "common.h"

namespace {

struct Factors {
  sh::qm31 f[4];
};

// in: M31 (first pass, in_m31 = 1) or QM31 AoS (4 words); out: QM31 AoS.
__global__ __launch_bounds__(256) void k_fold_pass(const uint32_t* __restrict__ in, int in_m31, uint32_t n_in,
                                                   uint32_t g, Factors fac, uint32_t* __restrict__ out) {
  const uint32_t n_out = n_in >> g;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_out) return;
  sh::qm31 v[16];
  const uint32_t cnt = 1u << g;
  for (uint32_t j = 0; j < cnt; j++) {
    uint32_t idx = (t << g) + j;
    v[j] = in_m31 ? sh::qfromm(in[idx]) : sh::qload(in + 4 * idx);
  }
  for (uint32_t k = 0; k < g; k++) {
    uint32_t half = cnt >> (k + 1);
    for (uint32_t j = 0; j < half; j++) v[j] = sh::qadd(v[2 * j], sh::qmul(v[2 * j + 1], fac.f[k]));
  }
  sh::qstore(v[0], out + 4 * t);
}

struct BaryConsts {
  uint32_t log_n;
  sh::qpt p;
  sh::qm31 vn_p;
  sh::mpt shift[31];  // coset_vanishing shift of CanonicCoset(j).coset, j = 1..log_n-1
  uint32_t four_pow;  // 4^(log_n - 1)
};

__global__ __launch_bounds__(256) void k_bary_weights(BaryConsts bc, sh_point_tab half, uint32_t* const* w) {
  const uint32_t n = 1u << bc.log_n;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sh::mpt q = sh_circle_domain_point(half, bc.log_n - 1, sh::bitrev(i, bc.log_n));
  // V'_n(q) = 4^(n-1) * prod_{j=1}^{n-1} coset_vanishing(CanonicCoset(j).coset, q)
  uint32_t d = bc.four_pow;
  for (uint32_t j = 1; j < bc.log_n; j++) {
    uint32_t x = sh::padd(q, bc.shift[j]).x;
    for (uint32_t r = 1; r < j; r++) x = sh::msub(sh::madd(sh::mmul(x, x), sh::mmul(x, x)), 1);
    d = sh::mmul(d, x);
  }
  uint32_t si = sh::mmul(sh::mneg(sh::madd(q.y, q.y)), d);
  // h = p - q = p + conj(q)
  sh::qm31 hx = sh::qadd(sh::qmulm(bc.p.x, q.x), sh::qmulm(bc.p.y, q.y));
  sh::qm31 hy = sh::qsub(sh::qmulm(bc.p.y, q.x), sh::qmulm(bc.p.x, q.y));
  sh::qm31 den = sh::qmulm(hy, si);
  sh::qm31 wv = sh::qmul(sh::qmul(bc.vn_p, sh::qadd(sh::qone(), hx)), sh::qinv(den));
  w[0][i] = wv.a.a;
  w[1][i] = wv.a.b;
  w[2][i] = wv.b.a;
  w[3][i] = wv.b.b;
}

__device__ __forceinline__ void lfold(uint64_t& v) { v = (v & sh::P) + (v >> 31); }
__device__ __forceinline__ uint32_t lfinal(uint64_t v) {
  lfold(v);
  lfold(v);
  return sh::reduce_once((uint32_t)v);
}

// partial[block] = sum_i w_i * e_i over a grid-stride range (lazy 64-bit accumulation).
__global__ __launch_bounds__(256) void k_bary_dot(const uint32_t* __restrict__ e, const uint32_t* const* w,
                                                  uint32_t n, uint32_t* partial) {
  __shared__ uint32_t red[256][4];
  uint64_t acc[4] = {0, 0, 0, 0};
  uint32_t k = 0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    uint32_t v = e[i];
#pragma unroll
    for (int c = 0; c < 4; c++) acc[c] += (uint64_t)w[c][i] * v;
    if (++k == 3) {
      k = 0;
#pragma unroll
      for (int c = 0; c < 4; c++) lfold(acc[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) red[threadIdx.x][c] = lfinal(acc[c]);
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int c = 0; c < 4; c++) red[threadIdx.x][c] = sh::madd(red[threadIdx.x][c], red[threadIdx.x + s][c]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 4; c++) partial[4 * blockIdx.x + c] = red[0][c];
}

__global__ void k_sum_partials(const uint32_t* partial, uint32_t nblk, uint32_t* out) {
  uint32_t s[4] = {0, 0, 0, 0};
  for (uint32_t b = 0; b < nblk; b++)
    for (int c = 0; c < 4; c++) s[c] = sh::madd(s[c], partial[4 * b + c]);
  for (int c = 0; c < 4; c++) out[c] = s[c];
}

// coset_vanishing(coset, p) for QM31 p on the host (constraints.rs:12-35).
sh::qm31 coset_vanishing_host(uint32_t init_index, uint32_t step_index, uint32_t log_size, sh::qpt p) {
  sh::mpt ci = sh::index_to_point_host(init_index);
  sh::mpt hs = sh::index_to_point_host(step_index >> 1);
  sh::qpt neg_init = {sh::qfromm(ci.x), sh::qfromm(sh::mneg(ci.y))};
  sh::qpt q = sh::qpadd(sh::qpadd(p, neg_init), {sh::qfromm(hs.x), sh::qfromm(hs.y)});
  sh::qm31 x = q.x;
  for (uint32_t i = 1; i < log_size; i++) x = sh::qdouble_x(x);
  return x;
}

}  // namespace

extern "C" int sh_eval_at_point(sh_ctx* ctx, const uint32_t* d_coeffs, uint32_t log_size, const uint32_t point[8],
                                uint32_t out[4]) {
  if (!ctx || !d_coeffs || !point || !out) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, log_size <= 30, "log size %u too large", log_size);
  sh::qpt p = {sh::qload(point), sh::qload(point + 4)};
  if (log_size == 0) {
    uint32_t c0;
    SH_CHECK_HIP(ctx, hipMemcpyAsync(&c0, d_coeffs, 4, hipMemcpyDeviceToHost, ctx->stream));
    SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    sh::qstore(sh::qfromm(c0), out);
    return SH_OK;
  }
  // factor for index bit k: bit 0 -> y, bit 1 -> x, bit k -> pi^(k-1)(x)
  std::vector<sh::qm31> f(log_size);
  f[0] = p.y;
  sh::qm31 x = p.x;
  for (uint32_t k = 1; k < log_size; k++) {
    f[k] = x;
    x = sh::qdouble_x(x);
  }
  uint64_t n = 1ull << log_size;
  void* scr;
  SH_TRY(sh_scratch_get(ctx, SCR_MISC, 16ull * ((n + 15) / 16) * 2 + 64, &scr));
  uint32_t* bufs[2] = {(uint32_t*)scr, (uint32_t*)scr + 4 * ((n + 15) / 16)};
  const uint32_t* in = d_coeffs;
  int in_m31 = 1;
  uint32_t n_in = (uint32_t)n, bit = 0, which = 0;
  SpanGuard span(ctx, "eval_at_point");
  while (n_in > 1) {
    uint32_t g = std::min<uint32_t>(4, log_size - bit);
    Factors fac;
    for (uint32_t k = 0; k < g; k++) fac.f[k] = f[bit + k];
    uint32_t n_out = n_in >> g;
    k_fold_pass<<<(n_out + 255) / 256, 256, 0, ctx->stream>>>(in, in_m31, n_in, g, fac, bufs[which]);
    SH_TRY(sh_launch_check(ctx, "eval_at_point fold"));
    in = bufs[which];
    in_m31 = 0;
    which ^= 1;
    n_in = n_out;
    bit += g;
  }
  SH_CHECK_HIP(ctx, hipMemcpyAsync(out, in, 16, hipMemcpyDeviceToHost, ctx->stream));
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SH_OK;
}

extern "C" int sh_barycentric_weights(sh_ctx* ctx, uint32_t log_size, const uint32_t point[8], uint32_t* const d_w[4]) {
  if (!ctx || !point || !d_w) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, log_size >= 1 && log_size <= 30, "log size %u out of range", log_size);
  BaryConsts bc = {};
  bc.log_n = log_size;
  bc.p = {sh::qload(point), sh::qload(point + 4)};
  // V_n(p) on CanonicCoset(n).coset = Coset::odds(n)
  bc.vn_p = coset_vanishing_host(sh::subgroup_gen(log_size + 1), sh::subgroup_gen(log_size), log_size, bc.p);
  for (uint32_t j = 1; j < log_size; j++) {
    // q - initial + step/2 for CanonicCoset(j).coset: a single point shift
    sh::mpt ci = sh::index_to_point_host(sh::subgroup_gen(j + 1));
    sh::mpt hs = sh::index_to_point_host(sh::subgroup_gen(j) >> 1);
    bc.shift[j] = sh::padd({ci.x, sh::mneg(ci.y)}, hs);
  }
  bc.four_pow = sh::mpow(4, log_size - 1);
  sh_point_tab half;
  SH_TRY(sh_point_tab_make(ctx, sh::subgroup_gen(log_size + 1), log_size - 1, &half));
  void* d_wt;
  SH_TRY(sh_stage(ctx, d_w, 4 * sizeof(void*), &d_wt));
  SpanGuard span(ctx, "barycentric_weights");
  k_bary_weights<<<((1u << log_size) + 255) / 256, 256, 0, ctx->stream>>>(bc, half, (uint32_t* const*)d_wt);
  return sh_launch_check(ctx, "barycentric_weights");
}

extern "C" int sh_barycentric_eval_at_point(sh_ctx* ctx, const uint32_t* d_evals, uint32_t log_size,
                                            const uint32_t* const d_w[4], uint32_t out[4]) {
  if (!ctx || !d_evals || !d_w || !out) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, log_size <= 30, "log size %u too large", log_size);
  uint32_t n = 1u << log_size;
  uint32_t nblk = std::min<uint32_t>((n + 255) / 256, 1024);
  void* scr;
  SH_TRY(sh_scratch_get(ctx, SCR_MISC, 16ull * nblk + 16, &scr));
  uint32_t* partial = (uint32_t*)scr;
  uint32_t* res = partial + 4 * nblk;
  void* d_wt;
  SH_TRY(sh_stage(ctx, d_w, 4 * sizeof(void*), &d_wt));
  SpanGuard span(ctx, "barycentric_eval");
  k_bary_dot<<<nblk, 256, 0, ctx->stream>>>(d_evals, (const uint32_t* const*)d_wt, n, partial);
  k_sum_partials<<<1, 1, 0, ctx->stream>>>(partial, nblk, res);
  SH_TRY(sh_launch_check(ctx, "barycentric_eval"));
  SH_CHECK_HIP(ctx, hipMemcpyAsync(out, res, 16, hipMemcpyDeviceToHost, ctx->stream));
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SH_OK;
}
