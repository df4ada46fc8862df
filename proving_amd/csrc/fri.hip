// FRI folds (FriOps::fold_line, FriOps::fold_circle_into_line, FriOps::decompose).
//
// CpuBackend (crates/stwo/src/prover/backend/cpu/fri.rs:66-116): for output i,
// (f0, f1) = ibutterfly(e[2i], e[2i+1], t^-1) and out[i] = f0 + alpha * f1, with
// t = x of LineDomain.at(bitrev(2i)) for a line fold and t = y of
// CircleDomain.at(bitrev(2i)) for the circle-to-line fold.  Those inverses are
// exactly the inverse-twiddle tree entries the SIMD backend reads
// (crates/stwo/src/prover/backend/simd/fri.rs:32-166): line layer of the
// domain for x^-1, the circle layer [y, -y, -x, x] derivation for y^-1.
//
// One lane per output element; a lane folds 2^fold_step inputs in registers
// (fold_step <= 4), reading each coordinate array once.  HBM-bound.
#include "blake2s.h"
#include "common.h"

namespace {

struct Alphas {
  sh::qm31 a[4];
};

__device__ __forceinline__ uint32_t circle_itw(const uint32_t* l1, uint32_t h) {
  uint32_t x = l1[(h >> 2) * 2], y = l1[(h >> 2) * 2 + 1];
  switch (h & 3) {
    case 0: return y;
    case 1: return sh::mneg(y);
    case 2: return sh::mneg(x);
    default: return x;
  }
}

__device__ __forceinline__ sh::qm31 ld4(const uint32_t* const* c, uint32_t i) {
  return {{c[0][i], c[1][i]}, {c[2][i], c[3][i]}};
}
__device__ __forceinline__ void st4(uint32_t* const* c, uint32_t i, sh::qm31 v) {
  c[0][i] = v.a.a;
  c[1][i] = v.a.b;
  c[2][i] = v.b.a;
  c[3][i] = v.b.b;
}

struct Cols4 {
  const uint32_t* s[4];
  uint32_t* d[4];
};

// The leaf of a folded value in the next layer's tree (one value per leaf: Blake2s of its 4
// coordinates, 16 bytes; FriInnerLayerProver::new, crates/stwo/src/prover/fri.rs:357-368), so
// that a fold also writes the next layer's leaf layer and the next commit starts above it.
__device__ __forceinline__ void leaf_of(sh::qm31 v, uint32_t* __restrict__ leaves, uint32_t i) {
  uint32_t m[16], h[8];
  m[0] = v.a.a;
  m[1] = v.a.b;
  m[2] = v.b.a;
  m[3] = v.b.b;
#pragma unroll
  for (int k = 4; k < 16; k++) m[k] = 0;
  sh::b2_init(h);
  sh::b2_compress(h, m, 16, true);
  uint4* o = (uint4*)(leaves + 8ull * i);
  o[0] = make_uint4(h[0], h[1], h[2], h[3]);
  o[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

// Alphas come by value, or from device memory (d_al != nullptr: 4 words per fold, written by an
// earlier kernel on the stream, e.g. the device-side channel of the FRI driver).
template <int F>
__global__ __launch_bounds__(256) void k_fold_line(Cols4 io, const uint32_t* __restrict__ itw, uint32_t K,
                                                   uint32_t n, Alphas al, const uint32_t* __restrict__ d_al,
                                                   uint32_t* __restrict__ leaves) {
  const uint32_t n_out = 1u << (n - F);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  if (d_al) {
#pragma unroll
    for (int f = 0; f < F; f++) al.a[f] = sh::qload(d_al + 4 * f);
  }
  sh::qm31 v[1 << F];
#pragma unroll
  for (int j = 0; j < (1 << F); j++) v[j] = ld4(io.s, (i << F) + j);
#pragma unroll
  for (int f = 0; f < F; f++) {
    // the domain has been doubled f times: a coset of log (n - f), tree layer at 2^K - 2^(n-f)
    const uint32_t* tab = itw + (1u << K) - (1u << (n - f));
    const int half = 1 << (F - f - 1);
#pragma unroll
    for (int jj = 0; jj < half; jj++) {
      uint32_t t = tab[i * half + jj];
      sh::qm31 a = v[2 * jj], b = v[2 * jj + 1];
      sh::qm31 f0 = sh::qadd(a, b);
      sh::qm31 f1 = sh::qmulm(sh::qsub(a, b), t);
      v[jj] = sh::qadd(f0, sh::qmul(al.a[f], f1));
    }
  }
  st4(io.d, i, v[0]);
  if (leaves) leaf_of(v[0], leaves, i);
}

// y^-1 from the circle layer of the inverse tree; domains of log size <= 2 take
// host-computed inverses (c0, c1) for outputs 0 and 1.
__global__ __launch_bounds__(256) void k_fold_circle(Cols4 io, const uint32_t* __restrict__ itw, uint32_t K,
                                                     uint32_t m, sh::qm31 alpha, const uint32_t* __restrict__ d_al,
                                                     uint32_t c0, uint32_t c1, uint32_t* __restrict__ leaves) {
  const uint32_t n_out = 1u << (m - 1);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  if (d_al) alpha = sh::qload(d_al);
  uint32_t yinv;
  if (m <= 2) {
    yinv = i == 0 ? c0 : c1;
  } else {
    yinv = circle_itw(itw + (1u << K) - (1u << (m - 1)), i);
  }
  sh::qm31 a = ld4(io.s, 2 * i), b = ld4(io.s, 2 * i + 1);
  sh::qm31 f0 = sh::qadd(a, b);
  sh::qm31 f1 = sh::qmulm(sh::qsub(a, b), yinv);
  const sh::qm31 r = sh::qadd(f0, sh::qmul(alpha, f1));
  st4(io.d, i, r);
  if (leaves) leaf_of(r, leaves, i);
}

// decompose: per-block sums of the first and second halves.
__global__ __launch_bounds__(256) void k_half_sums(Cols4 io, uint32_t n, uint32_t* partial) {
  __shared__ uint32_t red[256][4];
  const uint32_t half = 1u << (n - 1);
  sh::qm31 acc = sh::qzero();
  const uint32_t which = blockIdx.y;  // 0: first half, 1: second half
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < half; i += gridDim.x * 256)
    acc = sh::qadd(acc, ld4(io.s, which * half + i));
  sh::qstore(acc, red[threadIdx.x]);
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sh::qm31 x = sh::qload(red[threadIdx.x]), y = sh::qload(red[threadIdx.x + s]);
      sh::qstore(sh::qadd(x, y), red[threadIdx.x]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) sh::qstore(sh::qload(red[0]), partial + 4 * (which * gridDim.x + blockIdx.x));
}

__global__ void k_decompose_finish(const uint32_t* partial, uint32_t nblk, uint32_t inv_n, uint32_t* lambda_out) {
  sh::qm31 a = sh::qzero(), b = sh::qzero();
  for (uint32_t k = 0; k < nblk; k++) {
    a = sh::qadd(a, sh::qload(partial + 4 * k));
    b = sh::qadd(b, sh::qload(partial + 4 * (nblk + k)));
  }
  sh::qstore(sh::qmulm(sh::qsub(a, b), inv_n), lambda_out);
}

__global__ __launch_bounds__(256) void k_decompose_apply(Cols4 io, uint32_t n, const uint32_t* lambda_w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1u << n)) return;
  sh::qm31 lam = sh::qload(lambda_w);
  sh::qm31 x = ld4(io.s, i);
  st4(io.d, i, i < (1u << (n - 1)) ? sh::qsub(x, lam) : sh::qadd(x, lam));
}

}  // namespace

static int fold_line_launch(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const src[4],
                            uint32_t line_initial_index, uint32_t log_size, const Alphas& al, const uint32_t* d_al,
                            uint32_t fold_step, uint32_t* const dst[4], uint32_t* leaves = nullptr) {
  SH_REQUIRE(ctx, !leaves || fold_step == 1, "leaf hashes come with single folds only");
  SH_REQUIRE(ctx, fold_step >= 1 && fold_step <= 4, "fold_step %u not in [1, 4]", fold_step);
  SH_REQUIRE(ctx, log_size >= fold_step && log_size <= 30, "Evaluation too small (log %u, fold_step %u)",
             log_size, fold_step);
  SH_REQUIRE(ctx, sh::is_doubling_of(line_initial_index, log_size, tw->root_initial_index, tw->root_log_size),
             "line domain (initial %u, log %u) is not in the twiddle tower", line_initial_index, log_size);
  Cols4 io;
  for (int c = 0; c < 4; c++) {
    io.s[c] = src[c];
    io.d[c] = dst[c];
  }
  uint32_t n_out = 1u << (log_size - fold_step);
  unsigned grid = (n_out + 255) / 256;
  SpanGuard span(ctx, "fri_fold_line");
  switch (fold_step) {
    case 1: k_fold_line<1><<<grid, 256, 0, ctx->stream>>>(io, tw->itw, tw->root_log_size, log_size, al, d_al, leaves); break;
    case 2: k_fold_line<2><<<grid, 256, 0, ctx->stream>>>(io, tw->itw, tw->root_log_size, log_size, al, d_al, leaves); break;
    case 3: k_fold_line<3><<<grid, 256, 0, ctx->stream>>>(io, tw->itw, tw->root_log_size, log_size, al, d_al, leaves); break;
    default: k_fold_line<4><<<grid, 256, 0, ctx->stream>>>(io, tw->itw, tw->root_log_size, log_size, al, d_al, leaves); break;
  }
  return sh_launch_check(ctx, "fold_line");
}

static int fold_circle_launch(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const src[4], uint32_t log_size,
                              sh::qm31 alpha, const uint32_t* d_al, uint32_t* const dst[4],
                              uint32_t* leaves = nullptr) {
  SH_REQUIRE(ctx, log_size >= 1 && log_size <= 30, "circle domain log size %u out of range", log_size);
  SH_REQUIRE(ctx, sh::is_doubling_of(sh::subgroup_gen(log_size + 1), log_size - 1, tw->root_initial_index,
                                     tw->root_log_size),
             "circle domain of log size %u is not in the twiddle tower", log_size);
  uint32_t c0 = 0, c1 = 0;
  if (log_size <= 2) {
    // CircleDomain(half_odds(m - 1)).at(bitrev(2i, m)).y^-1 for i = 0, 1
    uint32_t half_init = sh::subgroup_gen(log_size + 1);
    uint32_t half_step = sh::subgroup_gen(log_size - 1);
    c0 = sh::minv(sh::index_to_point_host(half_init).y);
    if (log_size == 2) c1 = sh::minv(sh::index_to_point_host(sh::idx_reduce((uint64_t)half_init + half_step)).y);
  }
  Cols4 io;
  for (int c = 0; c < 4; c++) {
    io.s[c] = src[c];
    io.d[c] = dst[c];
  }
  uint32_t n_out = 1u << (log_size - 1);
  SpanGuard span(ctx, "fri_fold_circle");
  k_fold_circle<<<(n_out + 255) / 256, 256, 0, ctx->stream>>>(io, tw->itw, tw->root_log_size, log_size, alpha, d_al,
                                                             c0, c1, leaves);
  return sh_launch_check(ctx, "fold_circle_into_line");
}

extern "C" int sh_fold_line(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const src[4],
                            uint32_t line_initial_index, uint32_t log_size, const uint32_t* alphas,
                            uint32_t fold_step, uint32_t* const dst[4]) {
  if (!ctx || !tw || !src || !dst || !alphas) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, fold_step >= 1 && fold_step <= 4, "fold_step %u not in [1, 4]", fold_step);
  Alphas al;
  for (uint32_t f = 0; f < fold_step; f++) al.a[f] = sh::qload(alphas + 4 * f);
  return fold_line_launch(ctx, tw, src, line_initial_index, log_size, al, nullptr, fold_step, dst);
}

extern "C" int sh_fold_circle_into_line(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const src[4],
                                        uint32_t log_size, const uint32_t alpha[4], uint32_t* const dst[4]) {
  if (!ctx || !tw || !src || !dst || !alpha) return SH_ERR_INVALID;
  return fold_circle_launch(ctx, tw, src, log_size, sh::qload(alpha), nullptr, dst);
}

// The same folds with the alphas read on the device (4 words per fold at d_alphas), for drivers
// that draw them with a device-side channel and so never wait for the host between layers.
// leaves (optional, fold_step 1): also the leaf hash of every output value (32 bytes each), the
// leaf layer of the next FRI layer's tree.
int sh_fold_line_dev(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const src[4], uint32_t line_initial_index,
                     uint32_t log_size, const uint32_t* d_alphas, uint32_t fold_step, uint32_t* const dst[4],
                     uint32_t* leaves) {
  return fold_line_launch(ctx, tw, src, line_initial_index, log_size, Alphas{}, d_alphas, fold_step, dst, leaves);
}
int sh_fold_circle_dev(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const src[4], uint32_t log_size,
                       const uint32_t* d_alpha, uint32_t* const dst[4], uint32_t* leaves) {
  return fold_circle_launch(ctx, tw, src, log_size, sh::qzero(), d_alpha, dst, leaves);
}

extern "C" int sh_fri_decompose(sh_ctx* ctx, const uint32_t* const src[4], uint32_t log_size,
                                uint32_t* const dst[4], uint32_t lambda[4]) {
  if (!ctx || !src || !dst || !lambda) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, log_size >= 1 && log_size <= 30, "log size %u out of range", log_size);
  Cols4 io;
  for (int c = 0; c < 4; c++) {
    io.s[c] = src[c];
    io.d[c] = dst[c];
  }
  uint32_t half = 1u << (log_size - 1);
  uint32_t nblk = (half + 255) / 256;
  if (nblk > 256) nblk = 256;
  void* scr;
  SH_TRY(sh_scratch_get(ctx, SCR_MISC, 4ull * 4 * (2 * nblk + 1), &scr));
  uint32_t* partial = (uint32_t*)scr;
  uint32_t* lam = partial + 4 * 2 * nblk;
  SpanGuard span(ctx, "fri_decompose");
  k_half_sums<<<dim3(nblk, 2), 256, 0, ctx->stream>>>(io, log_size, partial);
  k_decompose_finish<<<1, 1, 0, ctx->stream>>>(partial, nblk, sh::minv(1u << log_size), lam);
  k_decompose_apply<<<((1u << log_size) + 255) / 256, 256, 0, ctx->stream>>>(io, log_size, lam);
  SH_TRY(sh_launch_check(ctx, "decompose"));
  SH_CHECK_HIP(ctx, hipMemcpyAsync(lambda, lam, 16, hipMemcpyDeviceToHost, ctx->stream));
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SH_OK;
}
