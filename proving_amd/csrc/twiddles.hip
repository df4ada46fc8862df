// Twiddle tree (PolyOps::precompute_twiddles).
//
// Layout, as CpuBackend builds it (crates/stwo/src/prover/backend/cpu/circle.rs:
// 274-285): for root coset C of log size K, layer l (l = 0..K-1) holds the x
// coordinates of the first half of C.double()^l in bit-reversed order, at
// offset 2^K - 2^(K-l); one padding word (1) completes 2^K words.  itwiddles are
// the element-wise inverses.  The SIMD backend stores 2x these values for its
// multiply (crates/stwo/src/prover/backend/simd/circle.rs:452-459); this
// backend keeps canonical words.
//
// One lane per word: the point index is computed from the coset, converted to a
// point with at most 31 additions of precomputed G * 2^b, and inverted by
// Fermat (x^(P-2)).
#include "common.h"

struct GenTable {
  sh::mpt p[31];
};

__global__ void k_twiddles(uint32_t* tw, uint32_t* itw, uint32_t root_init, uint32_t root_step,
                           uint32_t K, GenTable gen) {
  uint64_t n = 1ull << K;
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  if (e == n - 1) {
    tw[e] = 1;
    itw[e] = 1;
    return;
  }
  uint64_t r = n - e;  // in (2^(K-l-1), 2^(K-l)]
  uint32_t clog = 64 - __clzll(r - 1);  // ceil(log2 r) = K - l
  uint32_t l = K - clog;
  uint64_t off = n - (1ull << clog);
  uint32_t j = (uint32_t)(e - off);
  uint32_t half_log = clog - 1;
  uint32_t init_l = sh::idx_reduce((uint64_t)root_init << l);
  uint32_t step_l = sh::idx_reduce((uint64_t)root_step << l);
  uint32_t idx = sh::idx_reduce((uint64_t)init_l + (uint64_t)step_l * sh::bitrev(j, half_log));
  sh::mpt p = sh_index_to_point(idx, gen.p);
  tw[e] = p.x;
  itw[e] = sh::minv(p.x);
}

extern "C" int sh_twiddles_create(sh_ctx* ctx, uint32_t root_initial_index, uint32_t root_log_size,
                                  sh_twiddles** out) {
  if (!ctx || !out) return SH_ERR_INVALID;
  *out = nullptr;
  SH_REQUIRE(ctx, root_log_size <= 30, "root coset log size %u exceeds 30", root_log_size);
  sh_twiddles* t = new sh_twiddles();
  t->root_initial_index = root_initial_index & 0x7fffffffu;
  t->root_log_size = root_log_size;
  uint64_t n = 1ull << root_log_size;
  if (hipMalloc(&t->tw, n * 4) != hipSuccess || hipMalloc(&t->itw, n * 4) != hipSuccess) {
    if (t->tw) hipFree(t->tw);
    delete t;
    return sh_fail(ctx, SH_ERR_OOM, "twiddle allocation failed");
  }
  GenTable g;
  const sh::mpt* tab = sh::gen_pow2_table_host();
  for (int b = 0; b < 31; b++) g.p[b] = tab[b];
  {
    SpanGuard span(ctx, "twiddles");
    k_twiddles<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(
        t->tw, t->itw, t->root_initial_index, sh::subgroup_gen(root_log_size), root_log_size, g);
  }
  int rc = sh_launch_check(ctx, "twiddles");
  if (rc != SH_OK) {
    hipFree(t->tw);
    hipFree(t->itw);
    delete t;
    return rc;
  }
  *out = t;
  return SH_OK;
}

extern "C" int sh_twiddles_destroy(sh_ctx* ctx, sh_twiddles* t) {
  if (!ctx) return SH_ERR_INVALID;
  if (!t) return SH_OK;
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  hipFree(t->tw);
  hipFree(t->itw);
  delete t;
  return SH_OK;
}

extern "C" int sh_twiddles_download(sh_ctx* ctx, const sh_twiddles* t, uint32_t* h_tw, uint32_t* h_itw) {
  if (!ctx || !t) return SH_ERR_INVALID;
  uint64_t bytes = 4ull << t->root_log_size;
  if (h_tw) SH_CHECK_HIP(ctx, hipMemcpyAsync(h_tw, t->tw, bytes, hipMemcpyDeviceToHost, ctx->stream));
  if (h_itw) SH_CHECK_HIP(ctx, hipMemcpyAsync(h_itw, t->itw, bytes, hipMemcpyDeviceToHost, ctx->stream));
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SH_OK;
}
