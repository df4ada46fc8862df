// Batched trace commitment: interpolate + LDE + lifted Merkle tree over many
// columns of uneven sizes (TreeBuilder::extend_evals followed by
// CommitmentTreeProver::new, crates/stwo/src/prover/pcs/mod.rs:325-334, 360-386).
#include <algorithm>

#include "common.h"

int sh_interpolate_grouped(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* h_src,
                           uint32_t* const* h_dst, const uint32_t* log_sizes, uint64_t n);
int sh_evaluate_grouped(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* h_src,
                        const uint32_t* coeff_logs, const uint32_t* domain_logs, uint32_t* const* h_dst,
                        uint64_t n);
int sh_build_leaves_impl(sh_ctx* ctx, const uint32_t* const* h_cols, const uint32_t* log_sizes, uint64_t n,
                         uint32_t L, uint32_t* d_out);
int sh_build_layers_impl(sh_ctx* ctx, uint32_t L, uint32_t* const* layers);

extern "C" int sh_commit_trace(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_evals,
                               uint32_t* const* d_coeffs, const uint32_t* log_sizes, uint64_t n,
                               uint32_t log_blowup, uint32_t* const* d_ldes, uint8_t* const* d_layers,
                               uint8_t root[32]) {
  if (!ctx || !tw || !d_layers || (n && (!d_evals || !d_coeffs || !log_sizes || !d_ldes)))
    return SH_ERR_INVALID;
  SH_REQUIRE(ctx, n > 0, "sh_commit_trace needs at least one column");
  SH_TRY(sh_interpolate_grouped(ctx, tw, d_evals, d_coeffs, log_sizes, n));
  std::vector<uint32_t> dl(n);
  uint32_t L = 0;
  for (uint64_t i = 0; i < n; i++) {
    dl[i] = log_sizes[i] + log_blowup;
    L = std::max(L, dl[i]);
  }
  SH_TRY(sh_evaluate_grouped(ctx, tw, (const uint32_t* const*)d_coeffs, log_sizes, dl.data(), d_ldes, n));
  std::vector<uint64_t> order(n);
  for (uint64_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) { return dl[x] < dl[y]; });
  std::vector<const uint32_t*> cols(n);
  std::vector<uint32_t> logs(n);
  for (uint64_t i = 0; i < n; i++) {
    cols[i] = d_ldes[order[i]];
    logs[i] = dl[order[i]];
  }
  SH_TRY(sh_build_leaves_impl(ctx, cols.data(), logs.data(), n, L, (uint32_t*)d_layers[L]));
  SH_TRY(sh_build_layers_impl(ctx, L, (uint32_t* const*)d_layers));
  if (root) {
    SH_CHECK_HIP(ctx, hipMemcpyAsync(root, d_layers[0], 32, hipMemcpyDeviceToHost, ctx->stream));
    SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return SH_OK;
}
