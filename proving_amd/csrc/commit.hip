// Batched trace commitment: interpolate + LDE + lifted Merkle tree over many
// columns of uneven sizes (TreeBuilder::extend_evals followed by
// CommitmentTreeProver::new, crates/stwo/src/prover/pcs/mod.rs:325-334, 360-386).
//
// Scheduling: the FFT chains (interpolate, then LDE) of the size classes run
// back to back, smallest first, on an auxiliary stream; the leaf hashing runs
// on the main stream, one size group at a time in ascending size, each group
// waiting only for its own class's LDE.  A small group (few rows, many
// columns, long per-row hash chains, too few lanes to fill the device) thus
// runs beside the next class's FFT instead of after it.  FFT classes are not
// run beside each other: each already fills the device, and running two
// concurrently was measured slower than running them in turn.  The inner
// layers follow on the main stream.
#include <algorithm>

#include "common.h"

int sh_interpolate_extend_grouped(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* h_src,
                                  uint32_t* const* h_coef, uint32_t* const* h_ext, uint32_t m, uint32_t b,
                                  uint64_t n);
int sh_build_leaves_impl(sh_ctx* ctx, const uint32_t* const* h_cols, const uint32_t* log_sizes, uint64_t n,
                         uint32_t L, uint32_t* d_out, const hipEvent_t* group_ready);
int sh_build_layers_impl(sh_ctx* ctx, uint32_t L, uint32_t* const* layers);
int sh_merkle_mixed_impl(sh_ctx* ctx, const uint32_t* const* h_cols, const uint32_t* logs, uint64_t n,
                         uint8_t* const* d_layers, const hipEvent_t* level_ready);

static int commit_forked(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_evals,
                         uint32_t* const* d_coeffs, const uint32_t* log_sizes, uint64_t n, uint32_t log_blowup,
                         uint32_t* const* d_ldes, uint8_t* const* d_layers, const std::vector<uint32_t>& classes,
                         uint32_t L, bool mixed, hipStream_t fft_stream);

// The lifted tree takes the size classes ascending (its leaf groups are absorbed smallest first);
// the mixed-degree tree descending (it is built from the largest columns' level down).
static int commit_pipeline(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_evals,
                           uint32_t* const* d_coeffs, const uint32_t* log_sizes, uint64_t n, uint32_t log_blowup,
                           uint32_t* const* d_ldes, uint8_t* const* d_layers, uint8_t root[32], bool mixed) {
  if (!ctx || !tw || !d_layers || (n && (!d_evals || !d_coeffs || !log_sizes || !d_ldes)))
    return SH_ERR_INVALID;
  SH_REQUIRE(ctx, n > 0, "sh_commit_trace needs at least one column");
  SH_REQUIRE(ctx, ctx->stream == ctx->main_stream, "sh_commit_trace called from a forked stream");
  std::vector<uint32_t> classes;
  for (uint64_t i = 0; i < n; i++) classes.push_back(log_sizes[i]);
  std::sort(classes.begin(), classes.end());
  classes.erase(std::unique(classes.begin(), classes.end()), classes.end());
  const uint32_t L = classes.back() + log_blowup;
  if (mixed) std::reverse(classes.begin(), classes.end());

  // fork (several size classes: a class's tree work overlaps the next class's FFTs): the
  // auxiliary stream starts after the work already on the main stream.  A single class has
  // nothing to overlap and stays on the main stream.
  hipStream_t fft_stream = ctx->main_stream;
  if (classes.size() > 1) {
    SH_TRY(sh_aux_stream(ctx, &fft_stream));
    hipEvent_t fork = sh_sync_event(ctx, 0);
    SH_CHECK_HIP(ctx, hipEventRecord(fork, ctx->main_stream));
    SH_CHECK_HIP(ctx, hipStreamWaitEvent(fft_stream, fork, 0));
  }
  const int rc = commit_forked(ctx, tw, d_evals, d_coeffs, log_sizes, n, log_blowup, d_ldes, d_layers, classes, L,
                               mixed, fft_stream);
  if (rc != SH_OK) {
    // join: buffers freed after this call may be reused at once, so nothing enqueued on the
    // auxiliary stream may still run unordered after the main stream's work
    if (fft_stream != ctx->main_stream) {
      hipEvent_t join = sh_sync_event(ctx, 1 + classes.size());
      if (hipEventRecord(join, fft_stream) == hipSuccess) (void)hipStreamWaitEvent(ctx->main_stream, join, 0);
    }
    return rc;
  }
  if (root) {
    SH_CHECK_HIP(ctx, hipMemcpyAsync(root, d_layers[0], 32, hipMemcpyDeviceToHost, ctx->stream));
    SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return SH_OK;
}

extern "C" int sh_commit_trace(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_evals,
                               uint32_t* const* d_coeffs, const uint32_t* log_sizes, uint64_t n,
                               uint32_t log_blowup, uint32_t* const* d_ldes, uint8_t* const* d_layers,
                               uint8_t root[32]) {
  return commit_pipeline(ctx, tw, d_evals, d_coeffs, log_sizes, n, log_blowup, d_ldes, d_layers, root, false);
}

extern "C" int sh_commit_trace_mixed(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_evals,
                                     uint32_t* const* d_coeffs, const uint32_t* log_sizes, uint64_t n,
                                     uint32_t log_blowup, uint32_t* const* d_ldes, uint8_t* const* d_layers,
                                     uint8_t root[32]) {
  return commit_pipeline(ctx, tw, d_evals, d_coeffs, log_sizes, n, log_blowup, d_ldes, d_layers, root, true);
}

// The part of sh_commit_trace after the fork: FFT classes on fft_stream, the tree on the main stream.
// Every leaf group (lifted) or column level (mixed) waits for its class's event, and the last
// class's event is waited on by the tree, so on success the auxiliary stream has joined back.
static int commit_forked(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_evals,
                         uint32_t* const* d_coeffs, const uint32_t* log_sizes, uint64_t n, uint32_t log_blowup,
                         uint32_t* const* d_ldes, uint8_t* const* d_layers, const std::vector<uint32_t>& classes,
                         uint32_t L, bool mixed, hipStream_t fft_stream) {
  std::vector<hipEvent_t> ready(classes.size());
  for (size_t k = 0; k < classes.size(); k++) {
    const uint32_t m = classes[k];
    std::vector<const uint32_t*> ev;
    std::vector<uint32_t*> co, ld;
    for (uint64_t i = 0; i < n; i++)
      if (log_sizes[i] == m) {
        ev.push_back(d_evals[i]);
        co.push_back(d_coeffs[i]);
        ld.push_back(d_ldes[i]);
      }
    StreamScope scope(ctx, fft_stream);
    SH_TRY(sh_interpolate_extend_grouped(ctx, tw, ev.data(), co.data(), ld.data(), m, log_blowup, ev.size()));
    ready[k] = sh_sync_event(ctx, 1 + k);
    SH_CHECK_HIP(ctx, hipEventRecord(ready[k], ctx->stream));
  }
  if (mixed) {
    // columns descending by size (stable); level k of the tree's column levels waits for class k
    std::vector<uint64_t> order(n);
    for (uint64_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint64_t x, uint64_t y) { return log_sizes[x] > log_sizes[y]; });
    std::vector<const uint32_t*> cols(n);
    std::vector<uint32_t> logs(n);
    for (uint64_t i = 0; i < n; i++) {
      cols[i] = d_ldes[order[i]];
      logs[i] = log_sizes[order[i]] + log_blowup;
    }
    return sh_merkle_mixed_impl(ctx, cols.data(), logs.data(), n, d_layers, ready.data());
  }
  // leaves on the main stream; group k (LDE log classes[k] + b) waits for class k
  std::vector<uint64_t> order(n);
  for (uint64_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint64_t x, uint64_t y) { return log_sizes[x] < log_sizes[y]; });
  std::vector<const uint32_t*> cols(n);
  std::vector<uint32_t> logs(n);
  for (uint64_t i = 0; i < n; i++) {
    cols[i] = d_ldes[order[i]];
    logs[i] = log_sizes[order[i]] + log_blowup;
  }
  SH_TRY(sh_build_leaves_impl(ctx, cols.data(), logs.data(), n, L, (uint32_t*)d_layers[L], ready.data()));
  return sh_build_layers_impl(ctx, L, (uint32_t* const*)d_layers);
}
