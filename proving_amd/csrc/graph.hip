// HIP-graph capture of the batched trace commitment (lifted or mixed-degree tree).
//
// A captured commit replays every launch of sh_commit_trace (both streams,
// with their fork/join dependencies) as one graph launch, removing per-launch
// host overhead.  Kernel arguments are baked into the graph, so everything the
// kernels dereference must stay where it was at capture: the graph owns its
// own staging arena (the device tables of column pointers) and its own scratch
// buffers.  Both are swapped into the context for an eager warm-up run (which
// stages the tables and sizes the scratch) and for the capture itself, then
// swapped back, so later eager calls never touch them.  Nothing in the
// captured work uses hipMemset.
#include <string.h>

#include <algorithm>

#include "common.h"

struct sh_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  sh_scratch scratch[8];
  uint8_t* h_ring = nullptr;
  uint8_t* d_ring = nullptr;
  uint64_t ring_size = 0, ring_off = 0;
  std::unordered_map<std::string, uint64_t> table_cache;
  std::unordered_map<uint64_t, sh_ctx::StageEv> stage_pending;
};

namespace {
void swap_state(sh_ctx* ctx, sh_graph* g) {
  for (int i = 0; i < 8; i++) std::swap(ctx->scratch[i], g->scratch[i]);
  std::swap(ctx->h_ring, g->h_ring);
  std::swap(ctx->d_ring, g->d_ring);
  std::swap(ctx->ring_size, g->ring_size);
  std::swap(ctx->ring_off, g->ring_off);
  std::swap(ctx->table_cache, g->table_cache);
  std::swap(ctx->stage_pending, g->stage_pending);
}
void release(sh_graph* g) {
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  for (auto& s : g->scratch)
    if (s.ptr) hipFree(s.ptr);
  if (g->d_ring) hipFree(g->d_ring);
  if (g->h_ring) hipHostFree(g->h_ring);
  for (auto& kv : g->stage_pending) hipEventDestroy(kv.second.ev);
  delete g;
}
}  // namespace

// Captures one batched commit (the lifted tree, or the mixed-degree tree when mixed) as a graph.
static int commit_graph_create(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_evals,
                               uint32_t* const* d_coeffs, const uint32_t* log_sizes, uint64_t n, uint32_t log_blowup,
                               uint32_t* const* d_ldes, uint8_t* const* d_layers, bool mixed, sh_graph** out) {
  auto commit = mixed ? sh_commit_trace_mixed : sh_commit_trace;
  if (!ctx || !out) return SH_ERR_INVALID;
  *out = nullptr;
  SH_REQUIRE(ctx, ctx->stream == ctx->main_stream, "graph capture must start on the main stream");
  sh_graph* g = new sh_graph();
  g->ring_size = 1ull << 20;
  if (hipHostMalloc((void**)&g->h_ring, g->ring_size, hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void**)&g->d_ring, g->ring_size) != hipSuccess) {
    release(g);
    return sh_fail(ctx, SH_ERR_OOM, "graph arena allocation failed");
  }
  // the context's streams must be idle before its state is swapped
  {
    hipError_t e = hipStreamSynchronize(ctx->main_stream);
    for (auto a : ctx->aux)
      if (a && e == hipSuccess) e = hipStreamSynchronize(a);
    if (e != hipSuccess) {
      release(g);
      return sh_fail(ctx, SH_ERR_HIP, "stream synchronisation before capture failed: %s", hipGetErrorString(e));
    }
  }
  const bool prof = ctx->profile;
  ctx->profile = false;
  swap_state(ctx, g);
  int rc = commit(ctx, tw, d_evals, d_coeffs, log_sizes, n, log_blowup, d_ldes, d_layers, nullptr);
  if (rc == SH_OK) {
    hipError_t e = hipStreamSynchronize(ctx->main_stream);
    for (auto a : ctx->aux)
      if (a && e == hipSuccess) e = hipStreamSynchronize(a);
    if (e != hipSuccess) rc = sh_fail(ctx, SH_ERR_HIP, "warm-up failed: %s", hipGetErrorString(e));
    sh_stage_all_complete(ctx);  // the capture below only reuses tables, with no event waits
  }
  if (rc == SH_OK) {
    hipError_t e = hipStreamBeginCapture(ctx->main_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      rc = sh_fail(ctx, SH_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    } else {
      rc = commit(ctx, tw, d_evals, d_coeffs, log_sizes, n, log_blowup, d_ldes, d_layers, nullptr);
      hipGraph_t graph = nullptr;
      hipError_t e2 = hipStreamEndCapture(ctx->main_stream, &graph);
      if (rc == SH_OK && e2 != hipSuccess)
        rc = sh_fail(ctx, SH_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
      g->graph = graph;
      if (rc == SH_OK) {
        e2 = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        if (e2 != hipSuccess) rc = sh_fail(ctx, SH_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e2));
      }
    }
  }
  swap_state(ctx, g);
  ctx->profile = prof;
  if (rc != SH_OK) {
    release(g);
    return rc;
  }
  *out = g;
  return SH_OK;
}

extern "C" int sh_commit_trace_graph_create(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_evals,
                                            uint32_t* const* d_coeffs, const uint32_t* log_sizes, uint64_t n,
                                            uint32_t log_blowup, uint32_t* const* d_ldes,
                                            uint8_t* const* d_layers, sh_graph** out) {
  return commit_graph_create(ctx, tw, d_evals, d_coeffs, log_sizes, n, log_blowup, d_ldes, d_layers, false, out);
}

extern "C" int sh_commit_trace_mixed_graph_create(sh_ctx* ctx, const sh_twiddles* tw,
                                                  const uint32_t* const* d_evals, uint32_t* const* d_coeffs,
                                                  const uint32_t* log_sizes, uint64_t n, uint32_t log_blowup,
                                                  uint32_t* const* d_ldes, uint8_t* const* d_layers,
                                                  sh_graph** out) {
  return commit_graph_create(ctx, tw, d_evals, d_coeffs, log_sizes, n, log_blowup, d_ldes, d_layers, true, out);
}

extern "C" int sh_graph_launch(sh_ctx* ctx, sh_graph* g) {
  if (!ctx || !g || !g->exec) return SH_ERR_INVALID;
  SH_CHECK_HIP(ctx, hipGraphLaunch(g->exec, ctx->main_stream));
  return SH_OK;
}

extern "C" int sh_graph_destroy(sh_ctx* ctx, sh_graph* g) {
  if (!ctx) return SH_ERR_INVALID;
  if (!g) return SH_OK;
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->main_stream));
  release(g);
  return SH_OK;
}
