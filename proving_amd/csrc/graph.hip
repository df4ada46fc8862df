// HIP-graph capture of the batched trace commitment (lifted or mixed-degree tree).
//
// A captured commit replays every launch of sh_commit_trace (both streams,
// with their fork/join dependencies) as one graph launch, removing per-launch
// host overhead.  Kernel arguments are baked into the graph, so everything the
// kernels dereference must stay where it was at capture: the graph owns its
// own staging arena (the device tables of column pointers), its own scratch
// buffers and its own device blocks (every pooled allocation made while it is
// open).  They are swapped into the context for an eager warm-up run (which
// stages the tables, sizes the scratch and creates the blocks) and for the
// capture itself, then swapped back, so later eager calls never touch them.
// The same warm-up/capture pair captures any sequence of entry points that
// only enqueue work (sh_graph_warmup_begin ... sh_graph_capture_end): the
// prover captures a whole proof's device work this way.  Nothing in the
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
  sh_arena arena;          // device blocks the captured work allocates (pooled allocations)
  int phase = 0;           // 0 idle, 1 warm-up, 2 capture: the graph's state is in the context
  bool saved_profile = false;
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
hipError_t sync_all(sh_ctx* ctx) {
  hipError_t e = hipStreamSynchronize(ctx->main_stream);
  for (auto a : ctx->aux)
    if (a && e == hipSuccess) e = hipStreamSynchronize(a);
  return e;
}
// The graph's state goes into the context (or comes back out): its staging ring, scratch
// buffers and arena serve every call in between; profiling spans are off meanwhile.
void enter(sh_ctx* ctx, sh_graph* g, int phase) {
  swap_state(ctx, g);
  ctx->arena = &g->arena;
  g->arena.mode = phase;
  g->arena.cursor = 0;
  if (phase == 1) g->arena.seq.clear();
  g->saved_profile = ctx->profile;
  ctx->profile = false;
  g->phase = phase;
}
void leave(sh_ctx* ctx, sh_graph* g) {
  swap_state(ctx, g);
  ctx->arena = nullptr;
  g->arena.mode = 0;
  ctx->profile = g->saved_profile;
  g->phase = 0;
}
void release(sh_ctx* ctx, sh_graph* g) {
  if (g->phase) leave(ctx, g);
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  for (auto& s : g->scratch)
    if (s.ptr) hipFree(s.ptr);
  if (g->d_ring) hipFree(g->d_ring);
  if (g->h_ring) hipHostFree(g->h_ring);
  for (auto& kv : g->stage_pending) hipEventDestroy(kv.second.ev);
  sh_arena_release(ctx, &g->arena);
  delete g;
}
}  // namespace

extern "C" int sh_graph_create(sh_ctx* ctx, sh_graph** out) {
  if (!ctx || !out) return SH_ERR_INVALID;
  *out = nullptr;
  sh_graph* g = new sh_graph();
  g->ring_size = 4ull << 20;
  if (hipHostMalloc((void**)&g->h_ring, g->ring_size, hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void**)&g->d_ring, g->ring_size) != hipSuccess) {
    (void)hipGetLastError();
    release(ctx, g);
    return sh_fail(ctx, SH_ERR_OOM, "graph arena allocation failed");
  }
  *out = g;
  return SH_OK;
}

extern "C" int sh_graph_warmup_begin(sh_ctx* ctx, sh_graph* g) {
  if (!ctx || !g) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, g->phase == 0 && !g->exec && !ctx->arena, "graph already captured or another graph is open");
  SH_REQUIRE(ctx, ctx->stream == ctx->main_stream, "graph work must run on the main stream");
  SH_CHECK_HIP(ctx, sync_all(ctx));  // the context's streams are idle before its state is swapped
  enter(ctx, g, 1);
  return SH_OK;
}

extern "C" int sh_graph_warmup_end(sh_ctx* ctx, sh_graph* g) {
  if (!ctx || !g) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, g->phase == 1, "no warm-up in progress");
  hipError_t e = sync_all(ctx);
  if (e == hipSuccess) sh_stage_all_complete(ctx);  // the capture reuses the tables with no event waits
  leave(ctx, g);
  if (e != hipSuccess) return sh_fail(ctx, SH_ERR_HIP, "graph warm-up failed: %s", hipGetErrorString(e));
  return SH_OK;
}

extern "C" int sh_graph_capture_begin(sh_ctx* ctx, sh_graph* g) {
  if (!ctx || !g) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, g->phase == 0 && !g->exec && !ctx->arena, "graph already captured or another graph is open");
  SH_REQUIRE(ctx, ctx->stream == ctx->main_stream, "graph capture must start on the main stream");
  SH_CHECK_HIP(ctx, sync_all(ctx));
  enter(ctx, g, 2);
  hipError_t e = hipStreamBeginCapture(ctx->main_stream, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    leave(ctx, g);
    return sh_fail(ctx, SH_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
  }
  return SH_OK;
}

extern "C" int sh_graph_capture_end(sh_ctx* ctx, sh_graph* g) {
  if (!ctx || !g) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, g->phase == 2, "no capture in progress");
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(ctx->main_stream, &graph);
  leave(ctx, g);
  if (e != hipSuccess || !graph) {
    if (graph) hipGraphDestroy(graph);
    (void)hipGetLastError();
    return sh_fail(ctx, SH_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  }
  g->graph = graph;
  e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    g->exec = nullptr;
    return sh_fail(ctx, SH_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  return SH_OK;
}

// Captures one batched commit (the lifted tree, or the mixed-degree tree when mixed) as a graph:
// an eager warm-up run (which stages the tables and sizes the scratch), then the capture.
static int commit_graph_create(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_evals,
                               uint32_t* const* d_coeffs, const uint32_t* log_sizes, uint64_t n, uint32_t log_blowup,
                               uint32_t* const* d_ldes, uint8_t* const* d_layers, bool mixed, sh_graph** out) {
  auto commit = mixed ? sh_commit_trace_mixed : sh_commit_trace;
  if (!ctx || !out) return SH_ERR_INVALID;
  *out = nullptr;
  sh_graph* g = nullptr;
  SH_TRY(sh_graph_create(ctx, &g));
  int rc = sh_graph_warmup_begin(ctx, g);
  if (rc == SH_OK) {
    rc = commit(ctx, tw, d_evals, d_coeffs, log_sizes, n, log_blowup, d_ldes, d_layers, nullptr);
    const int rc2 = sh_graph_warmup_end(ctx, g);
    if (rc == SH_OK) rc = rc2;
  }
  if (rc == SH_OK) rc = sh_graph_capture_begin(ctx, g);
  if (rc == SH_OK) {
    rc = commit(ctx, tw, d_evals, d_coeffs, log_sizes, n, log_blowup, d_ldes, d_layers, nullptr);
    const int rc2 = sh_graph_capture_end(ctx, g);
    if (rc == SH_OK) rc = rc2;
  }
  if (rc != SH_OK) {
    (void)sync_all(ctx);
    release(ctx, g);
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
  if (g->phase == 2) {  // a capture left open by an error: end it, dropping what was recorded
    hipGraph_t graph = nullptr;
    (void)hipStreamEndCapture(ctx->main_stream, &graph);
    if (graph) hipGraphDestroy(graph);
    (void)hipGetLastError();
  }
  const hipError_t e = sync_all(ctx);
  release(ctx, g);
  SH_CHECK_HIP(ctx, e);
  return SH_OK;
}
