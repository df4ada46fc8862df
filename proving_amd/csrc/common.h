// Internal definitions shared by the translation units of libstwo_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/stwo_hip.h"
#include "m31.h"

// One context per host thread: a device, its stream, a staging ring for the
// small per-call tables (column pointer lists) and reusable scratch buffers.
struct sh_prof_rec {
  const char* cls;
  hipEvent_t start, stop;
};

struct sh_scratch {
  void* ptr = nullptr;
  uint64_t bytes = 0;
};

// The device blocks of one captured graph (graph.hip): while sh_ctx::arena points at it, pooled
// allocations are served from it, and a block freed at any later time goes back to its own
// arena, never to the context's pool, so nothing outside the graph reuses memory its replays
// write.  The warm-up run records the block each allocation received; the capture that follows
// (the warm-up's blocks freed again) hands out the same blocks in the same order, so every
// address the captured work bakes in, and every staged table of addresses, is the warm-up's.
struct sh_arena {
  std::unordered_map<uint64_t, std::vector<void*>> free;          // size class -> free blocks
  std::unordered_map<void*, std::pair<uint64_t, bool>> blocks;    // block -> (size class, live)
  std::vector<void*> seq;  // the blocks the warm-up's allocations received, in order
  size_t cursor = 0;       // the capture's next allocation receives seq[cursor]
  int mode = 0;            // 1: warm-up (records seq), 2: capture (replays seq)
};

struct sh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;     // the stream every entry point orders its work on
  hipStream_t main_stream = nullptr;
  // fork/join helper inside one call (the batched commit of several column sizes runs its FFT
  // chain on it), created on first use: a context holds only the streams it uses, so several
  // contexts in one process stay within the device's few hardware queues
  hipStream_t aux[1] = {nullptr};
  std::string last_error;
  // pinned host arena + device mirror for per-call tables; identical tables are
  // reused (content-addressed), so steady-state calls issue no copies
  uint8_t* h_ring = nullptr;
  uint8_t* d_ring = nullptr;
  uint64_t ring_size = 0, ring_off = 0;
  std::unordered_map<std::string, uint64_t> table_cache;
  // pinned host buffer for the protocol drivers' bulk tables and read-backs (grown on demand;
  // every user waits for its copies before returning)
  uint8_t* h_xfer = nullptr;
  uint64_t h_xfer_size = 0;
  // staged tables whose copy is not known to be complete: ring offset -> (event after the copy,
  // stream that copied); a reuse from another stream waits on the event instead of the host
  struct StageEv {
    hipEvent_t ev;
    hipStream_t stream;
  };
  std::unordered_map<uint64_t, StageEv> stage_pending;
  std::vector<hipEvent_t> stage_ev_free;
  // profiling (hipEvent pairs around each kernel launch, tagged by span class)
  bool profile = false;
  std::vector<sh_prof_rec> prof;
  std::vector<hipEvent_t> event_pool;
  // named scratch buffers
  sh_scratch scratch[8];
  hipEvent_t timer_start = nullptr, timer_stop = nullptr;
  std::vector<hipEvent_t> sync_events;  // fork/join events (no timing)
  // caching device allocator (BaseColumnPool's role, crates/stwo/src/prover/mempool.rs:14-58):
  // freed blocks are kept per rounded size and handed out again; reuse is safe
  // without a sync because every entry point orders its work on the main
  // stream (auxiliary streams are joined back before a call returns)
  std::mutex pool_mu;  // sh_free may come from another host thread (a buffer handed between contexts)
  std::unordered_map<uint64_t, std::vector<void*>> pool;
  std::unordered_map<void*, uint64_t> pool_live;
  uint64_t pool_bytes = 0;
  sh_arena* arena = nullptr;                        // set during a graph's warm-up and capture
  std::unordered_map<void*, sh_arena*> arena_of;    // every live graph's blocks
  // FFT pass shape per transform log size (see fft.hip)
  struct FftTune {
    uint8_t block_log, stride_layers;
    uint8_t r_block, r_stride;  // rows bits per lane of the quad-layout passes (fft2.hip)
  } fft_tune[32];
  bool fft2 = true;  // quad-layout pass kernels where a specialisation exists
  // interpolation + extension of the same columns: 0 separate, 1 the interpolation's last strided
  // pass and the extension's first run as one pass where their plans share those layers, 2 the
  // extension also takes the interpolation's layer split so that they always do
  uint8_t fft_fuse = 1;
};

// The FRI channel step run by the launch that computes a tree's root (merkle.hip): mix the root
// into the device channel ch (digest, n_draws) and draw alpha, writing its k squared powers to
// alphas (4 words each) and a copy of the root to root_out.  ch == nullptr: no step.
struct sh_root_step {
  uint32_t* ch;
  uint32_t* alphas;
  uint32_t* root_out;
  uint32_t k;
};
// Layers L-1 .. 0 of a tree from layer L; with st, the launch that computes the root also runs
// the channel step (a tree of one leaf, L == 0, has no such launch: the caller steps itself).
int sh_build_layers_step_impl(sh_ctx* ctx, uint32_t L, uint32_t* const* layers, const sh_root_step* st);

// Every staged copy is complete (the caller has synchronised the context's streams).
void sh_stage_all_complete(sh_ctx* ctx);

struct sh_twiddles {
  uint32_t root_initial_index;
  uint32_t root_log_size;  // K: buffers hold 2^K words
  uint32_t* tw;            // device
  uint32_t* itw;           // device
};

enum {
  SCR_MERKLE_STATE_A = 0,
  SCR_MERKLE_STATE_B = 1,
  SCR_MERKLE_TMP = 2,
  SCR_FFT_TMP = 3,
  SCR_MISC = 4,
  SCR_QUOT = 5,
  SCR_GRIND = 6,
  SCR_EVAL = 7,
};

int sh_fail(sh_ctx* ctx, int code, const char* fmt, ...);
#define SH_CHECK_HIP(ctx, expr)                                                   \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess)                                                        \
      return sh_fail(ctx, SH_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(_e)); \
  } while (0)
#define SH_REQUIRE(ctx, cond, ...)                       \
  do {                                                   \
    if (!(cond)) return sh_fail(ctx, SH_ERR_INVALID, __VA_ARGS__); \
  } while (0)
#define SH_TRY(expr)            \
  do {                          \
    int _r = (expr);            \
    if (_r != SH_OK) return _r; \
  } while (0)

// Copies a small host table into the device arena (stream-ordered, once per
// distinct content) and returns its device address.
int sh_stage(sh_ctx* ctx, const void* host, uint64_t bytes, void** dev);
// The i-th fork/join event of the context (created on first use).
hipEvent_t sh_sync_event(sh_ctx* ctx, size_t i);
// The context's pinned transfer buffer, at least `bytes` long (see sh_ctx::h_xfer).
int sh_host_xfer(sh_ctx* ctx, uint64_t bytes, uint8_t** out);
// Ensures scratch buffer `slot` holds at least `bytes` and returns it.
int sh_scratch_get(sh_ctx* ctx, int slot, uint64_t bytes, void** out);
// The context's auxiliary stream (created on first use).
int sh_aux_stream(sh_ctx* ctx, hipStream_t* out);
// Pooled device allocation for buffers that outlive one launch (see sh_ctx::pool).
int sh_dev_alloc(sh_ctx* ctx, uint64_t bytes, void** out);
void sh_dev_free(sh_ctx* ctx, void* p);
// Frees every block of an arena (the caller has synchronised the context's streams).
void sh_arena_release(sh_ctx* ctx, sh_arena* a);
// Profiling spans around launches.
void sh_prof_begin(sh_ctx* ctx, const char* cls, sh_prof_rec* rec);
void sh_prof_end(sh_ctx* ctx, sh_prof_rec* rec);
int sh_launch_check(sh_ctx* ctx, const char* what);

// Profiling span: RAII pair of events on the context stream.
struct SpanGuard {
  sh_ctx* ctx;
  sh_prof_rec rec;
  SpanGuard(sh_ctx* c, const char* cls) : ctx(c) { sh_prof_begin(ctx, cls, &rec); }
  ~SpanGuard() { sh_prof_end(ctx, &rec); }
};

// Host helpers for circle-group indices (crates/stwo/src/core/circle.rs).
namespace sh {
// CirclePointIndex::to_point: M31_CIRCLE_GEN * index by double-and-add.
mpt index_to_point_host(uint32_t index);
// Table of M31_CIRCLE_GEN * 2^b for b = 0..30, for device-side index_to_point.
const mpt* gen_pow2_table_host();
// True if coset (init, log) is `root` doubled (root_log - log) times (Coset::is_doubling_of).
bool is_doubling_of(uint32_t init, uint32_t log, uint32_t root_init, uint32_t root_log);
}  // namespace sh

// Device-side index -> point from the 31-entry power table.
__device__ __forceinline__ sh::mpt sh_index_to_point(uint32_t idx, const sh::mpt* gen_pow2) {
  sh::mpt r = {1, 0};
#pragma unroll 1
  for (int b = 0; b < 31; b++)
    if ((idx >> b) & 1u) r = sh::padd(r, gen_pow2[b]);
  return r;
}

// Points of a coset C = init + <step> by table lookup: C.at(j) = init + sum_l T_l[digit_l(j)]
// with 11-bit digits, T_l[k] = (k << 11l) * step.  Built on the host, staged per call.
struct sh_point_tab {
  const sh::mpt* t;  // levels * 2048 points (device)
  uint32_t levels;
  sh::mpt init;
};
int sh_point_tab_make(sh_ctx* ctx, uint32_t init_index, uint32_t log_size, sh_point_tab* out);

__device__ __forceinline__ sh::mpt sh_tab_point(const sh_point_tab& pt, uint32_t j) {
  sh::mpt p = pt.init;
  for (uint32_t l = 0; l < pt.levels; l++) p = sh::padd(p, pt.t[l * 2048 + ((j >> (11 * l)) & 2047)]);
  return p;
}
// CircleDomain(half coset).at(j) (crates/stwo/src/core/poly/circle/domain.rs:50-60): the second
// half is the conjugate of the first.
__device__ __forceinline__ sh::mpt sh_circle_domain_point(const sh_point_tab& half, uint32_t half_log,
                                                          uint32_t j) {
  uint32_t hs = 1u << half_log;
  if (j < hs) return sh_tab_point(half, j);
  sh::mpt p = sh_tab_point(half, j - hs);
  return {p.x, sh::mneg(p.y)};
}

// Runs the enclosed launches on another stream of the context.
struct StreamScope {
  sh_ctx* ctx;
  hipStream_t saved;
  StreamScope(sh_ctx* c, hipStream_t s) : ctx(c), saved(c->stream) { c->stream = s; }
  ~StreamScope() { ctx->stream = saved; }
};
