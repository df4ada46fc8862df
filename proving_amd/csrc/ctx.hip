// Context, device memory, staging ring, profiling spans and column utilities.
#include <stdarg.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>

#include "common.h"

static const uint64_t kRingBytes = 16ull << 20;

int sh_static_scratch_init(sh_ctx* ctx);  // quotients.hip

int sh_fail(sh_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->last_error = buf;
  }
  return code;
}

int sh_launch_check(sh_ctx* ctx, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sh_fail(ctx, SH_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
  return SH_OK;
}

extern "C" int sh_ctx_create(int device, sh_ctx** out) {
  if (!out) return SH_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return SH_ERR_HIP;
  if (device < 0 || device >= count) return SH_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return SH_ERR_HIP;
  sh_ctx* ctx = new sh_ctx();
  ctx->device = device;
  bool ok = hipStreamCreateWithFlags(&ctx->main_stream, hipStreamNonBlocking) == hipSuccess;
  ctx->stream = ctx->main_stream;
  if (!ok || hipHostMalloc((void**)&ctx->h_ring, kRingBytes, hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void**)&ctx->d_ring, kRingBytes) != hipSuccess ||
      hipEventCreate(&ctx->timer_start) != hipSuccess || hipEventCreate(&ctx->timer_stop) != hipSuccess) {
    sh_ctx_destroy(ctx);
    return SH_ERR_HIP;
  }
  ctx->ring_size = kRingBytes;
  {
    // SH_FFT_IMPL=0 keeps every transform on the generic pass kernel (for A/B timing)
    const char* impl = getenv("SH_FFT_IMPL");
    ctx->fft2 = !(impl && impl[0] == '0');
  }
  if (sh_static_scratch_init(ctx) != SH_OK) {
    sh_ctx_destroy(ctx);
    return SH_ERR_HIP;
  }
  // Built-in pass shapes (the autotuner refines them per workload): a block pass of 2^13 words
  // up to that size, else 2^12 and strided passes of 8 layers; 8 quads per lane.
  for (uint32_t m = 0; m < 32; m++) ctx->fft_tune[m] = {(uint8_t)(m <= 13 ? 13 : 12), 8, 3, 3};
  *out = ctx;
  return SH_OK;
}

extern "C" int sh_ctx_destroy(sh_ctx* ctx) {
  if (!ctx) return SH_OK;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  for (auto& r : ctx->prof) {
    hipEventDestroy(r.start);
    hipEventDestroy(r.stop);
  }
  for (auto e : ctx->event_pool) hipEventDestroy(e);
  for (auto e : ctx->sync_events) hipEventDestroy(e);
  for (auto& kv : ctx->stage_pending) hipEventDestroy(kv.second.ev);
  for (auto e : ctx->stage_ev_free) hipEventDestroy(e);
  for (auto a : ctx->aux)
    if (a) hipStreamSynchronize(a);
  for (auto& s : ctx->scratch)
    if (s.ptr) hipFree(s.ptr);
  for (auto& kv : ctx->pool)
    for (void* p : kv.second) hipFree(p);
  for (auto& kv : ctx->pool_live) hipFree(kv.first);
  if (ctx->h_xfer) hipHostFree(ctx->h_xfer);
  if (ctx->d_ring) hipFree(ctx->d_ring);
  if (ctx->h_ring) hipHostFree(ctx->h_ring);
  if (ctx->timer_start) hipEventDestroy(ctx->timer_start);
  if (ctx->timer_stop) hipEventDestroy(ctx->timer_stop);
  for (auto a : ctx->aux)
    if (a) hipStreamDestroy(a);
  if (ctx->main_stream) hipStreamDestroy(ctx->main_stream);
  delete ctx;
  return SH_OK;
}

extern "C" const char* sh_last_error(const sh_ctx* ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

extern "C" int sh_sync(sh_ctx* ctx) {
  if (!ctx) return SH_ERR_INVALID;
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SH_OK;
}

// Tables are appended to the arena and never rewritten until the arena is
// reset, which first waits for every stream of the context: a table that an
// enqueued kernel reads stays valid, and a table seen before is reused without
// another copy (its bytes are compared in full, not just hashed).
void sh_stage_all_complete(sh_ctx* ctx) {
  for (auto& kv : ctx->stage_pending) ctx->stage_ev_free.push_back(kv.second.ev);
  ctx->stage_pending.clear();
}

int sh_stage(sh_ctx* ctx, const void* host, uint64_t bytes, void** dev) {
  uint64_t need = (bytes + 255) & ~255ull;
  SH_REQUIRE(ctx, need <= ctx->ring_size, "per-call table of %llu bytes exceeds the staging arena",
             (unsigned long long)bytes);
  std::string key((const char*)host, bytes);
  auto it = ctx->table_cache.find(key);
  if (it != ctx->table_cache.end()) {
    // the first copy was enqueued on some stream of the context: a reuse on another stream is
    // ordered after it by an event wait (same-stream reuse is ordered already)
    auto p = ctx->stage_pending.find(it->second);
    if (p != ctx->stage_pending.end() && p->second.stream != ctx->stream)
      SH_CHECK_HIP(ctx, hipStreamWaitEvent(ctx->stream, p->second.ev, 0));
    *dev = ctx->d_ring + it->second;
    return SH_OK;
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  SH_CHECK_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cap));
  SH_REQUIRE(ctx, cap == hipStreamCaptureStatusNone,
             "a per-call table was not staged before graph capture (run the call eagerly first)");
  if (ctx->ring_off + need > ctx->ring_size) {
    SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->main_stream));
    for (auto a : ctx->aux)
      if (a) SH_CHECK_HIP(ctx, hipStreamSynchronize(a));
    ctx->ring_off = 0;
    ctx->table_cache.clear();
    sh_stage_all_complete(ctx);
  }
  uint8_t* h = ctx->h_ring + ctx->ring_off;
  uint8_t* d = ctx->d_ring + ctx->ring_off;
  memcpy(h, host, bytes);
  SH_CHECK_HIP(ctx, hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
  if (key.size() <= 64 * 1024) {
    // A cached table may later be read from another stream of the context: record the copy's
    // completion for such reuses.  Completed records are reclaimed now and then.
    if (ctx->stage_pending.size() >= 256) {
      for (auto p = ctx->stage_pending.begin(); p != ctx->stage_pending.end();) {
        if (hipEventQuery(p->second.ev) == hipSuccess) {
          ctx->stage_ev_free.push_back(p->second.ev);
          p = ctx->stage_pending.erase(p);
        } else {
          ++p;
        }
      }
    }
    hipEvent_t ev = nullptr;
    if (!ctx->stage_ev_free.empty()) {
      ev = ctx->stage_ev_free.back();
      ctx->stage_ev_free.pop_back();
    } else {
      SH_CHECK_HIP(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    SH_CHECK_HIP(ctx, hipEventRecord(ev, ctx->stream));
    ctx->stage_pending[ctx->ring_off] = {ev, ctx->stream};
    ctx->table_cache.emplace(std::move(key), ctx->ring_off);
  }
  ctx->ring_off += need;
  *dev = d;
  return SH_OK;
}

int sh_host_xfer(sh_ctx* ctx, uint64_t bytes, uint8_t** out) {
  if (ctx->h_xfer_size < bytes) {
    if (ctx->h_xfer) SH_CHECK_HIP(ctx, hipHostFree(ctx->h_xfer));
    ctx->h_xfer = nullptr;
    ctx->h_xfer_size = 0;
    const uint64_t sz = std::max<uint64_t>(bytes, 1ull << 20);
    SH_CHECK_HIP(ctx, hipHostMalloc((void**)&ctx->h_xfer, sz, hipHostMallocDefault));
    ctx->h_xfer_size = sz;
  }
  *out = ctx->h_xfer;
  return SH_OK;
}

int sh_scratch_get(sh_ctx* ctx, int slot, uint64_t bytes, void** out) {
  sh_scratch& s = ctx->scratch[slot];
  if (s.bytes < bytes) {
    if (s.ptr) {
      SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->main_stream));
      for (auto a : ctx->aux)
      if (a) SH_CHECK_HIP(ctx, hipStreamSynchronize(a));
      SH_CHECK_HIP(ctx, hipFree(s.ptr));
      s.ptr = nullptr;
      s.bytes = 0;
    }
    if (hipMalloc(&s.ptr, bytes) != hipSuccess) {
      s.ptr = nullptr;
      return sh_fail(ctx, SH_ERR_OOM, "scratch allocation of %llu bytes failed", (unsigned long long)bytes);
    }
    s.bytes = bytes;
  }
  *out = s.ptr;
  return SH_OK;
}

// The auxiliary stream, created on first use: a context that never forks (every single-size
// commit, every other entry point) holds one stream.
int sh_aux_stream(sh_ctx* ctx, hipStream_t* out) {
  if (!ctx->aux[0]) SH_CHECK_HIP(ctx, hipStreamCreateWithFlags(&ctx->aux[0], hipStreamNonBlocking));
  *out = ctx->aux[0];
  return SH_OK;
}

hipEvent_t sh_sync_event(sh_ctx* ctx, size_t i) {
  while (ctx->sync_events.size() <= i) {
    hipEvent_t e = nullptr;
    hipEventCreateWithFlags(&e, hipEventDisableTiming);
    ctx->sync_events.push_back(e);
  }
  return ctx->sync_events[i];
}

static hipEvent_t pool_event(sh_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

void sh_prof_begin(sh_ctx* ctx, const char* cls, sh_prof_rec* rec) {
  rec->cls = cls;
  rec->start = rec->stop = nullptr;
  if (!ctx->profile) return;
  rec->start = pool_event(ctx);
  hipEventRecord(rec->start, ctx->stream);
}

void sh_prof_end(sh_ctx* ctx, sh_prof_rec* rec) {
  if (!ctx->profile || !rec->start) return;
  rec->stop = pool_event(ctx);
  hipEventRecord(rec->stop, ctx->stream);
  ctx->prof.push_back(*rec);
}

extern "C" int sh_profile_enable(sh_ctx* ctx, int on) {
  if (!ctx) return SH_ERR_INVALID;
  ctx->profile = on != 0;
  return SH_OK;
}

extern "C" int sh_profile_report(sh_ctx* ctx, char* buf, uint64_t cap) {
  if (!ctx || !buf || cap == 0) return SH_ERR_INVALID;
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::map<std::string, std::pair<uint64_t, double>> agg;
  for (auto& r : ctx->prof) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, r.start, r.stop);
    auto& a = agg[r.cls];
    a.first += 1;
    a.second += ms;
    ctx->event_pool.push_back(r.start);
    ctx->event_pool.push_back(r.stop);
  }
  ctx->prof.clear();
  std::string s = "{";
  bool first = true;
  for (auto& kv : agg) {
    char item[256];
    snprintf(item, sizeof(item), "%s\"%s\": {\"n\": %llu, \"ms\": %.6f}", first ? "" : ", ",
             kv.first.c_str(), (unsigned long long)kv.second.first, kv.second.second);
    s += item;
    first = false;
  }
  s += "}";
  if (s.size() + 1 > cap) return sh_fail(ctx, SH_ERR_INVALID, "profile report needs %zu bytes", s.size() + 1);
  memcpy(buf, s.c_str(), s.size() + 1);
  return SH_OK;
}

extern "C" int sh_timer_start(sh_ctx* ctx) {
  if (!ctx) return SH_ERR_INVALID;
  SH_CHECK_HIP(ctx, hipEventRecord(ctx->timer_start, ctx->stream));
  return SH_OK;
}

extern "C" int sh_timer_stop(sh_ctx* ctx, float* ms) {
  if (!ctx || !ms) return SH_ERR_INVALID;
  SH_CHECK_HIP(ctx, hipEventRecord(ctx->timer_stop, ctx->stream));
  SH_CHECK_HIP(ctx, hipEventSynchronize(ctx->timer_stop));
  SH_CHECK_HIP(ctx, hipEventElapsedTime(ms, ctx->timer_start, ctx->timer_stop));
  return SH_OK;
}

static constexpr uint64_t kPoolCapBytes = 64ull << 30;  // cached bytes kept for reuse

static int pool_release(sh_ctx* ctx) {  // caller holds pool_mu
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->main_stream));
  for (auto a : ctx->aux)
    if (a) SH_CHECK_HIP(ctx, hipStreamSynchronize(a));
  for (auto& kv : ctx->pool)
    for (void* p : kv.second) SH_CHECK_HIP(ctx, hipFree(p));
  ctx->pool.clear();
  ctx->pool_bytes = 0;
  return SH_OK;
}

int sh_dev_alloc(sh_ctx* ctx, uint64_t bytes, void** out) {
  std::lock_guard<std::mutex> lock(ctx->pool_mu);
  *out = nullptr;
  const uint64_t cls = ((bytes ? bytes : 4) + 255) & ~255ull;
  if (ctx->arena) {
    sh_arena* a = ctx->arena;
    auto& fl = a->free[cls];
    if (a->mode == 2) {  // capture: the block the warm-up's allocation at this point received
      void* p = a->cursor < a->seq.size() ? a->seq[a->cursor] : nullptr;
      auto b = p ? a->blocks.find(p) : a->blocks.end();
      if (b == a->blocks.end() || b->second.first != cls || b->second.second)
        return sh_fail(ctx, SH_ERR_INVALID,
                       "graph capture: allocation %zu (%llu bytes) differs from the warm-up's", a->cursor,
                       (unsigned long long)cls);
      const auto f = std::find(fl.begin(), fl.end(), p);  // a block not in use is on its free list
      if (f != fl.end()) fl.erase(f);
      b->second.second = true;
      a->cursor++;
      *out = p;
      return SH_OK;
    }
    if (!fl.empty()) {
      *out = fl.back();
      fl.pop_back();
      a->blocks[*out].second = true;
    } else {
      if (hipMalloc(out, cls) != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return sh_fail(ctx, SH_ERR_OOM, "hipMalloc of %llu bytes for a graph failed", (unsigned long long)cls);
      }
      a->blocks[*out] = {cls, true};
      ctx->arena_of[*out] = a;
    }
    if (a->mode == 1) a->seq.push_back(*out);
    return SH_OK;
  }
  auto it = ctx->pool.find(cls);
  if (it != ctx->pool.end() && !it->second.empty()) {
    *out = it->second.back();
    it->second.pop_back();
    ctx->pool_bytes -= cls;
  } else if (hipMalloc(out, cls) != hipSuccess) {
    (void)hipGetLastError();
    SH_TRY(pool_release(ctx));  // cached blocks of other sizes may be what is missing
    if (hipMalloc(out, cls) != hipSuccess) {
      *out = nullptr;
      return sh_fail(ctx, SH_ERR_OOM, "hipMalloc of %llu bytes failed", (unsigned long long)cls);
    }
  }
  ctx->pool_live[*out] = cls;
  return SH_OK;
}

void sh_dev_free(sh_ctx* ctx, void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lock(ctx->pool_mu);
  auto ao = ctx->arena_of.find(p);
  if (ao != ctx->arena_of.end()) {  // a graph's block goes back to that graph
    auto& b = ao->second->blocks[p];
    if (b.second) {
      b.second = false;
      ao->second->free[b.first].push_back(p);
    }
    return;
  }
  auto it = ctx->pool_live.find(p);
  // Not a live block of the pool or of a live graph: freed already, or a block of a graph that
  // was destroyed (every such block was released with it).  Nothing is released: the address may
  // by now belong to a newer allocation.
  if (it == ctx->pool_live.end()) return;
  const uint64_t cls = it->second;
  ctx->pool_live.erase(it);
  if (ctx->pool_bytes + cls > kPoolCapBytes) {
    hipStreamSynchronize(ctx->main_stream);
    for (auto a : ctx->aux)
      if (a) hipStreamSynchronize(a);
    hipFree(p);
    return;
  }
  ctx->pool[cls].push_back(p);
  ctx->pool_bytes += cls;
}

void sh_arena_release(sh_ctx* ctx, sh_arena* a) {
  std::lock_guard<std::mutex> lock(ctx->pool_mu);
  for (auto& kv : a->blocks) {
    hipFree(kv.first);
    ctx->arena_of.erase(kv.first);
  }
  a->blocks.clear();
  a->free.clear();
  if (ctx->arena == a) ctx->arena = nullptr;
}

extern "C" int sh_alloc(sh_ctx* ctx, uint64_t bytes, void** d_out) {
  if (!ctx || !d_out) return SH_ERR_INVALID;
  return sh_dev_alloc(ctx, bytes, d_out);
}

extern "C" int sh_free(sh_ctx* ctx, void* d) {
  if (!ctx) return SH_ERR_INVALID;
  sh_dev_free(ctx, d);
  return SH_OK;
}

extern "C" int sh_pool_trim(sh_ctx* ctx) {
  if (!ctx) return SH_ERR_INVALID;
  std::lock_guard<std::mutex> lock(ctx->pool_mu);
  return pool_release(ctx);
}

extern "C" int sh_upload(sh_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes) {
  if (!ctx) return SH_ERR_INVALID;
  SH_CHECK_HIP(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SH_OK;
}

extern "C" int sh_download(sh_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes) {
  if (!ctx) return SH_ERR_INVALID;
  SH_CHECK_HIP(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SH_OK;
}

// Fill kernel (not hipMemsetAsync, so the call stays safe inside a captured graph).
// Byte fill: 16-byte stores over the 16-byte-aligned middle, single bytes for the unaligned head
// and the tail (each under 16 bytes, so the first 16 lanes of the grid cover them).
__global__ void k_fill_u8(uint8_t* p, uint8_t v, uint64_t n) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t lead = (16 - ((uintptr_t)p & 15)) & 15;
  const uint64_t head = lead < n ? lead : n;
  const uint64_t nq = (n - head) >> 4;
  const uint32_t w = v * 0x01010101u;
  const uint4 x = make_uint4(w, w, w, w);
  uint4* q = (uint4*)(p + head);
  for (uint64_t i = tid; i < nq; i += stride) q[i] = x;
  const uint64_t done = head + (nq << 4);
  if (tid < head) p[tid] = v;
  if (tid < n - done) p[done + tid] = v;
}

extern "C" int sh_copy(sh_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes) {
  if (!ctx) return SH_ERR_INVALID;
  if (bytes == 0) return SH_OK;
  SH_CHECK_HIP(ctx, hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return SH_OK;
}

extern "C" int sh_memset(sh_ctx* ctx, void* d_dst, int value, uint64_t bytes) {
  if (!ctx) return SH_ERR_INVALID;
  if (bytes == 0) return SH_OK;
  uint64_t blocks = (bytes / 16 + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 65536) blocks = 65536;
  k_fill_u8<<<(unsigned)blocks, 256, 0, ctx->stream>>>((uint8_t*)d_dst, (uint8_t)value, bytes);
  return sh_launch_check(ctx, "fill");
}

// bit_reverse_column: swaps i <-> bitrev(i) for i < bitrev(i).
__global__ void k_bit_reverse(uint32_t* col, uint32_t log_n) {
  uint64_t n = 1ull << log_n;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t j = sh::bitrev((uint32_t)i, log_n);
  if (j > i) {
    uint32_t a = col[i], b = col[j];
    col[i] = b;
    col[j] = a;
  }
}

extern "C" int sh_bit_reverse_column(sh_ctx* ctx, uint32_t* d_col, uint32_t log_size) {
  if (!ctx) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, log_size <= 31, "log_size %u too large", log_size);
  uint64_t n = 1ull << log_size;
  SpanGuard span(ctx, "bit_reverse");
  k_bit_reverse<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(d_col, log_size);
  return sh_launch_check(ctx, "bit_reverse");
}

int sh_point_tab_make(sh_ctx* ctx, uint32_t init_index, uint32_t log_size, sh_point_tab* out) {
  uint32_t levels = log_size == 0 ? 0 : (log_size + 10) / 11;
  out->levels = levels;
  out->init = sh::index_to_point_host(init_index);
  out->t = nullptr;
  if (levels == 0) return SH_OK;
  std::vector<sh::mpt> tab(2048ull * levels);
  uint32_t step = sh::subgroup_gen(log_size);
  for (uint32_t l = 0; l < levels; l++) {
    sh::mpt unit = sh::index_to_point_host(sh::idx_reduce((uint64_t)step << (11 * l)));
    sh::mpt cur = {1, 0};
    for (uint32_t k = 0; k < 2048; k++) {
      tab[l * 2048 + k] = cur;
      cur = sh::padd(cur, unit);
    }
  }
  void* d;
  SH_TRY(sh_stage(ctx, tab.data(), tab.size() * sizeof(sh::mpt), &d));
  out->t = (const sh::mpt*)d;
  return SH_OK;
}

namespace sh {
mpt index_to_point_host(uint32_t index) {
  mpt res = {1, 0}, cur = {2, 1268011823u};  // M31_CIRCLE_GEN (crates/stwo/src/core/circle.rs:181-182)
  uint32_t s = index & 0x7fffffffu;
  while (s) {
    if (s & 1) res = padd(res, cur);
    cur = padd(cur, cur);
    s >>= 1;
  }
  return res;
}

const mpt* gen_pow2_table_host() {
  static mpt table[31];
  static bool init = false;
  if (!init) {
    mpt cur = {2, 1268011823u};
    for (int b = 0; b < 31; b++) {
      table[b] = cur;
      cur = padd(cur, cur);
    }
    init = true;
  }
  return table;
}

bool is_doubling_of(uint32_t init, uint32_t log, uint32_t root_init, uint32_t root_log) {
  if (log > root_log) return false;
  uint32_t d = root_log - log;
  uint64_t ri = root_init & 0x7fffffffu;
  for (uint32_t k = 0; k < d; k++) ri = (ri * 2) & 0x7fffffffull;
  return (uint32_t)ri == (init & 0x7fffffffu);
}
}  // namespace sh

extern "C" int sh_set_fft_tuning(sh_ctx* ctx, uint32_t log_size, uint32_t block_log, uint32_t stride_layers,
                                 uint32_t r_block, uint32_t r_stride) {
  if (!ctx) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, log_size < 32, "log size %u out of range", log_size);
  SH_REQUIRE(ctx, block_log >= 1 && block_log <= 13, "block_log %u not in [1, 13]", block_log);
  SH_REQUIRE(ctx, stride_layers >= 1 && stride_layers <= 8, "stride_layers %u not in [1, 8]", stride_layers);
  SH_REQUIRE(ctx, r_block >= 1 && r_block <= 4 && r_stride >= 1 && r_stride <= 4,
             "rows per lane 2^%u / 2^%u not in [2, 16]", r_block, r_stride);
  ctx->fft_tune[log_size] = {(uint8_t)block_log, (uint8_t)stride_layers, (uint8_t)r_block, (uint8_t)r_stride};
  return SH_OK;
}

extern "C" int sh_set_fft_fusion(sh_ctx* ctx, uint32_t mode) {
  if (!ctx) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, mode <= 2, "FFT fusion mode %u not in [0, 2]", mode);
  ctx->fft_fuse = (uint8_t)mode;
  return SH_OK;
}

extern "C" int sh_get_fft_tuning(sh_ctx* ctx, uint32_t log_size, uint32_t out[4]) {
  if (!ctx || !out || log_size >= 32) return SH_ERR_INVALID;
  const auto& t = ctx->fft_tune[log_size];
  out[0] = t.block_log;
  out[1] = t.stride_layers;
  out[2] = t.r_block;
  out[3] = t.r_stride;
  return SH_OK;
}
