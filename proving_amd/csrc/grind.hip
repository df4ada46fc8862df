// Proof-of-work grinding (GrindOps::grind for the Blake2s channel).
//
// Predicate (crates/stwo/src/core/channel/blake2s.rs:102-117):
//   prefixed = Blake2s(LE32(0x12345678) || digest || LE32(pow_bits))
//   ok(nonce) = trailing_zeros(LE u128 of Blake2s(prefixed || LE64(nonce))[0..16]) >= pow_bits
// The result is the smallest valid nonce, as CpuBackend's linear search returns
// (crates/stwo/src/prover/backend/cpu/grind.rs:5-15) and the SIMD backend's
// deterministic chunked search (crates/stwo/src/prover/backend/simd/grind.rs:78-122).
//
// Device search (pow_bits <= 32): one launch over 2^(pow_bits + 6) consecutive nonces.
// Workgroups claim chunks of 4096 nonces in increasing order (an atomic counter); each lane
// hashes 16 consecutive nonces of the chunk and the smallest hit is kept with atomicMin.  A
// workgroup stops claiming once the best hit lies below the next chunk, so the search costs
// about the position of the smallest valid nonce (plus the chunks in flight) instead of whole
// rounds.  Every chunk below the smallest hit is claimed before any chunk above it and is
// hashed completely, so the result is the smallest nonce.  Without a hit in the range the
// next range is searched.
//
// Per nonce, one 40-byte single-block compression.  Round 0's column step reads the 8
// prefixed-digest words and three of its four diagonal G functions read zero words: all of
// that is the same for every nonce and is computed once on the host; only G(0, 5, 10, 15) of
// round 0 sees the nonce.  Only digest word 0 is tested (pow_bits <= 32), so the compiler
// drops what of the last round does not feed v[0] or v[8].
//
// pow_bits above 32 (never used by the protocol) take rounds of a plain per-nonce
// compression with all four tested words.
#include <string.h>

#include <algorithm>

#include "blake2s.h"
#include "common.h"

namespace {

constexpr uint32_t kPerLane = 16;
constexpr uint32_t kRoundLog = 26;     // largest round
constexpr uint32_t kMinRoundLog = 12;  // kPerLane * 256 nonces: one workgroup
constexpr uint32_t kChunk = kPerLane * 256;
constexpr uint32_t kMaxGroups = 2048;  // 8 workgroups of 256 lanes per CU

// The compression state after the nonce-independent part of round 0, and the message words
// 0..7 (the prefixed digest).
struct GrindPre {
  uint32_t v[16];
  uint32_t m[8];
};

SH_HD void grind_prefix_state(const uint32_t pre[8], GrindPre* g) {
  uint32_t h[8], v[16];
  sh::b2_init(h);
  for (int i = 0; i < 8; i++) v[i] = h[i];
  v[8] = SH_B2_IV0;
  v[9] = SH_B2_IV1;
  v[10] = SH_B2_IV2;
  v[11] = SH_B2_IV3;
  v[12] = SH_B2_IV4 ^ 40u;  // t = 40 bytes, one block
  v[13] = SH_B2_IV5;
  v[14] = ~SH_B2_IV6;       // final block
  v[15] = SH_B2_IV7;
  using sh::rotr32;
  SH_B2_G(0, 4, 8, 12, pre[0], pre[1]);
  SH_B2_G(1, 5, 9, 13, pre[2], pre[3]);
  SH_B2_G(2, 6, 10, 14, pre[4], pre[5]);
  SH_B2_G(3, 7, 11, 15, pre[6], pre[7]);
  SH_B2_G(1, 6, 11, 12, 0u, 0u);
  SH_B2_G(2, 7, 8, 13, 0u, 0u);
  SH_B2_G(3, 4, 9, 14, 0u, 0u);
  for (int i = 0; i < 16; i++) g->v[i] = v[i];
  for (int i = 0; i < 8; i++) g->m[i] = pre[i];
}

// Digest word 0 of Blake2s(prefixed || LE64(nonce)) from the prefix state.
__device__ __forceinline__ uint32_t grind_word0(const GrindPre& g, uint32_t lo, uint32_t hi) {
  using sh::rotr32;
  uint32_t v[16];
#pragma unroll
  for (int i = 0; i < 16; i++) v[i] = g.v[i];
  const uint32_t m[16] = {g.m[0], g.m[1], g.m[2], g.m[3], g.m[4], g.m[5], g.m[6], g.m[7], lo, hi, 0, 0, 0, 0, 0, 0};
  SH_B2_G(0, 5, 10, 15, m[8], m[9]);  // the rest of round 0
  SH_B2_ROUND(14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3);
  SH_B2_ROUND(11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4);
  SH_B2_ROUND(7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8);
  SH_B2_ROUND(9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13);
  SH_B2_ROUND(2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9);
  SH_B2_ROUND(12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11);
  SH_B2_ROUND(13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10);
  SH_B2_ROUND(6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5);
  SH_B2_ROUND(10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0);
  return (SH_B2_IV0 ^ 0x01010020u) ^ v[0] ^ v[8];
}

// ctl[0..1]: best nonce (u64, all ones when none), ctl[2]: next chunk to claim.
__global__ __launch_bounds__(256) void k_grind32(const GrindPre* __restrict__ gp, uint32_t mask, uint64_t base,
                                                 uint32_t n_chunks, unsigned int* ctl) {
  const GrindPre g = *gp;
  unsigned long long* best = (unsigned long long*)ctl;
  unsigned int* next = ctl + 2;
  __shared__ uint32_t s_chunk;
  __shared__ int s_stop;
  for (;;) {
    if (threadIdx.x == 0) {
      const uint32_t c = atomicAdd(next, 1u);
      const unsigned long long b = __hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_chunk = c;
      s_stop = c >= n_chunks || b < base + (uint64_t)c * kChunk;
    }
    __syncthreads();
    const uint32_t c = s_chunk;
    const int stop = s_stop;
    __syncthreads();  // s_chunk and s_stop are rewritten by the next claim
    if (stop) return;
    const uint64_t first = base + (uint64_t)c * kChunk + (uint64_t)threadIdx.x * kPerLane;
    for (uint32_t k = 0; k < kPerLane; k++) {
      const uint64_t nonce = first + k;
      if ((grind_word0(g, (uint32_t)nonce, (uint32_t)(nonce >> 32)) & mask) == 0) {
        atomicMin(best, (unsigned long long)nonce);
        break;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_grind(const uint32_t* __restrict__ pre_w, uint32_t pow_bits,
                                               uint64_t base, unsigned long long* best) {
  uint32_t pre[8];
#pragma unroll
  for (int i = 0; i < 8; i++) pre[i] = pre_w[i];
  const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t first = base + lane * kPerLane;
  const uint32_t m0 = pow_bits >= 32 ? 0xffffffffu : ((1u << pow_bits) - 1);
  const uint32_t rest = pow_bits > 32 ? pow_bits - 32 : 0;
  const uint32_t m1 = rest >= 32 ? 0xffffffffu : ((1u << rest) - 1);
  const uint32_t rest2 = pow_bits > 64 ? pow_bits - 64 : 0;
  const uint32_t m2 = rest2 >= 32 ? 0xffffffffu : ((1u << rest2) - 1);
  const uint32_t rest3 = pow_bits > 96 ? pow_bits - 96 : 0;
  const uint32_t m3 = rest3 >= 32 ? 0xffffffffu : ((1u << rest3) - 1);
  for (uint32_t k = 0; k < kPerLane; k++) {
    const uint64_t nonce = first + k;
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 8; i++) m[i] = pre[i];
    m[8] = (uint32_t)nonce;
    m[9] = (uint32_t)(nonce >> 32);
#pragma unroll
    for (int i = 10; i < 16; i++) m[i] = 0;
    uint32_t h[8];
    sh::b2_init(h);
    sh::b2_compress(h, m, 40, true);
    if ((h[0] & m0) == 0 && (h[1] & m1) == 0 && (h[2] & m2) == 0 && (h[3] & m3) == 0) {
      atomicMin(best, (unsigned long long)nonce);
      break;
    }
  }
}

// The prefix state from a device channel's digest: prefixed = Blake2s(LE32(0x12345678) || digest
// || LE32(pow_bits)) (one 40-byte block), then the nonce-independent part of round 0.  Also resets
// the search control words (best nonce = none, next chunk = 0).
__global__ void k_grind_prep(const uint32_t* __restrict__ ch, uint32_t pow_bits, GrindPre* __restrict__ out,
                             unsigned int* __restrict__ ctl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t m[16], h[8];
  m[0] = 0x12345678u;
#pragma unroll
  for (int i = 0; i < 8; i++) m[1 + i] = ch[i];
  m[9] = pow_bits;
#pragma unroll
  for (int i = 10; i < 16; i++) m[i] = 0;
  sh::b2_init(h);
  sh::b2_compress(h, m, 40, true);
  grind_prefix_state(h, out);
  ctl[0] = 0xffffffffu;
  ctl[1] = 0xffffffffu;
  ctl[2] = 0;
  ctl[3] = 0;
}

uint32_t grind_groups(uint32_t pow_bits, uint64_t n_chunks) {
  // about four times the expected position of the first hit in flight at once: a small grind
  // does not hash a whole device's worth of chunks before the first claims see the hit
  const uint64_t want = std::max<uint64_t>(64, ((1ull << pow_bits) / kChunk) * 4);
  return (uint32_t)std::min<uint64_t>(std::min<uint64_t>(n_chunks, want), kMaxGroups);
}

}  // namespace

int sh_dch_mix_impl(sh_ctx* ctx, uint32_t* d_ch, const uint32_t* d_words, uint64_t n);

// GrindOps::grind on a device channel (crates/stwo/src/core/proof_of_work.rs:10-14) followed by
// mix_u64(nonce) (crates/stwo/src/prover/mod.rs:74-81, crates/stwo/src/prover/fri.rs:243-244),
// with nothing returned to the host: d_ctl (16 device bytes) receives the nonce as a
// little-endian u64 in its first 8 bytes.  One search launch over 2^(pow_bits + 12) nonces (the
// chance of no valid nonce in it is below e^-4000); if none were found the nonce reads as all
// ones, which no valid proof contains.
extern "C" int sh_grind_dchannel(sh_ctx* ctx, uint32_t* d_ch, uint32_t pow_bits, uint32_t* d_ctl) {
  if (!ctx || !d_ch || !d_ctl) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, pow_bits <= 32, "pow_bits %u above 32 has no device-channel search", pow_bits);
  void* scr;
  SH_TRY(sh_scratch_get(ctx, SCR_GRIND, 256, &scr));
  GrindPre* d_pre = (GrindPre*)((uint8_t*)scr + 64);
  SpanGuard span(ctx, "grind");
  k_grind_prep<<<1, 64, 0, ctx->stream>>>(d_ch, pow_bits, d_pre, (unsigned int*)d_ctl);
  SH_TRY(sh_launch_check(ctx, "grind_prep"));
  const uint32_t mask = pow_bits == 32 ? 0xffffffffu : ((1u << pow_bits) - 1);
  const uint32_t range_log = std::max<uint32_t>(pow_bits + 12, 16);
  const uint64_t n_chunks = std::min<uint64_t>((1ull << range_log) / kChunk, 0xffffffffull);
  k_grind32<<<grind_groups(pow_bits, n_chunks), 256, 0, ctx->stream>>>(d_pre, mask, 0, (uint32_t)n_chunks,
                                                                      (unsigned int*)d_ctl);
  SH_TRY(sh_launch_check(ctx, "grind"));
  return sh_dch_mix_impl(ctx, d_ch, d_ctl, 2);
}

extern "C" int sh_grind(sh_ctx* ctx, const uint8_t digest[32], uint32_t pow_bits, uint64_t* nonce) {
  if (!ctx || !digest || !nonce) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, pow_bits <= 64, "pow_bits %u above 64 is not searchable", pow_bits);
  uint8_t buf[40], pre[32];
  const uint32_t prefix = 0x12345678u;
  memcpy(buf, &prefix, 4);
  memcpy(buf + 4, digest, 32);
  memcpy(buf + 36, &pow_bits, 4);
  sh::b2s_hash(buf, 40, pre);
  void* scr;
  SH_TRY(sh_scratch_get(ctx, SCR_GRIND, 64, &scr));
  uint32_t* d_pre = (uint32_t*)scr;
  unsigned long long* d_best = (unsigned long long*)((uint8_t*)scr + 32);
  const unsigned long long none = ~0ull;
  if (pow_bits <= 32) {
    uint32_t pw[8];
    memcpy(pw, pre, 32);
    GrindPre g;
    grind_prefix_state(pw, &g);
    const uint32_t mask = pow_bits == 32 ? 0xffffffffu : ((1u << pow_bits) - 1);
    const uint32_t range_log = std::max<uint32_t>(pow_bits + 6, 16);  // 2^range_log nonces per launch
    const uint64_t n_chunks = (1ull << range_log) / kChunk;
    const unsigned groups = grind_groups(pow_bits, n_chunks);
    void* d_g;
    SH_TRY(sh_stage(ctx, &g, sizeof(g), &d_g));
    unsigned int* d_ctl = (unsigned int*)d_best;
    // the control words' initial value comes from the staging arena (staged once, then a device
    // copy per launch), not from pageable host memory
    const uint32_t init[4] = {0xffffffffu, 0xffffffffu, 0u, 0u};
    void* d_init;
    SH_TRY(sh_stage(ctx, init, sizeof(init), &d_init));
    SpanGuard span(ctx, "grind");
    for (uint64_t base = 0;; base += n_chunks * kChunk) {
      SH_CHECK_HIP(ctx, hipMemcpyAsync(d_ctl, d_init, 16, hipMemcpyDeviceToDevice, ctx->stream));
      k_grind32<<<groups, 256, 0, ctx->stream>>>((const GrindPre*)d_g, mask, base, (uint32_t)n_chunks, d_ctl);
      SH_TRY(sh_launch_check(ctx, "grind"));
      unsigned long long best = 0;
      SH_CHECK_HIP(ctx, hipMemcpyAsync(&best, d_best, 8, hipMemcpyDeviceToHost, ctx->stream));
      SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
      if (best != none) {
        *nonce = best;
        return SH_OK;
      }
      if (base > (~0ull) - 2 * n_chunks * kChunk) return sh_fail(ctx, SH_ERR_NOT_FOUND, "no nonce found");
    }
  }
  SH_CHECK_HIP(ctx, hipMemcpyAsync(d_pre, pre, 32, hipMemcpyHostToDevice, ctx->stream));
  uint32_t round_log = pow_bits + 5 < kMinRoundLog ? kMinRoundLog : (pow_bits + 5 > kRoundLog ? kRoundLog : pow_bits + 5);
  SpanGuard span(ctx, "grind");
  for (uint64_t base = 0;;) {
    const uint64_t round = 1ull << round_log;
    const unsigned grid = (unsigned)(round / kPerLane / 256);
    SH_CHECK_HIP(ctx, hipMemcpyAsync(d_best, &none, 8, hipMemcpyHostToDevice, ctx->stream));
    k_grind<<<grid, 256, 0, ctx->stream>>>(d_pre, pow_bits, base, d_best);
    SH_TRY(sh_launch_check(ctx, "grind"));
    unsigned long long best = 0;
    SH_CHECK_HIP(ctx, hipMemcpyAsync(&best, d_best, 8, hipMemcpyDeviceToHost, ctx->stream));
    SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (best != none) {
      *nonce = best;
      return SH_OK;
    }
    if (base > (~0ull) - 2 * round) return sh_fail(ctx, SH_ERR_NOT_FOUND, "no nonce found");
    base += round;
    if (round_log < kRoundLog) round_log++;
  }
}
