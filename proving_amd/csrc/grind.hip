// Proof-of-work grinding (GrindOps::grind for the Blake2s channel).
//
// Predicate (crates/stwo/src/core/channel/blake2s.rs:102-117):
//   prefixed = Blake2s(LE32(0x12345678) || digest || LE32(pow_bits))
//   ok(nonce) = trailing_zeros(LE u128 of Blake2s(prefixed || LE64(nonce))[0..16]) >= pow_bits
// The result is the smallest valid nonce, as CpuBackend's linear search returns
// (crates/stwo/src/prover/backend/cpu/grind.rs:5-15) and the SIMD backend's
// deterministic chunked search (crates/stwo/src/prover/backend/simd/grind.rs:78-122).
//
// Device search: rounds of consecutive nonces; each lane hashes 16 nonces (one
// 40-byte single-block compression each) and the round's smallest hit is kept
// with atomicMin.  Rounds run in increasing order and stop at the first round
// with a hit, so the answer is the global minimum.  The first round covers
// 2^(pow_bits + 5) nonces (a hit with probability 1 - e^-32), later rounds
// double up to 2^26, so a small grind costs a small launch.
#include <string.h>

#include "blake2s.h"
#include "common.h"

namespace {

constexpr uint32_t kPerLane = 16;
constexpr uint32_t kRoundLog = 26;     // largest round
constexpr uint32_t kMinRoundLog = 12;  // kPerLane * 256 nonces: one workgroup

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

}  // namespace

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
  SH_CHECK_HIP(ctx, hipMemcpyAsync(d_pre, pre, 32, hipMemcpyHostToDevice, ctx->stream));
  const unsigned long long none = ~0ull;
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
