// A device-resident Blake2sChannel (crates/stwo/src/core/channel/blake2s.rs:16-117).
//
// The state lives in 16 device words: the digest as 8 little-endian words and n_draws (word 8),
// the layout the FRI driver's device channel uses.  Every operation is a one-wave launch on the
// context stream, so a proof can mix commitment roots, proof-of-work nonces and sampled values
// and draw its randomness without the host waiting for any of them; the host reads the state (or
// the drawn values) whenever it next synchronises.
//
// mix: digest = Blake2s(digest || words), n_draws = 0 (mix_u32s / mix_u64 / mix_felts / mix_root
// are all this over their little-endian words).  A long input streams block by block through one
// quad of lanes (blake2s_quad.h): four lanes share each compression, so the 67 blocks of 264
// sampled values take about a quarter of a single lane's latency.
// draw: Blake2s(digest || LE32(n_draws) || 0x00) (37 bytes), n_draws += 1; a secure felt is
// f0 + 2 f1, f2 + 2 f3, ... of the 8 reduced words.
#include <cstring>

#include "blake2s.h"
#include "blake2s_quad.h"
#include "common.h"

namespace {

// Compression of one streamed block by the quad in lanes 0..3: h is the chaining value (lane q
// holds h[q] and h[4 + q]); t = bytes so far including this block, last = the final block.
__device__ __forceinline__ void quad_stream_block(const uint32_t* blk, const uint32_t (&sidx)[40], uint32_t q,
                                                  uint32_t& h_lo, uint32_t& h_hi, uint32_t t, bool last) {
  const uint32_t d0 = q == 0 ? (SH_B2_IV4 ^ t) : q == 1 ? SH_B2_IV5 : q == 2 ? (last ? ~SH_B2_IV6 : SH_B2_IV6)
                                                                               : SH_B2_IV7;
  uint32_t lo, hi;
  quad_node_hash(blk, sidx, q, h_lo, h_hi, d0, lo, hi);
  h_lo = lo;
  h_hi = hi;
}

__global__ __launch_bounds__(64) void k_dch_mix(uint32_t* __restrict__ ch, const uint32_t* __restrict__ words,
                                                uint32_t n) {
  __shared__ uint32_t dig[8];
  __shared__ uint32_t blk[2][16];
  const uint32_t lane = threadIdx.x, q = lane & 3;
  if (lane < 8) dig[lane] = ch[lane];
  __syncthreads();
  uint32_t sidx[40];
  quad_sigma(q, sidx);
  uint32_t h_lo = q == 0 ? (SH_B2_IV0 ^ 0x01010020u) : q == 1 ? SH_B2_IV1 : q == 2 ? SH_B2_IV2 : SH_B2_IV3;
  uint32_t h_hi = q == 0 ? SH_B2_IV4 : q == 1 ? SH_B2_IV5 : q == 2 ? SH_B2_IV6 : SH_B2_IV7;
  const uint32_t total = 32 + 4 * n;
  const uint32_t nb = (total + 63) / 64;
  // stream word k: the digest, then the input words, then zero padding of the last block
  auto word = [&](uint32_t k) { return k < 8 ? dig[k] : (k - 8 < n ? words[k - 8] : 0u); };
  if (lane < 16) blk[0][lane] = word(lane);
  __syncthreads();
  for (uint32_t j = 0; j < nb; j++) {
    // the next block is loaded by lanes 16..31 while the quad compresses this one
    if (lane >= 16 && lane < 32 && j + 1 < nb) blk[(j + 1) & 1][lane - 16] = word(16 * (j + 1) + lane - 16);
    if (lane < 4) {
      const bool last = j + 1 == nb;
      quad_stream_block(blk[j & 1], sidx, q, h_lo, h_hi, last ? total : 64 * (j + 1), last);
    }
    __syncthreads();
  }
  if (lane < 4) {
    ch[q] = h_lo;
    ch[4 + q] = h_hi;
  }
  if (lane == 0) ch[8] = 0;
}

// One draw: the 8 words of Blake2s(digest || LE32(n_draws) || 0x00); n_draws += 1.
__device__ void dch_draw_u32s(uint32_t* ch, uint32_t u[8]) {
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 8; i++) m[i] = ch[i];
  m[8] = ch[8];
#pragma unroll
  for (int i = 9; i < 16; i++) m[i] = 0;
  sh::b2_init(u);
  sh::b2_compress(u, m, 37, true);
  ch[8] = ch[8] + 1;
}
__device__ sh::qm31 dch_draw_felt(uint32_t* ch) {
  uint32_t u[8], w[4];
  dch_draw_u32s(ch, u);
  uint32_t f[8];
#pragma unroll
  for (int i = 0; i < 8; i++) f[i] = sh::reduce_once((u[i] & sh::P) + (u[i] >> 31));  // M31::reduce(u32)
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] = sh::madd(sh::madd(f[2 * i], f[2 * i + 1]), f[2 * i + 1]);
  return sh::qload(w);
}

// kind 0: n secure felts; kind 1: CirclePoint::get_random_point (crates/stwo/src/core/circle.rs:
// 150-160): t drawn, ((1 - t^2) / (1 + t^2), 2t / (1 + t^2)).
__global__ void k_dch_draw(uint32_t* __restrict__ ch, uint32_t kind, uint32_t n, uint32_t* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (kind == 0) {
    for (uint32_t i = 0; i < n; i++) sh::qstore(dch_draw_felt(ch), out + 4 * i);
    return;
  }
  const sh::qm31 t = dch_draw_felt(ch);
  const sh::qm31 t2 = sh::qmul(t, t);
  const sh::qm31 inv = sh::qinv(sh::qadd(t2, sh::qone()));
  sh::qstore(sh::qmul(sh::qsub(sh::qone(), t2), inv), out);
  sh::qstore(sh::qmul(sh::qadd(t, t), inv), out + 4);
}

// out[i] = felt^i for i < n (reversed: out[i] = felt^(n - 1 - i)); lane j of the wave computes
// the powers j, j + 64, j + 128, ... (felt^j by square-and-multiply, then steps of felt^64).
__global__ __launch_bounds__(64) void k_dch_powers(const uint32_t* __restrict__ felt, uint32_t n, uint32_t reversed,
                                                   uint32_t* __restrict__ out) {
  const uint32_t j = threadIdx.x;
  const sh::qm31 f = sh::qload(felt);
  sh::qm31 step = f;
  for (int k = 0; k < 6; k++) step = sh::qmul(step, step);  // f^64
  sh::qm31 p = sh::qone(), b = f;
  for (uint32_t e = j; e; e >>= 1) {
    if (e & 1) p = sh::qmul(p, b);
    b = sh::qmul(b, b);
  }
  for (uint32_t i = j; i < n; i += 64) {
    sh::qstore(p, out + 4ull * (reversed ? n - 1 - i : i));
    p = sh::qmul(p, step);
  }
}

}  // namespace

// Device channel operations (stream-ordered on the context stream; d_ch: 16 device words).
int sh_dch_mix_impl(sh_ctx* ctx, uint32_t* d_ch, const uint32_t* d_words, uint64_t n) {
  SH_REQUIRE(ctx, n < (1ull << 28), "too many words to mix (%llu)", (unsigned long long)n);
  SpanGuard span(ctx, "channel");
  k_dch_mix<<<1, 64, 0, ctx->stream>>>(d_ch, d_words, (uint32_t)n);
  return sh_launch_check(ctx, "channel_mix");
}

extern "C" int sh_dchannel_mix(sh_ctx* ctx, uint32_t* d_ch, const uint32_t* d_words, uint64_t n) {
  if (!ctx || !d_ch || (n && !d_words)) return SH_ERR_INVALID;
  return sh_dch_mix_impl(ctx, d_ch, d_words, n);
}

extern "C" int sh_dchannel_draw_secure_felts(sh_ctx* ctx, uint32_t* d_ch, uint32_t n, uint32_t* d_out) {
  if (!ctx || !d_ch || (n && !d_out)) return SH_ERR_INVALID;
  if (n == 0) return SH_OK;
  SpanGuard span(ctx, "channel");
  k_dch_draw<<<1, 64, 0, ctx->stream>>>(d_ch, 0, n, d_out);
  return sh_launch_check(ctx, "channel_draw");
}

extern "C" int sh_dchannel_draw_random_point(sh_ctx* ctx, uint32_t* d_ch, uint32_t* d_out) {
  if (!ctx || !d_ch || !d_out) return SH_ERR_INVALID;
  SpanGuard span(ctx, "channel");
  k_dch_draw<<<1, 64, 0, ctx->stream>>>(d_ch, 1, 1, d_out);
  return sh_launch_check(ctx, "channel_draw_point");
}

extern "C" int sh_secure_powers_dev(sh_ctx* ctx, const uint32_t* d_felt, uint64_t n, int reversed, uint32_t* d_out) {
  if (!ctx || !d_felt || (n && !d_out)) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, n < (1ull << 26), "too many powers (%llu)", (unsigned long long)n);
  if (n == 0) return SH_OK;
  SpanGuard span(ctx, "channel");
  k_dch_powers<<<1, 64, 0, ctx->stream>>>(d_felt, (uint32_t)n, reversed ? 1u : 0u, d_out);
  return sh_launch_check(ctx, "secure_powers");
}

// The device channel set to a host state (staged, stream-ordered: the host does not wait).
extern "C" int sh_dchannel_load(sh_ctx* ctx, uint32_t* d_ch, const uint8_t digest[32], uint32_t n_draws) {
  if (!ctx || !d_ch || !digest) return SH_ERR_INVALID;
  uint32_t w[16] = {};
  std::memcpy(w, digest, 32);
  w[8] = n_draws;
  void* staged;
  SH_TRY(sh_stage(ctx, w, sizeof(w), &staged));
  SH_CHECK_HIP(ctx, hipMemcpyAsync(d_ch, staged, sizeof(w), hipMemcpyDeviceToDevice, ctx->stream));
  return SH_OK;
}
