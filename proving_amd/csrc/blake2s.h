// Blake2s-256 compression (RFC 7693) for host and device.
//
// The Merkle hashers of the reference hash byte streams with the `blake2` crate
// (crates/stwo/src/core/vcs/blake2_hash.rs:71-99); the SIMD backend hashes the
// same streams 16 rows at a time (crates/stwo/src/prover/backend/simd/blake2s.rs:
// 148-299).  On CDNA4 each lane owns one message and runs the ten rounds in
// registers: 32-bit adds fold into v_add3_u32 and rotations into v_alignbit_b32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SH_HD __host__ __device__ __forceinline__

namespace sh {

#define SH_B2_IV0 0x6A09E667u
#define SH_B2_IV1 0xBB67AE85u
#define SH_B2_IV2 0x3C6EF372u
#define SH_B2_IV3 0xA54FF53Au
#define SH_B2_IV4 0x510E527Fu
#define SH_B2_IV5 0x9B05688Cu
#define SH_B2_IV6 0x1F83D9ABu
#define SH_B2_IV7 0x5BE0CD19u

SH_HD uint32_t rotr32(uint32_t x, uint32_t n) { return (x >> n) | (x << (32 - n)); }

// Initial chaining value: IV with the parameter block (digest 32, no key,
// fanout 1, depth 1) folded into word 0.
SH_HD void b2_init(uint32_t h[8]) {
  h[0] = SH_B2_IV0 ^ 0x01010020u;
  h[1] = SH_B2_IV1;
  h[2] = SH_B2_IV2;
  h[3] = SH_B2_IV3;
  h[4] = SH_B2_IV4;
  h[5] = SH_B2_IV5;
  h[6] = SH_B2_IV6;
  h[7] = SH_B2_IV7;
}

#define SH_B2_G(a, b, c, d, x, y) \
  do {                            \
    v[a] = v[a] + v[b] + (x);     \
    v[d] = rotr32(v[d] ^ v[a], 16); \
    v[c] = v[c] + v[d];           \
    v[b] = rotr32(v[b] ^ v[c], 12); \
    v[a] = v[a] + v[b] + (y);     \
    v[d] = rotr32(v[d] ^ v[a], 8);  \
    v[c] = v[c] + v[d];           \
    v[b] = rotr32(v[b] ^ v[c], 7);  \
  } while (0)

#define SH_B2_ROUND(s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15) \
  do {                                                                                    \
    SH_B2_G(0, 4, 8, 12, m[s0], m[s1]);                                                   \
    SH_B2_G(1, 5, 9, 13, m[s2], m[s3]);                                                   \
    SH_B2_G(2, 6, 10, 14, m[s4], m[s5]);                                                  \
    SH_B2_G(3, 7, 11, 15, m[s6], m[s7]);                                                  \
    SH_B2_G(0, 5, 10, 15, m[s8], m[s9]);                                                  \
    SH_B2_G(1, 6, 11, 12, m[s10], m[s11]);                                                \
    SH_B2_G(2, 7, 8, 13, m[s12], m[s13]);                                                 \
    SH_B2_G(3, 4, 9, 14, m[s14], m[s15]);                                                 \
  } while (0)

// Compresses one 64-byte block m (16 little-endian words) into h.
// t = bytes hashed so far including this block; last = final-block flag.
SH_HD void b2_compress(uint32_t h[8], const uint32_t m[16], uint32_t t, bool last) {
  uint32_t v[16];
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = h[i];
  v[8] = SH_B2_IV0;
  v[9] = SH_B2_IV1;
  v[10] = SH_B2_IV2;
  v[11] = SH_B2_IV3;
  v[12] = SH_B2_IV4 ^ t;
  v[13] = SH_B2_IV5;  // t >> 32 is zero for every stream hashed here (< 4 GiB)
  v[14] = last ? ~SH_B2_IV6 : SH_B2_IV6;
  v[15] = SH_B2_IV7;
  SH_B2_ROUND(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
  SH_B2_ROUND(14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3);
  SH_B2_ROUND(11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4);
  SH_B2_ROUND(7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8);
  SH_B2_ROUND(9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13);
  SH_B2_ROUND(2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9);
  SH_B2_ROUND(12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11);
  SH_B2_ROUND(13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10);
  SH_B2_ROUND(6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5);
  SH_B2_ROUND(10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0);
#pragma unroll
  for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[i + 8];
}

// Host-side streaming Blake2s over bytes (channel and small fallbacks).
struct B2Stream {
  uint32_t h[8];
  uint8_t buf[64];
  uint32_t buflen;
  uint32_t t;
};
inline void b2s_init(B2Stream* s) {
  b2_init(s->h);
  s->buflen = 0;
  s->t = 0;
}
inline void b2s_block(B2Stream* s, bool last) {
  uint32_t m[16];
  for (int i = 0; i < 16; i++)
    m[i] = (uint32_t)s->buf[4 * i] | ((uint32_t)s->buf[4 * i + 1] << 8) |
           ((uint32_t)s->buf[4 * i + 2] << 16) | ((uint32_t)s->buf[4 * i + 3] << 24);
  b2_compress(s->h, m, s->t, last);
}
inline void b2s_update(B2Stream* s, const void* data, uint64_t len) {
  const uint8_t* in = (const uint8_t*)data;
  while (len > 0) {
    if (s->buflen == 64) {
      s->t += 64;
      b2s_block(s, false);
      s->buflen = 0;
    }
    uint64_t take = 64 - s->buflen;
    if (take > len) take = len;
    for (uint64_t i = 0; i < take; i++) s->buf[s->buflen + i] = in[i];
    s->buflen += (uint32_t)take;
    in += take;
    len -= take;
  }
}
inline void b2s_final(B2Stream* s, uint8_t out[32]) {
  s->t += s->buflen;
  for (uint32_t i = s->buflen; i < 64; i++) s->buf[i] = 0;
  b2s_block(s, true);
  for (int i = 0; i < 8; i++) {
    out[4 * i] = (uint8_t)s->h[i];
    out[4 * i + 1] = (uint8_t)(s->h[i] >> 8);
    out[4 * i + 2] = (uint8_t)(s->h[i] >> 16);
    out[4 * i + 3] = (uint8_t)(s->h[i] >> 24);
  }
}
inline void b2s_hash(const void* data, uint64_t len, uint8_t out[32]) {
  B2Stream s;
  b2s_init(&s);
  b2s_update(&s, data, len);
  b2s_final(&s, out);
}

}  // namespace sh
