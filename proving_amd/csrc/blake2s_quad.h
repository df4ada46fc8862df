// Blake2s compression by the four lanes of a quad (node hashes at the top of Merkle trees and
// the FRI layers too small to fill the device).
//
// A level of up to 2^13 nodes cannot fill the device with one lane per node, so it costs one
// compression's latency (about 984 dependent VALU instructions).  Here four lanes
// share a compression: lane q of a quad holds row entries v[q], v[4+q], v[8+q], v[12+q]
// and runs one G function of each half-round; between the column and the diagonal step the
// b, c, d rows rotate across the quad by 1, 2, 3 lanes (DPP quad permutes) and back after.
// Each lane reads the four message words its G functions use in a round from LDS, through
// per-lane indices (SIGMA[r][2q], SIGMA[r][2q+1], SIGMA[r][8+2q], SIGMA[r][9+2q]).  About
// 300 VALU instructions per node on the critical path instead of 984.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "blake2s.h"

namespace {

struct QuadSigma {
  uint8_t v[4][40];  // lane q, round r: SIGMA[r][2q], SIGMA[r][2q+1], SIGMA[r][8+2q], SIGMA[r][9+2q]
};
constexpr uint8_t kSigma[10][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0}};
constexpr QuadSigma make_quad_sigma() {
  QuadSigma t{};
  for (int q = 0; q < 4; q++)
    for (int r = 0; r < 10; r++) {
      t.v[q][4 * r] = kSigma[r][2 * q];
      t.v[q][4 * r + 1] = kSigma[r][2 * q + 1];
      t.v[q][4 * r + 2] = kSigma[r][8 + 2 * q];
      t.v[q][4 * r + 3] = kSigma[r][9 + 2 * q];
    }
  return t;
}
__constant__ QuadSigma kQuadSigma = make_quad_sigma();

// value of lane (q + k) & 3 of the quad
template <int K>
__device__ __forceinline__ uint32_t quad_rot(uint32_t x) {
  constexpr int ctrl = K == 1 ? 0x39 : K == 2 ? 0x4E : 0x93;  // quad_perm [1,2,3,0] / [2,3,0,1] / [3,0,1,2]
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, ctrl, 0xF, 0xF, false);
}

#define SH_QG(x, y)                 \
  do {                              \
    a = a + b + (x);                \
    d = sh::rotr32(d ^ a, 16);      \
    c = c + d;                      \
    b = sh::rotr32(b ^ c, 12);      \
    a = a + b + (y);                \
    d = sh::rotr32(d ^ a, 8);       \
    c = c + d;                      \
    b = sh::rotr32(b ^ c, 7);       \
  } while (0)

// One node hash by the four lanes of a quad: msg = the 16 message words (LDS), sidx = this
// lane's 40 message indices.  Returns digest words q (lo) and 4 + q (hi).
__device__ __forceinline__ void quad_node_hash(const uint32_t* msg, const uint32_t (&sidx)[40], uint32_t q,
                                               uint32_t h_lo, uint32_t h_hi, uint32_t d0, uint32_t& lo,
                                               uint32_t& hi) {
  uint32_t mw[40];
#pragma unroll
  for (int i = 0; i < 40; i++) mw[i] = msg[sidx[i]];
  uint32_t a = h_lo, b = h_hi, c = q == 0 ? SH_B2_IV0 : q == 1 ? SH_B2_IV1 : q == 2 ? SH_B2_IV2 : SH_B2_IV3, d = d0;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    SH_QG(mw[4 * r], mw[4 * r + 1]);
    b = quad_rot<1>(b);
    c = quad_rot<2>(c);
    d = quad_rot<3>(d);
    SH_QG(mw[4 * r + 2], mw[4 * r + 3]);
    b = quad_rot<3>(b);
    c = quad_rot<2>(c);
    d = quad_rot<1>(d);
  }
  lo = h_lo ^ a ^ c;
  hi = h_hi ^ b ^ d;
}
#undef SH_QG

// The initial chaining value of a one-block hash of t bytes (final block), as a quad holds it:
// h_lo = h[q], h_hi = h[4 + q] (IV with the parameter word), d0 = row 3 entry q.
__device__ __forceinline__ void quad_init(uint32_t q, uint32_t t, uint32_t& h_lo, uint32_t& h_hi, uint32_t& d0) {
  h_lo = q == 0 ? (SH_B2_IV0 ^ 0x01010020u) : q == 1 ? SH_B2_IV1 : q == 2 ? SH_B2_IV2 : SH_B2_IV3;
  h_hi = q == 0 ? SH_B2_IV4 : q == 1 ? SH_B2_IV5 : q == 2 ? SH_B2_IV6 : SH_B2_IV7;
  d0 = q == 0 ? (SH_B2_IV4 ^ t) : q == 1 ? SH_B2_IV5 : q == 2 ? ~SH_B2_IV6 : SH_B2_IV7;
}

// This lane's 40 message indices.
__device__ __forceinline__ void quad_sigma(uint32_t q, uint32_t (&sidx)[40]) {
#pragma unroll
  for (int i = 0; i < 40; i++) sidx[i] = kQuadSigma.v[q][i];
}

}  // namespace
