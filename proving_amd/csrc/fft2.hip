// Circle FFT passes, compile-time specialised (the engine behind fft.hip's batched transforms).
//
// Same algorithm and pass split as fft.hip (CpuBackend's layer order, crates/stwo/src/prover/
// backend/cpu/circle.rs:26-79, 178-238): a transform of log m is a block pass over contiguous
// blocks of 2^G words (layers 0..G-1, the circle layer among them) and strided passes of G layers
// each over tiles of 2^G rows x 2^TL consecutive positions at stride 2^lo.  What differs is the
// data layout inside a pass:
//
// * Quads.  A lane owns 2^R "rows" x 4 consecutive positions (a quad).  A register stage of R
//   layers runs on the rows; the 4 positions of a quad ride along and share every twiddle of the
//   layers above 1 (layers 0 and 1 pair positions inside a quad, so the block pass folds them into
//   its stage over layers 2..R+1).  LDS moves whole quads (128-bit accesses).
// * Immediate addressing.  LDS holds quad Q at quad address Q + (Q >> 4) (one pad quad per 16).
//   The map is additive over disjoint bit sets, so a lane's 2^R quads sit at its own base plus
//   compile-time offsets: every LDS access is one ds_read/write_b128 with an immediate offset and
//   no address arithmetic.  The padding keeps 16-lane groups on distinct banks (2-way at worst).
// * Direct stages.  In a strided pass the lanes of the first stage read their quads straight from
//   HBM into registers (8..64 lanes cover whole 2^TL-word rows, so the reads are full 128-byte
//   lines) and the last stage writes straight back: two LDS round trips fewer per pass.
// * Everything else (stage count, shifts, twiddle offsets, thread count) is a template constant.
#include "common.h"

namespace {

typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) v4u lds_quad;
typedef __attribute__((address_space(1))) v4u gquad;  // global memory (not FLAT)
typedef __attribute__((address_space(1))) uint32_t gword;

constexpr int kPadShift = 4;
__host__ __device__ constexpr uint32_t padq(uint32_t q) { return q + (q >> kPadShift); }
__device__ __forceinline__ lds_quad* lq(uint32_t byte_addr) { return (lds_quad*)(size_t)byte_addr; }

// Butterflies with the twiddle given doubled (t2 = 2t): every twiddle is doubled once where it is
// loaded and then serves 4 * 2^b butterflies, each one instruction cheaper (sh::mmul_dbl).
template <bool INV>
__device__ __forceinline__ void bfly(uint32_t& a, uint32_t& b, uint32_t t2) {
  if constexpr (INV) {  // crates/stwo/src/core/fft.rs:14-21
    const uint32_t s = sh::madd(a, b);
    b = sh::mmul_dbl(sh::msub(a, b), t2);
    a = s;
  } else {  // crates/stwo/src/core/fft.rs:5-12
    const uint32_t m = sh::mmul_dbl(b, t2);
    b = sh::msub(a, m);
    a = sh::madd(a, m);
  }
}

// The R layers of a stage on rows e of the lane: layer b pairs rows e, e | 2^b (bit b clear) with
// twiddle t[off(b) + (e >> (b + 1))], off(b) = 2^R - 2^(R - b).  Inverse runs b upwards.
template <bool INV, int R>
__device__ __forceinline__ void row_layers(uint32_t (&v)[1 << R][4], const uint32_t (&t)[1 << R]) {
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int b = INV ? k : R - 1 - k;
    const int off = (1 << R) - (1 << (R - b));
#pragma unroll
    for (int eh = 0; eh < (1 << (R - 1 - b)); eh++) {
#pragma unroll
      for (int el = 0; el < (1 << b); el++) {
        const int a0 = (eh << (b + 1)) | el, a1 = a0 | (1 << b);
#pragma unroll
        for (int q = 0; q < 4; q++) bfly<INV>(v[a0][q], v[a1][q], t[off + eh]);
      }
    }
  }
}

// Layers 0 (circle) and 1 inside each quad: layer 1 pairs positions (0, 2), (1, 3) with t1[e];
// layer 0 pairs (0, 1) with ta[e] and (2, 3) with tb[e].  Inverse: layer 0 first.
template <bool INV, int R>
__device__ __forceinline__ void quad_layers(uint32_t (&v)[1 << R][4], const uint32_t (&t1)[1 << R],
                                            const uint32_t (&ta)[1 << R], const uint32_t (&tb)[1 << R]) {
#pragma unroll
  for (int e = 0; e < (1 << R); e++) {
    if constexpr (INV) {
      bfly<true>(v[e][0], v[e][1], ta[e]);
      bfly<true>(v[e][2], v[e][3], tb[e]);
      bfly<true>(v[e][0], v[e][2], t1[e]);
      bfly<true>(v[e][1], v[e][3], t1[e]);
    } else {
      bfly<false>(v[e][0], v[e][2], t1[e]);
      bfly<false>(v[e][1], v[e][3], t1[e]);
      bfly<false>(v[e][0], v[e][1], ta[e]);
      bfly<false>(v[e][2], v[e][3], tb[e]);
    }
  }
}

// Twiddle table of line layer i of a domain of log Pl inside a tree of 2^K words (the CpuBackend
// tree layout, outermost layer first).
__device__ __forceinline__ const uint32_t* layer_tab(const uint32_t* tw, uint32_t K, uint32_t Pl, uint32_t i) {
  return tw + ((1u << K) - (1u << (Pl - i)));
}

template <int R>
__device__ __forceinline__ void lds_read(uint32_t (&v)[1 << R][4], uint32_t lane_b, uint32_t qs) {
#pragma unroll
  for (int e = 0; e < (1 << R); e++) {
    const v4u x = *lq(lane_b + 16u * padq((uint32_t)e << qs));
    v[e][0] = x.x;
    v[e][1] = x.y;
    v[e][2] = x.z;
    v[e][3] = x.w;
  }
}
template <int R>
__device__ __forceinline__ void lds_write(const uint32_t (&v)[1 << R][4], uint32_t lane_b, uint32_t qs) {
#pragma unroll
  for (int e = 0; e < (1 << R); e++) *lq(lane_b + 16u * padq((uint32_t)e << qs)) = v4u{v[e][0], v[e][1], v[e][2], v[e][3]};
}

// Lane task tau of a stage whose rows occupy quad-index bits [qs, qs + R): its quad index with
// those bits zero.
__device__ __forceinline__ uint32_t lane_quad(uint32_t tau, uint32_t qs, uint32_t R) {
  return (tau & ((1u << qs) - 1)) | ((tau >> qs) << (qs + R));
}

__device__ __forceinline__ void scale_rows(uint32_t* v, int n, uint32_t scale) {
  const uint32_t s2 = scale << 1;
  for (int i = 0; i < n; i++) v[i] = sh::mmul_dbl(v[i], s2);
}

// ---------------------------------------------------------------- block pass
// Stage k of a block pass covers word bits [2 + kR, 2 + kR + Rk) (rows) plus the quad bits; stage 0
// also runs layers 0 and 1.  NT lanes, each task holds 2^Rk rows.
template <bool INV, int G, int R, int K>
__device__ __forceinline__ void block_stage(uint32_t lds_base, uint32_t base, const uint32_t* tw, uint32_t Kt,
                                            uint32_t Pl) {
  constexpr int S = 2 + K * R;                       // first word bit of the stage's rows
  constexpr int RK = (G - S) < R ? (G - S) : R;      // rows bits of this stage
  constexpr int QS = S - 2;                          // quad-index bit of the first row bit
  constexpr int NT = 1 << (G - 2 - R);
  constexpr int TASKS = 1 << (G - 2 - RK);
#pragma unroll
  for (int it = 0; it < TASKS / NT; it++) {
    const uint32_t tau = threadIdx.x + it * NT;
    const uint32_t q0 = lane_quad(tau, QS, RK);
    const uint32_t lane_b = lds_base + 16u * padq(q0);
    uint32_t v[1 << RK][4];
    lds_read<RK>(v, lane_b, QS);
    const uint32_t wpos = base + (q0 << 2);  // position of the lane's first word
    if constexpr (RK > 0) {
      uint32_t t[1 << RK];
#pragma unroll
      for (int b = 0; b < RK; b++) {
        const uint32_t i = S + b;
        const uint32_t* tab = layer_tab(tw, Kt, Pl, i);
        const uint32_t hb = wpos >> (i + 1);
        const int off = (1 << RK) - (1 << (RK - b));
#pragma unroll
        for (int eh = 0; eh < (1 << (RK - 1 - b)); eh++) t[off + eh] = tab[hb + eh] << 1;
      }
      t[(1 << RK) - 1] = 0;
      if constexpr (K == 0) {
        // layers 1 and 0: t1[e] = layer-1 twiddle of quad wpos/4 + e; the circle twiddles of the
        // quad's two pairs come from the (x, y) pair at layer-1 word 2 * ((wpos/4 + e) >> 1):
        // [y, -y] for an even quad, [-x, x] for an odd one (crates/stwo/src/prover/backend/cpu/
        // circle.rs:306-323)
        const uint32_t* tab1 = layer_tab(tw, Kt, Pl, 1);
        const uint32_t hq = wpos >> 2;  // even: wpos is a multiple of 2^(2 + RK)
        uint32_t t1[1 << RK], ta[1 << RK], tb[1 << RK];
#pragma unroll
        for (int e = 0; e < (1 << RK); e += 2) {
          const uint2 xy = *(const uint2*)(tab1 + hq + e);
          t1[e] = xy.x << 1;  // doubled, as every twiddle here
          t1[e + 1] = xy.y << 1;
          ta[e] = xy.y << 1;
          tb[e] = sh::mneg(xy.y) << 1;
          ta[e + 1] = sh::mneg(xy.x) << 1;
          tb[e + 1] = xy.x << 1;
        }
        if constexpr (INV) {
          quad_layers<true, RK>(v, t1, ta, tb);
          row_layers<true, RK>(v, t);
        } else {
          row_layers<false, RK>(v, t);
          quad_layers<false, RK>(v, t1, ta, tb);
        }
      } else {
        row_layers<INV, RK>(v, t);
      }
    }
    lds_write<RK>(v, lane_b, QS);
  }
}

template <bool INV, int G, int R, int K>
__device__ __forceinline__ void block_stages(uint32_t lds_base, uint32_t base, const uint32_t* tw, uint32_t Kt,
                                             uint32_t Pl) {
  constexpr int NS = (G - 2 + R - 1) / R;
  if constexpr (K < NS) {
    constexpr int KK = INV ? K : NS - 1 - K;
    block_stage<INV, G, R, KK>(lds_base, base, tw, Kt, Pl);
    __syncthreads();
    block_stages<INV, G, R, K + 1>(lds_base, base, tw, Kt, Pl);
  }
}

// One block pass: blockIdx.x = block of 2^G words, blockIdx.y = column.  src positions are masked
// by src_mask (the LDE's replication of its coefficients); scale != 0 multiplies the output.
template <bool INV, int G, int R>
__global__ __launch_bounds__(1 << (G - 2 - R)) void k_fft_block(const uint32_t* const* src_cols,
                                                             uint32_t* const* dst_cols, uint32_t src_mask,
                                                             const uint32_t* __restrict__ tw, uint32_t Kt,
                                                             uint32_t Pl, uint32_t scale) {
  constexpr int NT = 1 << (G - 2 - R);
  constexpr int NQ = 1 << (G - 2);
  __shared__ v4u lds[padq(NQ)];
  const uint32_t lds_base = (uint32_t)(size_t)(lds_quad*)lds;
  const uint32_t base = blockIdx.x << G;
  const gquad* src = (const gquad*)(src_cols[blockIdx.y] + (base & src_mask));
  gquad* dst = (gquad*)(dst_cols[blockIdx.y] + base);
  {
    v4u r[NQ / NT];
#pragma unroll
    for (int k = 0; k < NQ / NT; k++) r[k] = src[threadIdx.x + k * NT];
#pragma unroll
    for (int k = 0; k < NQ / NT; k++) *lq(lds_base + 16u * padq(threadIdx.x + k * NT)) = r[k];
  }
  __syncthreads();
  block_stages<INV, G, R, 0>(lds_base, base, tw, Kt, Pl);
  v4u r[NQ / NT];
#pragma unroll
  for (int k = 0; k < NQ / NT; k++) r[k] = *lq(lds_base + 16u * padq(threadIdx.x + k * NT));
  if (scale) {
    const uint32_t s2 = scale << 1;
#pragma unroll
    for (int k = 0; k < NQ / NT; k++) {
      r[k].x = sh::mmul_dbl(r[k].x, s2);
      r[k].y = sh::mmul_dbl(r[k].y, s2);
      r[k].z = sh::mmul_dbl(r[k].z, s2);
      r[k].w = sh::mmul_dbl(r[k].w, s2);
    }
  }
#pragma unroll
  for (int k = 0; k < NQ / NT; k++) dst[threadIdx.x + k * NT] = r[k];
}

// ---------------------------------------------------------------- strided pass
// Tile: rows j < 2^G at positions base + (j << lo) + t, t < 2^TL; quad index Q = (j << (TL - 2)) | t/4.
// Stage k covers row bits [kR, kR + Rk); the first stage in processing order reads HBM directly,
// the last writes HBM directly.
template <bool INV, int G, int TL, int R, int K, bool FIRST, bool LAST>
__device__ __forceinline__ void strided_stage(uint32_t lds_base, const uint32_t* src, uint32_t src_mask,
                                              uint32_t* dst, uint32_t base, uint32_t g_hi, uint32_t lo,
                                              const uint32_t* tw, uint32_t Kt, uint32_t Pl, uint32_t scale) {
  constexpr int S = K * R;
  constexpr int RK = (G - S) < R ? (G - S) : R;
  constexpr int QS = S + TL - 2;
  constexpr int NT = 1 << (G + TL - 2 - R);
  constexpr int TASKS = 1 << (G + TL - 2 - RK);
#pragma unroll
  for (int it = 0; it < TASKS / NT; it++) {
    const uint32_t tau = threadIdx.x + it * NT;
    const uint32_t q0 = lane_quad(tau, QS, RK);
    const uint32_t lane_b = lds_base + 16u * padq(q0);
    const uint32_t j0 = q0 >> (TL - 2);
    const uint32_t goff = base + (j0 << lo) + ((q0 & ((1u << (TL - 2)) - 1)) << 2);  // word position
    uint32_t v[1 << RK][4];
    if constexpr (FIRST) {
      const uint32_t so = goff & src_mask;
#pragma unroll
      for (int e = 0; e < (1 << RK); e++) {
        const v4u x = *(const gquad*)((const gword*)src + so + ((uint32_t)e << (S + lo)));
        v[e][0] = x.x;
        v[e][1] = x.y;
        v[e][2] = x.z;
        v[e][3] = x.w;
      }
    } else {
      lds_read<RK>(v, lane_b, QS);
    }
    uint32_t t[1 << RK];
#pragma unroll
    for (int b = 0; b < RK; b++) {
      const uint32_t i = lo + S + b;
      const uint32_t* tab = layer_tab(tw, Kt, Pl, i);
      const uint32_t hb = (g_hi << (G - S - b - 1)) + (j0 >> (S + b + 1));
      const int off = (1 << RK) - (1 << (RK - b));
#pragma unroll
      for (int eh = 0; eh < (1 << (RK - 1 - b)); eh++) t[off + eh] = tab[hb + eh] << 1;
    }
    t[(1 << RK) - 1] = 0;
    row_layers<INV, RK>(v, t);
    if constexpr (LAST) {
      if (scale) scale_rows(&v[0][0], 4 << RK, scale);
#pragma unroll
      for (int e = 0; e < (1 << RK); e++)
        *(gquad*)((gword*)dst + goff + ((uint32_t)e << (S + lo))) = v4u{v[e][0], v[e][1], v[e][2], v[e][3]};
    } else {
      lds_write<RK>(v, lane_b, QS);
    }
  }
}

template <bool INV, int G, int TL, int R, int K>
__device__ __forceinline__ void strided_stages(uint32_t lds_base, const uint32_t* src, uint32_t src_mask,
                                               uint32_t* dst, uint32_t base, uint32_t g_hi, uint32_t lo,
                                               const uint32_t* tw, uint32_t Kt, uint32_t Pl, uint32_t scale) {
  constexpr int NS = (G + R - 1) / R;
  if constexpr (K < NS) {
    constexpr int KK = INV ? K : NS - 1 - K;
    strided_stage<INV, G, TL, R, KK, K == 0, K == NS - 1>(lds_base, src, src_mask, dst, base, g_hi, lo, tw, Kt,
                                                           Pl, scale);
    if constexpr (K + 1 < NS) __syncthreads();
    strided_stages<INV, G, TL, R, K + 1>(lds_base, src, src_mask, dst, base, g_hi, lo, tw, Kt, Pl, scale);
  }
}

// One strided pass over layers [lo, lo + G): blockIdx.x = tile (low positions, then high bits),
// blockIdx.y = column.  src may equal dst (in place: every tile reads all its words before any
// write of the same words, and tiles are disjoint).
template <bool INV, int G, int TL, int R>
__global__ __launch_bounds__(1 << (G + TL - 2 - R)) void k_fft_strided(const uint32_t* const* src_cols,
                                                                    uint32_t* const* dst_cols, uint32_t src_mask,
                                                                    const uint32_t* __restrict__ tw, uint32_t Kt,
                                                                    uint32_t Pl, uint32_t lo, uint32_t scale) {
  constexpr int NQ = 1 << (G + TL - 2);
  __shared__ v4u lds[(G + R - 1) / R > 1 ? padq(NQ) : 1];
  const uint32_t lds_base = (uint32_t)(size_t)(lds_quad*)lds;
  const uint32_t lo_tiles_log = lo - TL;
  const uint32_t b = blockIdx.x;
  const uint32_t g_lo0 = (b & ((1u << lo_tiles_log) - 1)) << TL;
  const uint32_t g_hi = b >> lo_tiles_log;
  const uint32_t base = (g_hi << (lo + G)) | g_lo0;
  strided_stages<INV, G, TL, R, 0>(lds_base, src_cols[blockIdx.y], src_mask, dst_cols[blockIdx.y], base, g_hi, lo,
                                   tw, Kt, Pl, scale);
}

// Twiddles of a strided stage's RK layers (row bits [S, S + RK)), as strided_stage loads them.
template <int G, int RK, int S>
__device__ __forceinline__ void stage_twiddles(uint32_t (&t)[1 << RK], const uint32_t* tw, uint32_t Kt, uint32_t Pl,
                                               uint32_t lo, uint32_t g_hi, uint32_t j0) {
#pragma unroll
  for (int b = 0; b < RK; b++) {
    const uint32_t i = lo + S + b;
    const uint32_t* tab = layer_tab(tw, Kt, Pl, i);
    const uint32_t hb = (g_hi << (G - S - b - 1)) + (j0 >> (S + b + 1));
    const int off = (1 << RK) - (1 << (RK - b));
#pragma unroll
    for (int eh = 0; eh < (1 << (RK - 1 - b)); eh++) t[off + eh] = tab[hb + eh] << 1;
  }
  t[(1 << RK) - 1] = 0;
}

// Interpolation stages K .. KEND - 1 of a strided pass (processing order = stage order), the
// first reading src from HBM, all writing the LDS.
template <bool INV, int G, int TL, int R, int K, int KEND>
__device__ __forceinline__ void strided_stages_upto(uint32_t lds_base, const uint32_t* src, uint32_t* dst,
                                                    uint32_t base, uint32_t lo, const uint32_t* tw, uint32_t Kt,
                                                    uint32_t Pl) {
  if constexpr (K < KEND) {
    strided_stage<INV, G, TL, R, K, K == 0, false>(lds_base, src, 0xffffffffu, dst, base, 0u, lo, tw, Kt, Pl, 0u);
    if constexpr (K + 1 < KEND) __syncthreads();
    strided_stages_upto<INV, G, TL, R, K + 1, KEND>(lds_base, src, dst, base, lo, tw, Kt, Pl);
  }
}

// Evaluation stages in processing order K .. NS - 1 (stage NS - 1 - K), from the LDS; the last
// writes dst.
template <int G, int TL, int R, int K>
__device__ __forceinline__ void strided_stages_from(uint32_t lds_base, uint32_t* dst, uint32_t base, uint32_t g_hi,
                                                    uint32_t lo, const uint32_t* tw, uint32_t Kt, uint32_t Pl) {
  constexpr int NS = (G + R - 1) / R;
  if constexpr (K < NS) {
    strided_stage<false, G, TL, R, NS - 1 - K, false, K == NS - 1>(lds_base, nullptr, 0xffffffffu, dst, base, g_hi,
                                                                   lo, tw, Kt, Pl, 0u);
    if constexpr (K + 1 < NS) __syncthreads();
    strided_stages_from<G, TL, R, K + 1>(lds_base, dst, base, g_hi, lo, tw, Kt, Pl);
  }
}

// ---------------------------------------------------------------- interpolation + extension
// A column of 2^m words whose interpolation ends with a strided pass over layers [lo, m) and whose
// extension to 2^(m + nsub_log) words starts with a strided pass over the same layers (the
// extension's layers m.. only replicate the coefficients).  One workgroup per tile of the column:
// the interpolation's stages run as in k_fft_strided<true>; the last one (the pivot, row bits
// [(NS - 1) R, G)) scales its words, writes them as coefficients and keeps them in registers; the
// extension's stages then run once per subdomain h < 2^nsub_log, the first of them (the same row
// bits, so the same lane owns the same quads) from those registers, the last writing the
// subdomain's words at h * 2^m.  Coefficients are read once (as the interpolation's input) and
// written once, never read back for the extension.
template <int G, int TL, int R>
__global__ __launch_bounds__(1 << (G + TL - 2 - R)) void k_fft_strided_ie(
    uint32_t* const* coef_cols, uint32_t* const* ext_cols, const uint32_t* __restrict__ itw,
    const uint32_t* __restrict__ tw, uint32_t Kt, uint32_t m, uint32_t nsub_log, uint32_t lo, uint32_t scale) {
  constexpr int NQ = 1 << (G + TL - 2);
  constexpr int NS = (G + R - 1) / R;
  constexpr int SP = (NS - 1) * R;                  // first row bit of the pivot stage
  constexpr int RK = G - SP;                        // its row bits (1..R)
  constexpr int QS = SP + TL - 2;
  constexpr int NT = 1 << (G + TL - 2 - R);
  constexpr int TASKS = 1 << (G + TL - 2 - RK);
  constexpr int IT = TASKS / NT;
  __shared__ v4u lds[NS > 1 ? padq(NQ) : 1];
  const uint32_t lds_base = (uint32_t)(size_t)(lds_quad*)lds;
  const uint32_t g_lo0 = blockIdx.x << TL;          // lo + G = m: one tile row of high bits
  uint32_t* coef = coef_cols[blockIdx.y];
  uint32_t* ext = ext_cols[blockIdx.y];
  // interpolation, stages 0 .. NS - 2 (the first reads the column from HBM)
  if constexpr (NS > 1) {
    strided_stages_upto<true, G, TL, R, 0, NS - 1>(lds_base, coef, coef, g_lo0, lo, itw, Kt, m);
    __syncthreads();
  }
  uint32_t c[IT][1 << RK][4];
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const uint32_t tau = threadIdx.x + it * NT;
    const uint32_t q0 = lane_quad(tau, QS, RK);
    const uint32_t j0 = q0 >> (TL - 2);
    const uint32_t goff = g_lo0 + (j0 << lo) + ((q0 & ((1u << (TL - 2)) - 1)) << 2);
    uint32_t (&v)[1 << RK][4] = c[it];
    if constexpr (NS > 1) {
      lds_read<RK>(v, lds_base + 16u * padq(q0), QS);
    } else {
#pragma unroll
      for (int e = 0; e < (1 << RK); e++) {
        const v4u x = *(const gquad*)((const gword*)coef + goff + ((uint32_t)e << (SP + lo)));
        v[e][0] = x.x;
        v[e][1] = x.y;
        v[e][2] = x.z;
        v[e][3] = x.w;
      }
    }
    uint32_t t[1 << RK];
    stage_twiddles<G, RK, SP>(t, itw, Kt, m, lo, 0u, j0);
    row_layers<true, RK>(v, t);
    scale_rows(&v[0][0], 4 << RK, scale);
#pragma unroll
    for (int e = 0; e < (1 << RK); e++)
      *(gquad*)((gword*)coef + goff + ((uint32_t)e << (SP + lo))) = v4u{v[e][0], v[e][1], v[e][2], v[e][3]};
  }
  // extension, one subdomain at a time
  const uint32_t Pf = m + nsub_log;
  for (uint32_t h = 0; h < (1u << nsub_log); h++) {
    if (h > 0 && NS > 1) __syncthreads();  // the previous subdomain's last stage has read the LDS
    const uint32_t base = (h << m) | g_lo0;
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const uint32_t tau = threadIdx.x + it * NT;
      const uint32_t q0 = lane_quad(tau, QS, RK);
      const uint32_t j0 = q0 >> (TL - 2);
      uint32_t v[1 << RK][4];
#pragma unroll
      for (int e = 0; e < (1 << RK); e++)
#pragma unroll
        for (int q = 0; q < 4; q++) v[e][q] = c[it][e][q];
      uint32_t t[1 << RK];
      stage_twiddles<G, RK, SP>(t, tw, Kt, Pf, lo, h, j0);
      row_layers<false, RK>(v, t);
      if constexpr (NS > 1) {
        lds_write<RK>(v, lds_base + 16u * padq(q0), QS);
      } else {
        const uint32_t goff = base + (j0 << lo) + ((q0 & ((1u << (TL - 2)) - 1)) << 2);
#pragma unroll
        for (int e = 0; e < (1 << RK); e++)
          *(gquad*)((gword*)ext + goff + ((uint32_t)e << (SP + lo))) = v4u{v[e][0], v[e][1], v[e][2], v[e][3]};
      }
    }
    if constexpr (NS > 1) {
      __syncthreads();
      strided_stages_from<G, TL, R, 1>(lds_base, ext, base, h, lo, tw, Kt, Pf);
    }
  }
}

// ---------------------------------------------------------------- dispatch
typedef void (*block_fn)(const uint32_t* const*, uint32_t* const*, uint32_t, const uint32_t*, uint32_t, uint32_t,
                         uint32_t);
typedef void (*strided_fn)(const uint32_t* const*, uint32_t* const*, uint32_t, const uint32_t*, uint32_t, uint32_t,
                           uint32_t, uint32_t);

template <bool INV, int G, int R>
block_fn block_inst() {
  if constexpr (G >= R + 2 && G - 2 - R <= 10 && G - 2 - R >= 6)
    return k_fft_block<INV, G, R>;
  else
    return nullptr;
}
template <bool INV, int G>
block_fn block_by_r(int r) {
  switch (r) {
    case 1: return block_inst<INV, G, 1>();
    case 2: return block_inst<INV, G, 2>();
    case 3: return block_inst<INV, G, 3>();
    case 4: return block_inst<INV, G, 4>();
    default: return nullptr;
  }
}
template <bool INV>
block_fn block_lookup(int g, int r) {
  switch (g) {
    case 8: return block_by_r<INV, 8>(r);
    case 9: return block_by_r<INV, 9>(r);
    case 10: return block_by_r<INV, 10>(r);
    case 11: return block_by_r<INV, 11>(r);
    case 12: return block_by_r<INV, 12>(r);
    case 13: return block_by_r<INV, 13>(r);
    default: return nullptr;
  }
}

template <int G>
constexpr int strided_tl() { return G >= 5 ? 13 - G : 8; }
template <bool INV, int G, int R>
strided_fn strided_inst() {
  constexpr int TL = strided_tl<G>();
  if constexpr (R <= G && G + TL - 2 - R <= 10 && G + TL - 2 - R >= 6)
    return k_fft_strided<INV, G, TL, R>;
  else
    return nullptr;
}
template <bool INV, int G>
strided_fn strided_by_r(int r) {
  switch (r) {
    case 1: return strided_inst<INV, G, 1>();
    case 2: return strided_inst<INV, G, 2>();
    case 3: return strided_inst<INV, G, 3>();
    case 4: return strided_inst<INV, G, 4>();
    default: return nullptr;
  }
}
template <bool INV>
strided_fn strided_lookup(int g, int r) {
  switch (g) {
    case 1: return strided_by_r<INV, 1>(r);
    case 2: return strided_by_r<INV, 2>(r);
    case 3: return strided_by_r<INV, 3>(r);
    case 4: return strided_by_r<INV, 4>(r);
    case 5: return strided_by_r<INV, 5>(r);
    case 6: return strided_by_r<INV, 6>(r);
    case 7: return strided_by_r<INV, 7>(r);
    case 8: return strided_by_r<INV, 8>(r);
    default: return nullptr;
  }
}
uint32_t strided_tl_host(uint32_t g) { return g >= 5 ? 13 - g : 8; }

typedef void (*ie_fn)(uint32_t* const*, uint32_t* const*, const uint32_t*, const uint32_t*, uint32_t, uint32_t,
                      uint32_t, uint32_t, uint32_t);
template <int G, int R>
ie_fn ie_inst() {
  constexpr int TL = strided_tl<G>();
  if constexpr (R <= G && G + TL - 2 - R <= 10 && G + TL - 2 - R >= 6)
    return k_fft_strided_ie<G, TL, R>;
  else
    return nullptr;
}
template <int G>
ie_fn ie_by_r(int r) {
  switch (r) {
    case 1: return ie_inst<G, 1>();
    case 2: return ie_inst<G, 2>();
    case 3: return ie_inst<G, 3>();
    case 4: return ie_inst<G, 4>();
    default: return nullptr;
  }
}
ie_fn ie_lookup(int g, int r) {
  switch (g) {
    case 4: return ie_by_r<4>(r);
    case 5: return ie_by_r<5>(r);
    case 6: return ie_by_r<6>(r);
    case 7: return ie_by_r<7>(r);
    case 8: return ie_by_r<8>(r);
    default: return nullptr;
  }
}

}  // namespace

// Launches one block pass (layers [0, g)) over n_cols columns of 2^m words, or returns
// SH_ERR_NOT_FOUND (nothing launched) when no specialised kernel has that shape.
int sh_fft2_block(sh_ctx* ctx, bool inverse, uint32_t g, uint32_t r, const uint32_t* const* d_src,
                  uint32_t* const* d_dst, uint32_t n_cols, uint32_t m, uint32_t src_mask, const uint32_t* tw,
                  uint32_t K, uint32_t Pl, uint32_t scale) {
  block_fn f = inverse ? block_lookup<true>((int)g, (int)r) : block_lookup<false>((int)g, (int)r);
  if (!f || g > m) return SH_ERR_NOT_FOUND;
  dim3 grid(1u << (m - g), n_cols);
  hipLaunchKernelGGL(f, grid, dim3(1u << (g - 2 - r)), 0, ctx->stream, d_src, d_dst, src_mask, tw, K, Pl, scale);
  return sh_launch_check(ctx, "fft_block");
}

// One strided pass over layers [lo, lo + g).
int sh_fft2_strided(sh_ctx* ctx, bool inverse, uint32_t g, uint32_t r, uint32_t lo, const uint32_t* const* d_src,
                    uint32_t* const* d_dst, uint32_t n_cols, uint32_t m, uint32_t src_mask, const uint32_t* tw,
                    uint32_t K, uint32_t Pl, uint32_t scale) {
  strided_fn f = inverse ? strided_lookup<true>((int)g, (int)r) : strided_lookup<false>((int)g, (int)r);
  const uint32_t tl = strided_tl_host(g);
  if (!f || lo < tl || lo + g > m) return SH_ERR_NOT_FOUND;
  dim3 grid(1u << (m - g - tl), n_cols);
  hipLaunchKernelGGL(f, grid, dim3(1u << (g + tl - 2 - r)), 0, ctx->stream, d_src, d_dst, src_mask, tw, K, Pl, lo,
                     scale);
  return sh_launch_check(ctx, "fft_strided");
}

// The interpolation's last strided pass (layers [lo, m), scaled by `scale`) fused with the first
// strided pass of the extension of the same columns to 2^(m + nsub_log) words: d_coef holds the
// interpolation's intermediate and receives the coefficients, d_ext receives the extension's
// intermediate.  SH_ERR_NOT_FOUND (nothing launched) when no kernel has that shape.
int sh_fft2_strided_ie(sh_ctx* ctx, uint32_t g, uint32_t r, uint32_t lo, uint32_t* const* d_coef,
                       uint32_t* const* d_ext, uint32_t n_cols, uint32_t m, uint32_t nsub_log, const uint32_t* itw,
                       const uint32_t* tw, uint32_t K, uint32_t scale) {
  ie_fn f = ie_lookup((int)g, (int)r);
  const uint32_t tl = strided_tl_host(g);
  if (!f || lo < tl || lo + g != m || nsub_log == 0 || m + nsub_log > 31) return SH_ERR_NOT_FOUND;
  dim3 grid(1u << (m - g - tl), n_cols);
  hipLaunchKernelGGL(f, grid, dim3(1u << (g + tl - 2 - r)), 0, ctx->stream, d_coef, d_ext, itw, tw, K, m, nsub_log,
                     lo, scale);
  return sh_launch_check(ctx, "fft_strided_ie");
}
