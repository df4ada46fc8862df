// Circle FFT: interpolate (iFFT) and evaluate (FFT / low-degree extension).
//
// Algorithm (CpuBackend, crates/stwo/src/prover/backend/cpu/circle.rs:26-79,
// 178-238): on a circle domain of log size m, evaluations in bit-reversed
// order go through the circle layer (layer 0, pairs (2h, 2h+1), twiddle from
// the y-coordinates) and line layers 1..m-1 (pairs at distance 2^i, twiddle
// h = pos >> (i+1) of the line layer table), then a scale by 2^-m.  Evaluate
// runs the layers in reverse with butterfly instead of ibutterfly.
//
// Evaluating coefficients of length 2^n on a domain of log m > n: layers
// i >= n of the zero-padded transform only copy (v1 = 0), so the output is
// the coefficient vector replicated 2^(m-n) times followed by layers n-1..0 of
// the size-2^m transform.  This is the SIMD backend's subdomain evaluation
// (crates/stwo/src/prover/backend/simd/fft/rfft.rs:121-260) reading each
// coefficient once per subdomain.
//
// GPU structure: a transform of log m is split into passes, each covering a
// contiguous range of layers [lo, hi) held in LDS.  The first pass works on
// contiguous blocks of up to 2^13 words (layers 0..12); later passes work on
// groups of 2^(hi-lo) words at stride 2^lo, 32 consecutive groups per
// workgroup so that every global access is a 128-byte coalesced row.  One
// launch covers every column of one size class (blockIdx.y = column).
#include <algorithm>

#include "common.h"

namespace {

constexpr uint32_t kThreads = 256;

struct FftPass {
  uint32_t lo, hi, tl;
};

// Layers [0, top) split into passes, ascending: a contiguous block pass of up
// to 2^block_log words, then strided passes of stride_layers layers.
std::vector<FftPass> plan_passes(uint32_t top, const sh_ctx::FftTune& t) {
  std::vector<FftPass> v;
  // One table entry serves both directions of a size: an LDE plans its 2^(m-1) computed layers,
  // an interpolation all m, so an entry tuned for one can leave the other a last strided pass of
  // one or two layers (a whole pass over the array whose workgroups are mostly idle).  Such a
  // remainder goes into the block pass while the block stays within 2^13 words.
  // Otherwise, with two or more strided passes, the last two share their layers evenly.
  uint32_t block = t.block_log;
  uint32_t rem = 0;
  if (top > block) {
    rem = (top - block) % t.stride_layers;
    if (rem != 0 && rem <= 2 && block + rem <= 13) {
      block += rem;
      rem = 0;
    }
  }
  uint32_t bh = top < block ? top : block;
  v.push_back({0, bh, 0});
  uint32_t l = bh;
  while (l < top) {
    uint32_t g = top - l < t.stride_layers ? top - l : t.stride_layers;
    if (rem != 0 && rem <= 2 && top - l > t.stride_layers && top - l < 2 * t.stride_layers)
      g = (top - l + 1) / 2;  // the last two passes: ceil and floor of their layers
    // tile of the generic kernel (the fallback): up to 32 groups, within 2^13 words
    const uint32_t tl = std::min<uint32_t>(std::min<uint32_t>(l, 5u), 13u - g);
    v.push_back({l, l + g, tl});
    l += g;
  }
  return v;
}

// LDS word-address swizzle: bits 2-5 ^= bits 6-9.  It permutes words inside
// each 64-word row and keeps aligned 4-word quads intact (so 128-bit LDS
// accesses still work), and it is linear over XOR, so the address of a lane's
// element e is sw(lane part) ^ sw(uniform part).  Without it, lanes whose
// addresses differ by multiples of 64 words hit the same bank.
__device__ __forceinline__ uint32_t sw(uint32_t a) { return a ^ (((a >> 6) & 15u) << 2); }

// LDS words by byte address in the LDS address space.  The pass kernels declare no static LDS, so
// their dynamic array starts at byte 0 and a word's byte address is 4 * its index: the XOR of a
// lane's address with a stage's constant element offset is then the whole address computation.
typedef uint32_t v4u __attribute__((ext_vector_type(4)));  // plain vector: address-space qualifiable
typedef __attribute__((address_space(1))) uint32_t gword;    // global memory (not FLAT)
typedef __attribute__((address_space(1))) v4u gquad;
constexpr int kMaxQuads = 8;  // 2^13 words (the largest tile sh_set_fft_tuning allows) / (4 * kThreads)
typedef __attribute__((address_space(3))) uint32_t lds_word;
typedef __attribute__((address_space(3))) v4u lds_quad;
__device__ __forceinline__ lds_word* lds_at(uint32_t byte_addr) { return (lds_word*)(size_t)byte_addr; }
__device__ __forceinline__ lds_quad* lds_quad_at(uint32_t byte_addr) { return (lds_quad*)(size_t)byte_addr; }

__device__ __forceinline__ uint32_t circle_twiddle(const uint32_t* l1, uint32_t h) {
  // crates/stwo/src/prover/backend/cpu/circle.rs:306-323: [x, y] -> [y, -y, -x, x];
  // branch-free (lanes of one wave take different cases), one 64-bit load
  const uint2 xy = ((const uint2*)l1)[h >> 2];
  const uint32_t v = (h & 2) ? xy.x : xy.y;
  return ((h ^ (h >> 1)) & 1) ? sh::mneg(v) : v;
}

// Butterflies of one register stage: R consecutive local layers [s, s+R) of
// the group.  A lane owns a subgroup of 2^R words whose j differ only in bits
// [s, s+R) (the words every butterfly of these layers connects), loads them
// from LDS once, runs the R layers in registers and writes them back: one
// LDS access per butterfly for R = 4 instead of four.  At layer i the twiddle
// index h = pos >> (i+1) is shared by the 2^b pairs below the pair bit b, so a
// stage loads 2^R - 1 twiddles for R * 2^(R-1) butterflies.
template <bool INVERSE, bool CIRCLE, int R>
__device__ __forceinline__ void fft_stage(uint32_t* lds, uint32_t G, uint32_t tl, uint32_t s, uint32_t base,
                                          uint32_t lo, const uint32_t* __restrict__ tw, uint32_t tw_end,
                                          uint32_t Pl) {
  const uint32_t n_sub = 1u << (G + tl - R);
  const uint32_t tmask = (1u << tl) - 1;
  for (uint32_t q = threadIdx.x; q < n_sub; q += kThreads) {
    const uint32_t t = q & tmask;
    const uint32_t jr = q >> tl;
    const uint32_t jf = ((jr >> s) << (s + R)) | (jr & ((1u << s) - 1));
    uint32_t v[1 << R];
    // byte addresses: the element offsets XOR straight into the address the LDS access takes
    // (one VALU op per access instead of an XOR and a scale)
    uint32_t lane_b = sw((jf << tl) | t) << 2;
    // opaque to the optimiser: otherwise (x << 2) ^ (c << 2) is refactored into ((x ^ c) << 2),
    // one more op per access
    asm("" : "+v"(lane_b));
    // s == 0 in a contiguous block: the lane's 2^R words are adjacent, so they
    // move as 128-bit LDS accesses instead of 2^R lanes-strided words (which
    // would serialise on a few banks)
    const bool vec = s == 0 && tl == 0;
    if constexpr (R >= 2) {
      if (vec) {
#pragma unroll
        for (int e = 0; e < (1 << R); e += 4) {
          v4u q4 = *lds_quad_at(lane_b ^ (sw((uint32_t)e) << 2));
          v[e] = q4.x;
          v[e + 1] = q4.y;
          v[e + 2] = q4.z;
          v[e + 3] = q4.w;
        }
      }
    }
    if (!vec || R < 2) {
#pragma unroll
      for (int e = 0; e < (1 << R); e++)
        v[e] = *lds_at(lane_b ^ (sw((uint32_t)e << (s + tl)) << 2));
    }
    // position of the lane's first word; the pair (eh, el) of layer i = lo+s+b
    // sits at pb + eh * 2^(i+1) + el * 2^(lo+s), since jf is zero in bits
    // [s, s+R), so its twiddle index is (pb >> (i+1)) + eh: one address per layer
    const uint32_t pb = base + (jf << lo) + t;
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int b = INVERSE ? k : R - 1 - k;
      const uint32_t i = lo + s + b;
      // CIRCLE: this stage holds layer 0 (lo = s = 0), the circle layer, read
      // from the pairs of line layer 1
      const bool circ = CIRCLE && b == 0;
      const uint32_t* tab = tw + tw_end - (1u << (Pl - (circ ? 1 : i)));
      const uint32_t hb = pb >> (i + 1);
      const uint32_t* tp = tab + hb;  // twiddle eh at a constant offset
#pragma unroll
      for (int eh = 0; eh < (1 << (R - 1 - b)); eh++) {
        const uint32_t twv = circ ? circle_twiddle(tab, hb + eh) : tp[eh];
#pragma unroll
        for (int el = 0; el < (1 << b); el++) {
          const int a0 = (eh << (b + 1)) | el, a1 = a0 | (1 << b);
          if (INVERSE) {  // crates/stwo/src/core/fft.rs:14-21
            uint32_t x = v[a0], y = v[a1];
            v[a0] = sh::madd(x, y);
            v[a1] = sh::mmul(sh::msub(x, y), twv);
          } else {  // crates/stwo/src/core/fft.rs:5-12
            uint32_t m = sh::mmul(v[a1], twv);
            uint32_t x = v[a0];
            v[a0] = sh::madd(x, m);
            v[a1] = sh::msub(x, m);
          }
        }
      }
    }
    if constexpr (R >= 2) {
      if (vec) {
#pragma unroll
        for (int e = 0; e < (1 << R); e += 4)
          *lds_quad_at(lane_b ^ (sw((uint32_t)e) << 2)) = v4u{v[e], v[e + 1], v[e + 2], v[e + 3]};
      }
    }
    if (!vec || R < 2) {
#pragma unroll
      for (int e = 0; e < (1 << R); e++) *lds_at(lane_b ^ (sw((uint32_t)e << (s + tl)) << 2)) = v[e];
    }
  }
}

// A tile of NQ quads per lane between HBM and LDS: every global load of the lane is issued before
// the first LDS write (and every LDS read before the first global store), so a lane waits for
// memory once per tile.  Position of tile word w: base + ((w >> tl) << lo) + (w & tmask).
template <int NQ>
__device__ __forceinline__ void tile_to_lds(const gword* __restrict__ src, uint32_t base, uint32_t lo, uint32_t tl,
                                            uint32_t tmask, uint32_t src_mask) {
  v4u r[NQ];
#pragma unroll
  for (int k = 0; k < NQ; k++) {
    const uint32_t w = 4 * (threadIdx.x + k * kThreads);
    r[k] = *(const gquad*)(src + ((base + ((w >> tl) << lo) + (w & tmask)) & src_mask));
  }
#pragma unroll
  for (int k = 0; k < NQ; k++) *lds_quad_at(sw(4 * (threadIdx.x + k * kThreads)) << 2) = r[k];
}
template <int NQ>
__device__ __forceinline__ void tile_from_lds(gword* __restrict__ dst, uint32_t base, uint32_t lo, uint32_t tl,
                                              uint32_t tmask, uint32_t scale) {
  v4u r[NQ];
#pragma unroll
  for (int k = 0; k < NQ; k++) r[k] = *lds_quad_at(sw(4 * (threadIdx.x + k * kThreads)) << 2);
#pragma unroll
  for (int k = 0; k < NQ; k++) {
    v4u v = r[k];
    if (scale) {
      v.x = sh::mmul(v.x, scale);
      v.y = sh::mmul(v.y, scale);
      v.z = sh::mmul(v.z, scale);
      v.w = sh::mmul(v.w, scale);
    }
    const uint32_t w = 4 * (threadIdx.x + k * kThreads);
    *(gquad*)(dst + base + ((w >> tl) << lo) + (w & tmask)) = v;
  }
}

// One pass of layers [lo, hi) over groups at stride 2^lo, 2^tl consecutive
// groups per workgroup.  Layer i's table starts at tw + 2^K - 2^(Pl - i)
// (Pl = log size of the domain whose tree layers are used; the transform of
// a subdomain uses prefixes of its parent's layers).  Global words move as
// 128-bit rows whenever four consecutive LDS words are four consecutive
// positions (tile of >= 4 groups, or a contiguous block).
// One register stage of r layers, r <= MAXR (a kernel compiled for a smaller MAXR holds fewer
// values in registers and runs at a higher occupancy).
template <bool INVERSE, bool CIRCLE, int MAXR>
__device__ __forceinline__ void run_stage(uint32_t r, uint32_t* lds, uint32_t G, uint32_t tl, uint32_t s,
                                          uint32_t base, uint32_t lo, const uint32_t* __restrict__ tw,
                                          uint32_t tw_end, uint32_t Pl) {
  if constexpr (MAXR >= 5) {
    if (r == 5) return fft_stage<INVERSE, CIRCLE, 5>(lds, G, tl, s, base, lo, tw, tw_end, Pl);
  }
  if constexpr (MAXR >= 4) {
    if (r == 4) return fft_stage<INVERSE, CIRCLE, 4>(lds, G, tl, s, base, lo, tw, tw_end, Pl);
  }
  if (r == 3) return fft_stage<INVERSE, CIRCLE, 3>(lds, G, tl, s, base, lo, tw, tw_end, Pl);
  if (r == 2) return fft_stage<INVERSE, CIRCLE, 2>(lds, G, tl, s, base, lo, tw, tw_end, Pl);
  fft_stage<INVERSE, CIRCLE, 1>(lds, G, tl, s, base, lo, tw, tw_end, Pl);
}

template <bool INVERSE, int MAXR>
__global__ __launch_bounds__(kThreads) void k_fft_pass(const uint32_t* const* src_cols,
                                                       uint32_t* const* dst_cols, uint32_t src_mask,
                                                       const uint32_t* __restrict__ tw, uint32_t K,
                                                       uint32_t Pl, uint32_t lo, uint32_t hi,
                                                       uint32_t tl, uint32_t scale, uint32_t max_stage) {
  extern __shared__ uint32_t lds[];
  const uint32_t G = hi - lo;
  const uint32_t col = blockIdx.y;
  const uint32_t* __restrict__ src = src_cols[col];
  uint32_t* __restrict__ dst = dst_cols[col];
  const uint32_t lo_tiles_log = lo - tl;
  const uint32_t b = blockIdx.x;
  const uint32_t g_lo0 = (b & ((1u << lo_tiles_log) - 1)) << tl;
  const uint32_t g_hi = b >> lo_tiles_log;
  const uint32_t base = (g_hi << hi) | g_lo0;
  const uint32_t n_words = 1u << (G + tl);
  const uint32_t tmask = (1u << tl) - 1;
  const bool vec = (tl >= 2 || lo == 0) && G + tl >= 2 && src_mask >= 3 &&
                   ((((uintptr_t)src) | ((uintptr_t)dst)) & 15) == 0;

  // Tile load: a lane's quads are loaded in batches of four, each batch issued before its first
  // LDS write (a tile is at most 2^13 words, 8 quads per lane), from the global address space
  // (no FLAT counter waits).
  const gword* gsrc = (const gword*)src;
  // whole quads per lane (tiles of 2^10..2^13 words) take the straight-line path
  const uint32_t nq = n_words >> 10;
  if (vec && nq == 4) {
    tile_to_lds<4>(gsrc, base, lo, tl, tmask, src_mask);
  } else if (vec && nq == 8) {
    tile_to_lds<8>(gsrc, base, lo, tl, tmask, src_mask);
  } else if (vec && nq == 2) {
    tile_to_lds<2>(gsrc, base, lo, tl, tmask, src_mask);
  } else if (vec && nq == 1) {
    tile_to_lds<1>(gsrc, base, lo, tl, tmask, src_mask);
  } else if (vec) {
    for (uint32_t w = 4 * threadIdx.x; w < n_words; w += 4 * kThreads)
      *lds_quad_at(sw(w) << 2) = *(const gquad*)(gsrc + ((base + ((w >> tl) << lo) + (w & tmask)) & src_mask));
  } else {
    for (uint32_t w = threadIdx.x; w < n_words; w += kThreads)
      *lds_at(sw(w) << 2) = gsrc[(base + ((w >> tl) << lo) + (w & tmask)) & src_mask];
  }
  __syncthreads();

  const uint32_t tw_end = 1u << K;
  // G layers in stages of at most max_stage, as even as possible (13 = 5 + 4 + 4);
  // inverse runs them upwards, forward downwards
  const uint32_t n_stages = (G + max_stage - 1) / max_stage;
  const uint32_t st_base = G / n_stages, st_extra = G % n_stages;
  for (uint32_t st = 0; st < n_stages; st++) {
    const uint32_t si = INVERSE ? st : n_stages - 1 - st;
    const uint32_t s = si * st_base + (si < st_extra ? si : st_extra);
    const uint32_t r = st_base + (si < st_extra ? 1 : 0);
    if (lo + s == 0)
      run_stage<INVERSE, true, MAXR>(r, lds, G, tl, s, base, lo, tw, tw_end, Pl);
    else
      run_stage<INVERSE, false, MAXR>(r, lds, G, tl, s, base, lo, tw, tw_end, Pl);
    __syncthreads();
  }

  gword* gdst = (gword*)dst;
  if (vec && nq == 4) {
    tile_from_lds<4>(gdst, base, lo, tl, tmask, scale);
  } else if (vec && nq == 8) {
    tile_from_lds<8>(gdst, base, lo, tl, tmask, scale);
  } else if (vec && nq == 2) {
    tile_from_lds<2>(gdst, base, lo, tl, tmask, scale);
  } else if (vec && nq == 1) {
    tile_from_lds<1>(gdst, base, lo, tl, tmask, scale);
  } else if (vec) {
    for (uint32_t w = 4 * threadIdx.x; w < n_words; w += 4 * kThreads) {
      {
        v4u v = *lds_quad_at(sw(w) << 2);
        if (scale) {
          v.x = sh::mmul(v.x, scale);
          v.y = sh::mmul(v.y, scale);
          v.z = sh::mmul(v.z, scale);
          v.w = sh::mmul(v.w, scale);
        }
        *(gquad*)(gdst + base + ((w >> tl) << lo) + (w & tmask)) = v;
      }
    }
  } else {
    for (uint32_t w = threadIdx.x; w < n_words; w += kThreads) {
      const uint32_t v = *lds_at(sw(w) << 2);
      gdst[base + ((w >> tl) << lo) + (w & tmask)] = scale ? sh::mmul(v, scale) : v;
    }
  }
}

// Domains of log size 1 and 2 (crates/stwo/src/prover/backend/cpu/circle.rs:33-58,
// 203-223); c0..c2 are host-computed constants, one lane per column.
// For evaluation, coeff_log = n < m: the layers i >= n only copy, so the
// replicated input skips their butterflies.
__global__ void k_fft_tiny(const uint32_t* const* src_cols, uint32_t* const* dst_cols,
                           uint32_t src_mask, uint32_t n_cols, uint32_t m, int inverse, uint32_t c0,
                           uint32_t c1, uint32_t c2, uint32_t n) {
  uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n_cols) return;
  const uint32_t* src = src_cols[col];
  uint32_t* dst = dst_cols[col];
  if (m == 1) {
    uint32_t v0 = src[0 & src_mask], v1 = src[1 & src_mask];
    if (inverse) {  // c0 = y_inv, c1 = n_inv
      uint32_t a = sh::madd(v0, v1), b = sh::mmul(sh::msub(v0, v1), c0);
      dst[0] = sh::mmul(a, c1);
      dst[1] = sh::mmul(b, c1);
    } else {  // c0 = y
      if (n >= 1) {
        uint32_t t = sh::mmul(v1, c0);
        dst[0] = sh::madd(v0, t);
        dst[1] = sh::msub(v0, t);
      } else {
        dst[0] = v0;
        dst[1] = v1;
      }
    }
    return;
  }
  uint32_t v[4];
  for (int k = 0; k < 4; k++) v[k] = src[k & src_mask];
  if (inverse) {  // c0 = x_inv, c1 = y_inv, c2 = n_inv
    auto ib = [](uint32_t& a, uint32_t& b, uint32_t t) {
      uint32_t s = sh::madd(a, b);
      b = sh::mmul(sh::msub(a, b), t);
      a = s;
    };
    ib(v[0], v[1], c1);
    ib(v[2], v[3], sh::mneg(c1));
    ib(v[0], v[2], c0);
    ib(v[1], v[3], c0);
    for (int k = 0; k < 4; k++) dst[k] = sh::mmul(v[k], c2);
  } else {  // c0 = x, c1 = y
    auto bf = [](uint32_t& a, uint32_t& b, uint32_t t) {
      uint32_t p = sh::mmul(b, t);
      b = sh::msub(a, p);
      a = sh::madd(a, p);
    };
    if (n >= 2) {
      bf(v[0], v[2], c0);
      bf(v[1], v[3], c0);
    }
    if (n >= 1) {
      bf(v[0], v[1], c1);
      bf(v[2], v[3], sh::mneg(c1));
    }
    for (int k = 0; k < 4; k++) dst[k] = v[k];
  }
}

// Replicates a constant polynomial (coeff_log = 0) over the domain.
__global__ void k_replicate(const uint32_t* const* src_cols, uint32_t* const* dst_cols, uint32_t m) {
  uint32_t col = blockIdx.y;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (1u << m)) dst_cols[col][i] = src_cols[col][0];
}

int check_domain(sh_ctx* ctx, const sh_twiddles* tw, uint32_t m) {
  SH_REQUIRE(ctx, m >= 1 && m <= 30, "circle domain log size %u out of range [1, 30]", m);
  // CanonicCoset(m).circle_domain().half_coset = half_odds(m - 1)
  SH_REQUIRE(ctx, sh::is_doubling_of(sh::subgroup_gen(m + 1), m - 1, tw->root_initial_index, tw->root_log_size),
             "domain of log size %u is not a doubling of the twiddle root coset (log %u)", m,
             tw->root_log_size);
  return SH_OK;
}

}  // namespace

int sh_fft2_block(sh_ctx* ctx, bool inverse, uint32_t g, uint32_t r, const uint32_t* const* d_src,
                  uint32_t* const* d_dst, uint32_t n_cols, uint32_t m, uint32_t src_mask, const uint32_t* tw,
                  uint32_t K, uint32_t Pl, uint32_t scale);
int sh_fft2_strided(sh_ctx* ctx, bool inverse, uint32_t g, uint32_t r, uint32_t lo, const uint32_t* const* d_src,
                    uint32_t* const* d_dst, uint32_t n_cols, uint32_t m, uint32_t src_mask, const uint32_t* tw,
                    uint32_t K, uint32_t Pl, uint32_t scale);

int sh_fft2_strided_ie(sh_ctx* ctx, uint32_t g, uint32_t r, uint32_t lo, uint32_t* const* d_coef,
                       uint32_t* const* d_ext, uint32_t n_cols, uint32_t m, uint32_t nsub_log, const uint32_t* itw,
                       const uint32_t* tw, uint32_t K, uint32_t scale);

// One pass with the quad-layout kernels of fft2.hip (the tuned rows-per-lane first, then the
// others); SH_ERR_NOT_FOUND when none of them has this shape.
static int fft2_pass(sh_ctx* ctx, bool inverse, const FftPass& p, const sh_ctx::FftTune& t,
                     const uint32_t* const* s, uint32_t* const* d, uint32_t n_cols, uint32_t m, uint32_t src_mask,
                     const uint32_t* tw, uint32_t K, uint32_t Pl, uint32_t scale) {
  if (!ctx->fft2) return SH_ERR_NOT_FOUND;
  const uint32_t g = p.hi - p.lo;
  const uint32_t first = p.lo == 0 ? t.r_block : t.r_stride;
  const uint32_t order[5] = {first, 3, 2, 4, 1};
  for (uint32_t k = 0; k < 5; k++) {
    if (k > 0 && order[k] == first) continue;
    const int rc = p.lo == 0 ? sh_fft2_block(ctx, inverse, g, order[k], s, d, n_cols, m, src_mask, tw, K, Pl, scale)
                             : sh_fft2_strided(ctx, inverse, g, order[k], p.lo, s, d, n_cols, m, src_mask, tw, K, Pl,
                                               scale);
    if (rc != SH_ERR_NOT_FOUND) return rc;
  }
  return SH_ERR_NOT_FOUND;
}

// One interpolation pass (layers [p.lo, p.hi) of a column of log m, domain of log Pl).
static int ifft_pass(sh_ctx* ctx, const sh_twiddles* tw, const FftPass& p, const sh_ctx::FftTune& tune,
                     const uint32_t* const* s, uint32_t* const* d, uint32_t n_cols, uint32_t m, uint32_t Pl,
                     uint32_t scale) {
  const uint32_t G = p.hi - p.lo;
  SpanGuard span(ctx, "fft_ifft_pass");
  const int rc2 = fft2_pass(ctx, true, p, tune, s, d, n_cols, m, 0xffffffffu, tw->itw, tw->root_log_size, Pl, scale);
  if (rc2 != SH_ERR_NOT_FOUND) return rc2;
  // the generic kernel: shapes with no quad-layout specialisation (blocks below 2^9 words)
  dim3 grid(1u << (m - G - p.tl), n_cols);
  const size_t lds = 4u << (G + p.tl);
  k_fft_pass<true, 4><<<grid, kThreads, lds, ctx->stream>>>(s, d, 0xffffffffu, tw->itw, tw->root_log_size, Pl, p.lo,
                                                            p.hi, p.tl, scale, 4u);
  return sh_launch_check(ctx, "ifft_pass");
}

// One evaluation pass over a column of log m (reads through src_mask: the first pass replicates the
// coefficients).
static int rfft_pass(sh_ctx* ctx, const sh_twiddles* tw, const FftPass& p, const sh_ctx::FftTune& tune,
                     const uint32_t* const* s, uint32_t* const* d, uint32_t n_cols, uint32_t m, uint32_t Pl,
                     uint32_t src_mask) {
  const uint32_t G = p.hi - p.lo;
  SpanGuard span(ctx, "fft_rfft_pass");
  const int rc2 = fft2_pass(ctx, false, p, tune, s, d, n_cols, m, src_mask, tw->tw, tw->root_log_size, Pl, 0u);
  if (rc2 != SH_ERR_NOT_FOUND) return rc2;
  dim3 grid(1u << (m - G - p.tl), n_cols);
  const size_t lds = 4u << (G + p.tl);
  k_fft_pass<false, 4><<<grid, kThreads, lds, ctx->stream>>>(s, d, src_mask, tw->tw, tw->root_log_size, Pl, p.lo,
                                                             p.hi, p.tl, 0u, 4u);
  return sh_launch_check(ctx, "rfft_pass");
}

// Interpolation of n_cols columns of log size m, all passes; src/dst are
// device tables of column pointers (src may equal dst: in place).
// Pl and half_init describe the domain: canonic domains pass Pl = m.
int sh_fft_interpolate_batch(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_src,
                             uint32_t* const* d_dst, uint32_t n_cols, uint32_t m, uint32_t Pl,
                             uint32_t half_init) {
  if (n_cols == 0) return SH_OK;
  if (m <= 2) {
    SpanGuard span(ctx, "fft_tiny");
    sh::mpt p = sh::index_to_point_host(half_init);
    uint32_t c0, c1, c2 = 0;
    if (m == 1) {
      uint32_t yn_inv = sh::minv(sh::mmul(p.y, 2));
      c0 = sh::mmul(yn_inv, 2);
      c1 = sh::mmul(yn_inv, p.y);
    } else {
      uint32_t xyn_inv = sh::minv(sh::mmul(sh::mmul(p.x, p.y), 4));
      c0 = sh::mmul(sh::mmul(xyn_inv, p.y), 4);
      c1 = sh::mmul(sh::mmul(xyn_inv, p.x), 4);
      c2 = sh::mmul(sh::mmul(xyn_inv, p.x), p.y);
    }
    k_fft_tiny<<<(n_cols + 63) / 64, 64, 0, ctx->stream>>>(d_src, d_dst, 0xffffffffu, n_cols, m, 1,
                                                           c0, c1, c2, m);
    return sh_launch_check(ctx, "fft_tiny");
  }
  uint32_t scale = sh::minv(1u << m);
  const sh_ctx::FftTune tune = ctx->fft_tune[m];
  auto passes = plan_passes(m, tune);
  for (size_t k = 0; k < passes.size(); k++)
    SH_TRY(ifft_pass(ctx, tw, passes[k], tune, k == 0 ? d_src : (const uint32_t* const*)d_dst, d_dst, n_cols, m, Pl,
                     k + 1 == passes.size() ? scale : 0u));
  return SH_OK;
}

// Evaluation of n_cols polynomials with 2^n coefficients on domains of log m.
int sh_fft_evaluate_batch(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_src,
                          uint32_t* const* d_dst, uint32_t n_cols, uint32_t n, uint32_t m,
                          uint32_t Pl, uint32_t half_init) {
  if (n_cols == 0) return SH_OK;
  uint32_t src_mask = (1u << n) - 1;
  if (m <= 2) {
    SpanGuard span(ctx, "fft_tiny");
    sh::mpt p = sh::index_to_point_host(half_init);
    uint32_t c0 = m == 1 ? p.y : p.x, c1 = p.y;
    k_fft_tiny<<<(n_cols + 63) / 64, 64, 0, ctx->stream>>>(d_src, d_dst, src_mask, n_cols, m, 0, c0,
                                                           c1, 0, n);
    return sh_launch_check(ctx, "fft_tiny");
  }
  if (n == 0) {
    dim3 grid(((1u << m) + 255) / 256, n_cols);
    SpanGuard span(ctx, "fft_replicate");
    k_replicate<<<grid, 256, 0, ctx->stream>>>(d_src, d_dst, m);
    return sh_launch_check(ctx, "replicate");
  }
  const sh_ctx::FftTune tune = ctx->fft_tune[m];
  auto passes = plan_passes(n, tune);
  for (int k = (int)passes.size() - 1; k >= 0; k--) {
    const bool first = k == (int)passes.size() - 1;
    SH_TRY(rfft_pass(ctx, tw, passes[k], tune, first ? d_src : (const uint32_t* const*)d_dst, d_dst, n_cols, m, Pl,
                     first ? src_mask : 0xffffffffu));
  }
  return SH_OK;
}

// Interpolation of n_cols columns of log m (canonic domain; d_src -> d_coef, in place allowed) and
// their extension to log m + b (d_coef -> d_ext), with the interpolation's last strided pass and the
// extension's first fused (sh_fft2_strided_ie) when the context's fusion mode and the two pass plans
// allow; otherwise the two transforms one after the other.  Same words either way.
int sh_fft_interpolate_extend_batch(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_src,
                                    uint32_t* const* d_coef, uint32_t* const* d_ext, uint32_t n_cols, uint32_t m,
                                    uint32_t b) {
  if (n_cols == 0) return SH_OK;
  const sh_ctx::FftTune ti = ctx->fft_tune[m], te = ctx->fft_tune[m + b];
  std::vector<FftPass> pi, pe;
  if (m > 2 && b >= 1) {
    pi = plan_passes(m, ti);
    pe = ctx->fft_fuse == 2 ? pi : plan_passes(m, te);
  }
  const bool fuse = ctx->fft_fuse != 0 && ctx->fft2 && pi.size() >= 2 && pe.size() >= 2 && m + b <= 30 &&
                    pi.back().lo == pe.back().lo && pi.back().hi == m && pe.back().hi == m;
  if (!fuse) {
    SH_TRY(sh_fft_interpolate_batch(ctx, tw, d_src, d_coef, n_cols, m, m, sh::subgroup_gen(m + 1)));
    return sh_fft_evaluate_batch(ctx, tw, (const uint32_t* const*)d_coef, d_ext, n_cols, m, m + b, m + b,
                                 sh::subgroup_gen(m + b + 1));
  }
  const uint32_t scale = sh::minv(1u << m);
  for (size_t k = 0; k + 1 < pi.size(); k++)
    SH_TRY(ifft_pass(ctx, tw, pi[k], ti, k == 0 ? d_src : (const uint32_t* const*)d_coef, d_coef, n_cols, m, m, 0u));
  const FftPass& last = pi.back();
  int rc = SH_ERR_NOT_FOUND;
  {
    SpanGuard span(ctx, "fft_ie_pass");
    const uint32_t order[6] = {te.r_stride, ti.r_stride, 3, 2, 4, 1};
    for (uint32_t k = 0; k < 6 && rc == SH_ERR_NOT_FOUND; k++)
      rc = sh_fft2_strided_ie(ctx, last.hi - last.lo, order[k], last.lo, d_coef, d_ext, n_cols, m, b, tw->itw, tw->tw,
                              tw->root_log_size, scale);
  }
  if (rc == SH_ERR_NOT_FOUND) {  // no fused kernel of this shape: the two passes apart
    SH_TRY(ifft_pass(ctx, tw, last, ti, pi.size() == 1 ? d_src : (const uint32_t* const*)d_coef, d_coef, n_cols, m, m,
                     scale));
    SH_TRY(rfft_pass(ctx, tw, pe.back(), te, (const uint32_t* const*)d_coef, d_ext, n_cols, m + b, m + b,
                     (1u << m) - 1));
  } else {
    SH_TRY(rc);
  }
  for (int k = (int)pe.size() - 2; k >= 0; k--)
    SH_TRY(rfft_pass(ctx, tw, pe[k], te, (const uint32_t* const*)d_ext, d_ext, n_cols, m + b, m + b, 0xffffffffu));
  return SH_OK;
}

// Groups columns by size class and runs one batched launch chain per class.
static int interpolate_grouped(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* h_src,
                               uint32_t* const* h_dst, const uint32_t* log_sizes, uint64_t n) {
  std::vector<uint32_t> classes;
  for (uint64_t i = 0; i < n; i++) {
    SH_TRY(check_domain(ctx, tw, log_sizes[i]));
    bool seen = false;
    for (auto c : classes) seen |= c == log_sizes[i];
    if (!seen) classes.push_back(log_sizes[i]);
  }
  for (uint32_t m : classes) {
    std::vector<const void*> ptrs;
    for (uint64_t i = 0; i < n; i++)
      if (log_sizes[i] == m) ptrs.push_back(h_src[i]);
    for (uint64_t i = 0; i < n; i++)
      if (log_sizes[i] == m) ptrs.push_back(h_dst[i]);
    void* d_tab;
    SH_TRY(sh_stage(ctx, ptrs.data(), ptrs.size() * sizeof(void*), &d_tab));
    uint32_t cnt = (uint32_t)(ptrs.size() / 2);
    const uint32_t* const* d_src = (const uint32_t* const*)d_tab;
    uint32_t* const* d_dst = (uint32_t* const*)((const void**)d_tab + cnt);
    SH_TRY(sh_fft_interpolate_batch(ctx, tw, d_src, d_dst, cnt, m, m, sh::subgroup_gen(m + 1)));
  }
  return SH_OK;
}

int sh_interpolate_grouped(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* h_src,
                           uint32_t* const* h_dst, const uint32_t* log_sizes, uint64_t n) {
  return interpolate_grouped(ctx, tw, h_src, h_dst, log_sizes, n);
}

int sh_evaluate_grouped(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* h_src,
                        const uint32_t* coeff_logs, const uint32_t* domain_logs,
                        uint32_t* const* h_dst, uint64_t n) {
  std::vector<std::pair<uint32_t, uint32_t>> classes;
  for (uint64_t i = 0; i < n; i++) {
    SH_TRY(check_domain(ctx, tw, domain_logs[i]));
    SH_REQUIRE(ctx, coeff_logs[i] <= domain_logs[i], "Can only evaluate on larger domains (%u > %u)",
               coeff_logs[i], domain_logs[i]);
    auto key = std::make_pair(coeff_logs[i], domain_logs[i]);
    bool seen = false;
    for (auto& c : classes) seen |= c == key;
    if (!seen) classes.push_back(key);
  }
  for (auto& c : classes) {
    std::vector<const void*> ptrs;
    for (uint64_t i = 0; i < n; i++)
      if (coeff_logs[i] == c.first && domain_logs[i] == c.second) ptrs.push_back(h_src[i]);
    for (uint64_t i = 0; i < n; i++)
      if (coeff_logs[i] == c.first && domain_logs[i] == c.second) ptrs.push_back(h_dst[i]);
    void* d_tab;
    SH_TRY(sh_stage(ctx, ptrs.data(), ptrs.size() * sizeof(void*), &d_tab));
    uint32_t cnt = (uint32_t)(ptrs.size() / 2);
    const uint32_t* const* d_src = (const uint32_t* const*)d_tab;
    uint32_t* const* d_dst = (uint32_t* const*)((const void**)d_tab + cnt);
    SH_TRY(sh_fft_evaluate_batch(ctx, tw, d_src, d_dst, cnt, c.first, c.second, c.second,
                                 sh::subgroup_gen(c.second + 1)));
  }
  return SH_OK;
}

// One size class of a batched commit: interpolation of n columns of log m (h_src -> h_coef) and their
// extension to log m + b (h_coef -> h_ext) through sh_fft_interpolate_extend_batch.
int sh_interpolate_extend_grouped(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* h_src,
                                  uint32_t* const* h_coef, uint32_t* const* h_ext, uint32_t m, uint32_t b,
                                  uint64_t n) {
  if (n == 0) return SH_OK;
  SH_TRY(check_domain(ctx, tw, m));
  SH_TRY(check_domain(ctx, tw, m + b));
  std::vector<const void*> ptrs;
  for (uint64_t i = 0; i < n; i++) ptrs.push_back(h_src[i]);
  for (uint64_t i = 0; i < n; i++) ptrs.push_back(h_coef[i]);
  for (uint64_t i = 0; i < n; i++) ptrs.push_back(h_ext[i]);
  void* d_tab;
  SH_TRY(sh_stage(ctx, ptrs.data(), ptrs.size() * sizeof(void*), &d_tab));
  const void** t = (const void**)d_tab;
  return sh_fft_interpolate_extend_batch(ctx, tw, (const uint32_t* const*)t, (uint32_t* const*)(t + n),
                                         (uint32_t* const*)(t + 2 * n), (uint32_t)n, m, b);
}

extern "C" int sh_interpolate(sh_ctx* ctx, const sh_twiddles* tw, uint32_t* d_values, uint32_t log_size) {
  if (!ctx || !tw || !d_values) return SH_ERR_INVALID;
  const uint32_t* s = d_values;
  return interpolate_grouped(ctx, tw, &s, &d_values, &log_size, 1);
}

extern "C" int sh_interpolate_columns(sh_ctx* ctx, const sh_twiddles* tw, uint32_t* const* d_cols,
                                      const uint32_t* log_sizes, uint64_t n) {
  if (!ctx || !tw || (n && (!d_cols || !log_sizes))) return SH_ERR_INVALID;
  return interpolate_grouped(ctx, tw, (const uint32_t* const*)d_cols, d_cols, log_sizes, n);
}

extern "C" int sh_evaluate(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* d_coeffs, uint32_t coeff_log,
                           uint32_t* d_out, uint32_t domain_log) {
  if (!ctx || !tw || !d_coeffs || !d_out) return SH_ERR_INVALID;
  return sh_evaluate_grouped(ctx, tw, &d_coeffs, &coeff_log, &domain_log, &d_out, 1);
}

extern "C" int sh_evaluate_polynomials(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_coeffs,
                                       const uint32_t* coeff_logs, uint64_t n, uint32_t log_blowup,
                                       uint32_t* const* d_outs) {
  if (!ctx || !tw || (n && (!d_coeffs || !coeff_logs || !d_outs))) return SH_ERR_INVALID;
  std::vector<uint32_t> dl(n);
  for (uint64_t i = 0; i < n; i++) dl[i] = coeff_logs[i] + log_blowup;
  return sh_evaluate_grouped(ctx, tw, d_coeffs, coeff_logs, dl.data(), d_outs, n);
}

// Autotuner: for every transform size the workload uses (interpolations of
// log n, evaluations onto log n + log_blowup), times each candidate pass shape
// on a batch of the class's real width and keeps the fastest.  Inputs are
// scratch columns; the shape only changes speed, never results.
extern "C" int sh_autotune_fft(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* log_sizes, uint64_t n,
                               uint32_t log_blowup) {
  if (!ctx || !tw || (n && !log_sizes)) return SH_ERR_INVALID;
  std::vector<std::pair<uint32_t, uint32_t>> cls;  // (log, count)
  for (uint64_t i = 0; i < n; i++) {
    SH_TRY(check_domain(ctx, tw, log_sizes[i] + log_blowup));
    bool seen = false;
    for (auto& c : cls)
      if (c.first == log_sizes[i]) {
        c.second++;
        seen = true;
      }
    if (!seen) cls.push_back({log_sizes[i], 1});
  }
  std::vector<sh_ctx::FftTune> cands;
  for (uint8_t blk : {11, 12, 13})
    for (uint8_t st : {5, 6, 7, 8})
      for (uint8_t rb : {1, 2, 3})
        for (uint8_t rs : {1, 2, 3}) cands.push_back({blk, st, rb, rs});
  hipEvent_t e0, e1;
  SH_CHECK_HIP(ctx, hipEventCreate(&e0));
  SH_CHECK_HIP(ctx, hipEventCreate(&e1));
  int rc = SH_OK;
  for (auto& c : cls) {
    const uint32_t lg = c.first, cnt = c.second;
    if (lg + log_blowup <= 2) continue;
    uint32_t* src = nullptr;
    uint32_t* dst = nullptr;
    if (hipMalloc(&src, (4ull << lg) * cnt) != hipSuccess ||
        hipMalloc(&dst, (4ull << (lg + log_blowup)) * cnt) != hipSuccess) {
      if (src) hipFree(src);
      rc = sh_fail(ctx, SH_ERR_OOM, "autotune scratch allocation failed");
      break;
    }
    std::vector<const void*> tab;
    for (uint32_t k = 0; k < cnt; k++) tab.push_back(src + ((uint64_t)k << lg));
    for (uint32_t k = 0; k < cnt; k++) tab.push_back(dst + ((uint64_t)k << (lg + log_blowup)));
    void* d_tab = nullptr;
    rc = sh_stage(ctx, tab.data(), tab.size() * sizeof(void*), &d_tab);
    const uint32_t* const* ts = (const uint32_t* const*)d_tab;
    uint32_t* const* td = (uint32_t* const*)((const void**)d_tab + cnt);
    for (int which = 0; which < 2 && rc == SH_OK; which++) {
      // which 0: interpolation of log lg (in place on src); 1: evaluation lg -> lg + b
      const uint32_t m = which == 0 ? lg : lg + log_blowup;
      if (m <= 2 || (which == 1 && log_blowup == 0)) continue;
      float best = 1e30f;
      sh_ctx::FftTune best_t = ctx->fft_tune[m];
      for (const auto& t : cands) {
        if (t.block_log >= m && (t.stride_layers != 8 || t.r_stride != 3)) continue;  // same plan
        ctx->fft_tune[m] = t;
        float total = 1e30f;  // the fastest of 4 timed runs (the first, untimed, warms caches)
        for (int rep = 0; rep < 5 && rc == SH_OK; rep++) {
          SH_CHECK_HIP(ctx, hipEventRecord(e0, ctx->stream));
          if (which == 0)
            rc = sh_fft_interpolate_batch(ctx, tw, ts, (uint32_t* const*)ts, cnt, m, m, sh::subgroup_gen(m + 1));
          else
            rc = sh_fft_evaluate_batch(ctx, tw, ts, td, cnt, lg, m, m, sh::subgroup_gen(m + 1));
          SH_CHECK_HIP(ctx, hipEventRecord(e1, ctx->stream));
          SH_CHECK_HIP(ctx, hipEventSynchronize(e1));
          float ms = 0.f;
          SH_CHECK_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
          if (rep > 0) total = std::min(total, ms);
        }
        if (total < best) {
          best = total;
          best_t = t;
        }
      }
      ctx->fft_tune[m] = best_t;
    }
    SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(src);
    hipFree(dst);
    if (rc != SH_OK) break;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return rc;
}

// PolyOps::split_at_mid (crates/stwo/src/prover/poly/circle/ops.rs:145-165;
// CpuBackend crates/stwo/src/prover/backend/cpu/circle.rs:266-271): with the
// coefficients in natural FFT-basis order the split is the two halves of the
// vector (p = p_left + pi^(L-2)(x) * p_right).  d_left may equal d_coeffs (the
// left half then stays in place); neither output may overlap the right half.
extern "C" int sh_split_at_mid(sh_ctx* ctx, const uint32_t* d_coeffs, uint32_t log_size, uint32_t* d_left,
                               uint32_t* d_right) {
  if (!ctx || !d_coeffs || !d_left || !d_right) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, log_size >= 1 && log_size < 32, "split_at_mid needs at least 2 coefficients (log %u)", log_size);
  const uint64_t half = 1ull << (log_size - 1);
  const uint32_t* right = d_coeffs + half;
  SH_REQUIRE(ctx, !(d_right < right + half && right < d_right + half) && !(d_left < right + half && right < d_left + half),
             "split_at_mid outputs overlap the right half of the input");
  SpanGuard span(ctx, "split_at_mid");
  if (d_left != d_coeffs)
    SH_CHECK_HIP(ctx, hipMemcpyAsync(d_left, d_coeffs, 4 * half, hipMemcpyDeviceToDevice, ctx->stream));
  SH_CHECK_HIP(ctx, hipMemcpyAsync(d_right, right, 4 * half, hipMemcpyDeviceToDevice, ctx->stream));
  return SH_OK;
}
