// Out-of-domain sampling: PolyOps::eval_at_point, barycentric_weights_into,
// barycentric_eval_at_point (crates/stwo/src/prover/poly/circle/ops.rs:44-80).
//
// eval_at_point (CpuBackend, crates/stwo/src/prover/backend/cpu/circle.rs:81-98):
// value = sum_i c_i * prod_{bit k of i} f_k with f_0 = p.y, f_1 = p.x,
// f_k = pi^(k-1)(p.x) (the fold of crates/stwo/src/core/poly/utils.rs:27-38).
// Batched on the device: the basis value of index i = hi * 2^s + lo factors as
// B_hi[hi] * B_lo[lo] (products of the factors of the high and the low bits), so
// with B_lo (2^s words) and B_hi (2^(n-s)) tabulated per point on the device (each entry
// the product of its bits' factors, k_eval_tables; the host computes only the factors), lane
// lo of a block sums c[hi * 2^s + lo] * B_hi[hi] over a range of hi (4 lazy
// 32x32->64 multiply-adds per coefficient and point; B_hi is uniform across the
// block), multiplies by B_lo[lo] once, and the block's sum is added into 64-bit
// per-(column, point) accumulators.  One launch per column size covers every
// column of that size and up to 4 points, reading each coefficient once.
//
// Barycentric weights (cpu/circle.rs:100-131), for the bit-reversed domain point q_i:
//   W_i = V_n(p) / (-2 q_i.y * V'_n(q_i) * (h.y / (1 + h.x))),  h = p - q_i,
// with V'_n the coset-vanishing derivative (crates/stwo/src/core/constraints.rs:37-55).
// q_i has M31 coordinates, so V'_n(q_i) is evaluated in M31.
#include <algorithm>

#include "common.h"

namespace {

constexpr uint32_t kEvalLoLog = 8;   // 256 lanes: the low 8 index bits
constexpr uint32_t kEvalHiPerBlock = 256;
constexpr uint32_t kEvalMaxPoints = 4;

// The factors f_k of up to kEvalMaxPoints points (QM31, k < 31), passed by value.
struct EvalFactors {
  uint32_t w[kEvalMaxPoints][31][4];
};

// b_lo[(i * np + pt)] = prod_{bit j of i} f[pt][j] for i < 2^sl, and
// b_hi[(i * np + pt)] = prod_{bit j of i} f[pt][sl + j] for i < 2^nh (QM31 words).
__device__ __forceinline__ void eval_table_entry(const EvalFactors& f, uint32_t np, uint32_t sl, uint32_t nh,
                                                 uint32_t* __restrict__ b_lo, uint32_t* __restrict__ b_hi);
__global__ __launch_bounds__(256) void k_eval_tables(EvalFactors f, uint32_t np, uint32_t sl, uint32_t nh,
                                                     uint32_t* __restrict__ b_lo, uint32_t* __restrict__ b_hi) {
  eval_table_entry(f, np, sl, nh, b_lo, b_hi);
}
// The same with the factors in device memory (a point the device drew).
__global__ __launch_bounds__(256) void k_eval_tables_dev(const EvalFactors* __restrict__ f, uint32_t np, uint32_t sl,
                                                         uint32_t nh, uint32_t* __restrict__ b_lo,
                                                         uint32_t* __restrict__ b_hi) {
  eval_table_entry(*f, np, sl, nh, b_lo, b_hi);
}
__device__ __forceinline__ void eval_table_entry(const EvalFactors& f, uint32_t np, uint32_t sl, uint32_t nh,
                                                 uint32_t* __restrict__ b_lo, uint32_t* __restrict__ b_hi) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n_lo = np << sl, n_hi = np << nh;
  if (t >= n_lo + n_hi) return;
  const bool lo = t < n_lo;
  const uint32_t e = lo ? t : t - n_lo;
  const uint32_t i = e / np, pt = e - i * np;
  const uint32_t bits = lo ? sl : nh, first = lo ? 0 : sl;
  sh::qm31 acc = sh::qone();
  for (uint32_t j = 0; j < bits; j++)
    if ((i >> j) & 1) acc = sh::qmul(acc, sh::qload(f.w[pt][first + j]));
  sh::qstore(acc, (lo ? b_lo : b_hi) + 4ull * e);
}

__device__ __forceinline__ void lfold(uint64_t& v) { v = (v & sh::P) + (v >> 31); }
__device__ __forceinline__ uint32_t lfinal(uint64_t v) {
  lfold(v);
  lfold(v);
  return sh::reduce_once((uint32_t)v);
}

// blockIdx.y = column (table of pointers), blockIdx.x = range of hi; NP points in one pass
// over the coefficients.  b_lo: [lo][pt] QM31, b_hi: [hi][pt] QM31; acc: [col][pt][4] u64
// (zeroed by the caller).
template <int NP>
__global__ __launch_bounds__(256) void k_eval_points(const uint32_t* const* __restrict__ cols, uint32_t log_n,
                                                     const uint32_t* __restrict__ b_lo,
                                                     const uint32_t* __restrict__ b_hi,
                                                     unsigned long long* __restrict__ acc) {
  __shared__ uint32_t red[4 * kEvalMaxPoints][4];  // [point * 4 + wave]
  const uint32_t s = log_n < kEvalLoLog ? log_n : kEvalLoLog;
  const uint32_t lo = threadIdx.x;
  const uint32_t n_hi = 1u << (log_n - s);
  const uint32_t h0 = blockIdx.x * kEvalHiPerBlock;
  const uint32_t h1 = h0 + kEvalHiPerBlock < n_hi ? h0 + kEvalHiPerBlock : n_hi;
  // global address space (no FLAT counter waits); a column is below 2^30 words
  const __attribute__((address_space(1))) uint32_t* __restrict__ c =
      (const __attribute__((address_space(1))) uint32_t*)cols[blockIdx.y];
  const bool live = lo < (1u << s);
  uint64_t a[NP][4];
#pragma unroll
  for (int pt = 0; pt < NP; pt++) a[pt][0] = a[pt][1] = a[pt][2] = a[pt][3] = 0;
  auto mac = [&](uint32_t v, uint32_t hi) {
#pragma unroll
    for (int pt = 0; pt < NP; pt++) {
      const uint4 bh = *(const uint4*)&b_hi[4 * (hi * NP + pt)];
      a[pt][0] += (uint64_t)v * bh.x;
      a[pt][1] += (uint64_t)v * bh.y;
      a[pt][2] += (uint64_t)v * bh.z;
      a[pt][3] += (uint64_t)v * bh.w;
    }
  };
  auto fold_all = [&]() {
#pragma unroll
    for (int pt = 0; pt < NP; pt++)
#pragma unroll
      for (int q = 0; q < 4; q++) lfold(a[pt][q]);
  };
  if (live) {
    uint32_t hi = h0;
    // six coefficients per step, their loads issued together; a fold after every three
    // products (three products < 2^62 on a folded value < 2^34 stay below 2^64)
    for (; hi + 6 <= h1; hi += 6) {
      uint32_t v[6];
#pragma unroll
      for (int u = 0; u < 6; u++) v[u] = c[((hi + u) << s) + lo];
#pragma unroll
      for (int u = 0; u < 6; u++) {
        mac(v[u], hi + u);
        if (u == 2 || u == 5) fold_all();
      }
    }
    for (; hi < h1; hi++) {
      mac(c[(hi << s) + lo], hi);
      fold_all();
    }
  }
  // block sum per point: within each wave by lane shuffles, then the four wave sums through LDS
  // (one barrier), then one 64-bit atomic add per coordinate
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int pt = 0; pt < NP; pt++) {
    sh::qm31 part = {{lfinal(a[pt][0]), lfinal(a[pt][1])}, {lfinal(a[pt][2]), lfinal(a[pt][3])}};
    part = live ? sh::qmul(part, sh::qload(b_lo + 4 * (lo * NP + pt))) : sh::qzero();
    uint32_t w[4] = {part.a.a, part.a.b, part.b.a, part.b.b};
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
      for (int q = 0; q < 4; q++) w[q] = sh::madd(w[q], __shfl_xor(w[q], off, 64));
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < 4; q++) red[pt * 4 + wave][q] = w[q];
  }
  __syncthreads();
  if (threadIdx.x < 4 * NP) {
    const uint32_t pt = threadIdx.x >> 2, q = threadIdx.x & 3;
    uint32_t sum = 0;
    for (uint32_t wv = 0; wv < 4; wv++) sum = sh::madd(sum, red[pt * 4 + wv][q]);
    atomicAdd(&acc[4 * ((uint64_t)blockIdx.y * NP + pt) + q], (unsigned long long)sum);
  }
}

struct BaryConsts {
  uint32_t log_n;
  sh::qpt p;
  sh::qm31 vn_p;
  sh::mpt shift[31];  // coset_vanishing shift of CanonicCoset(j).coset, j = 1..log_n-1
  uint32_t four_pow;  // 4^(log_n - 1)
};

__global__ __launch_bounds__(256) void k_bary_weights(BaryConsts bc, sh_point_tab half, uint32_t* const* w) {
  const uint32_t n = 1u << bc.log_n;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sh::mpt q = sh_circle_domain_point(half, bc.log_n - 1, sh::bitrev(i, bc.log_n));
  // V'_n(q) = 4^(n-1) * prod_{j=1}^{n-1} coset_vanishing(CanonicCoset(j).coset, q)
  uint32_t d = bc.four_pow;
  for (uint32_t j = 1; j < bc.log_n; j++) {
    uint32_t x = sh::padd(q, bc.shift[j]).x;
    for (uint32_t r = 1; r < j; r++) x = sh::msub(sh::madd(sh::mmul(x, x), sh::mmul(x, x)), 1);
    d = sh::mmul(d, x);
  }
  uint32_t si = sh::mmul(sh::mneg(sh::madd(q.y, q.y)), d);
  // h = p - q = p + conj(q)
  sh::qm31 hx = sh::qadd(sh::qmulm(bc.p.x, q.x), sh::qmulm(bc.p.y, q.y));
  sh::qm31 hy = sh::qsub(sh::qmulm(bc.p.y, q.x), sh::qmulm(bc.p.x, q.y));
  sh::qm31 den = sh::qmulm(hy, si);
  sh::qm31 wv = sh::qmul(sh::qmul(bc.vn_p, sh::qadd(sh::qone(), hx)), sh::qinv(den));
  w[0][i] = wv.a.a;
  w[1][i] = wv.a.b;
  w[2][i] = wv.b.a;
  w[3][i] = wv.b.b;
}

// partial[block] = sum_i w_i * e_i over a grid-stride range (lazy 64-bit accumulation).
__global__ __launch_bounds__(256) void k_bary_dot(const uint32_t* __restrict__ e, const uint32_t* const* w,
                                                  uint32_t n, uint32_t* partial) {
  __shared__ uint32_t red[256][4];
  uint64_t acc[4] = {0, 0, 0, 0};
  uint32_t k = 0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    uint32_t v = e[i];
#pragma unroll
    for (int c = 0; c < 4; c++) acc[c] += (uint64_t)w[c][i] * v;
    if (++k == 3) {
      k = 0;
#pragma unroll
      for (int c = 0; c < 4; c++) lfold(acc[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) red[threadIdx.x][c] = lfinal(acc[c]);
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int c = 0; c < 4; c++) red[threadIdx.x][c] = sh::madd(red[threadIdx.x][c], red[threadIdx.x + s][c]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 4; c++) partial[4 * blockIdx.x + c] = red[0][c];
}

__global__ void k_sum_partials(const uint32_t* partial, uint32_t nblk, uint32_t* out) {
  uint32_t s[4] = {0, 0, 0, 0};
  for (uint32_t b = 0; b < nblk; b++)
    for (int c = 0; c < 4; c++) s[c] = sh::madd(s[c], partial[4 * b + c]);
  for (int c = 0; c < 4; c++) out[c] = s[c];
}

// coset_vanishing(coset, p) for QM31 p on the host (constraints.rs:12-35).
sh::qm31 coset_vanishing_host(uint32_t init_index, uint32_t step_index, uint32_t log_size, sh::qpt p) {
  sh::mpt ci = sh::index_to_point_host(init_index);
  sh::mpt hs = sh::index_to_point_host(step_index >> 1);
  sh::qpt neg_init = {sh::qfromm(ci.x), sh::qfromm(sh::mneg(ci.y))};
  sh::qpt q = sh::qpadd(sh::qpadd(p, neg_init), {sh::qfromm(hs.x), sh::qfromm(hs.y)});
  sh::qm31 x = q.x;
  for (uint32_t i = 1; i < log_size; i++) x = sh::qdouble_x(x);
  return x;
}

}  // namespace

// Many columns (any sizes) at many points: out[(col * n_points + pt) * 4 ..] (host).
extern "C" int sh_eval_at_points(sh_ctx* ctx, const uint32_t* const* d_coeffs, const uint32_t* log_sizes,
                                 uint64_t n_cols, const uint32_t* points, uint32_t n_points, uint32_t* out) {
  if (!ctx || (n_cols && (!d_coeffs || !log_sizes || !out)) || (n_points && !points)) return SH_ERR_INVALID;
  if (n_cols == 0 || n_points == 0) return SH_OK;
  for (uint64_t i = 0; i < n_cols; i++) SH_REQUIRE(ctx, log_sizes[i] <= 30, "log size %u too large", log_sizes[i]);
  std::vector<uint32_t> classes(log_sizes, log_sizes + n_cols);
  std::sort(classes.begin(), classes.end());
  classes.erase(std::unique(classes.begin(), classes.end()), classes.end());
  SpanGuard span(ctx, "eval_at_point");
  for (uint32_t p0 = 0; p0 < n_points; p0 += kEvalMaxPoints) {
    const uint32_t np = std::min(kEvalMaxPoints, n_points - p0);
    for (uint32_t n : classes) {
      // factor of index bit k: bit 0 -> y, bit 1 -> x, bit k -> pi^(k-1)(x)
      const uint32_t sl = std::min(n, kEvalLoLog), nh = n - sl;
      EvalFactors fac;
      for (uint32_t pt = 0; pt < np; pt++) {
        const uint32_t* P8 = points + 8ull * (p0 + pt);
        sh::qm31 x = sh::qload(P8);
        for (uint32_t k = 0; k < n; k++) {
          if (k == 0) {
            sh::qstore(sh::qload(P8 + 4), fac.w[pt][0]);
          } else {
            sh::qstore(x, fac.w[pt][k]);
            x = sh::qdouble_x(x);
          }
        }
      }
      std::vector<const void*> ptrs;
      std::vector<uint64_t> idx;
      for (uint64_t i = 0; i < n_cols; i++)
        if (log_sizes[i] == n) {
          ptrs.push_back(d_coeffs[i]);
          idx.push_back(i);
        }
      // columns of this class may not be contiguous in the output: accumulate into a class slice
      void *d_ptrs, *d_tab, *d_cls;
      SH_TRY(sh_stage(ctx, ptrs.data(), ptrs.size() * sizeof(void*), &d_ptrs));
      const uint64_t lo_words = (4ull * np) << sl, hi_words = (4ull * np) << nh;
      SH_TRY(sh_scratch_get(ctx, SCR_EVAL, 4 * (lo_words + hi_words), &d_tab));
      uint32_t* d_lo = (uint32_t*)d_tab;
      uint32_t* d_hi = d_lo + lo_words;
      const uint32_t n_tab = (np << sl) + (np << nh);
      k_eval_tables<<<(n_tab + 255) / 256, 256, 0, ctx->stream>>>(fac, np, sl, nh, d_lo, d_hi);
      SH_TRY(sh_launch_check(ctx, "eval_tables"));
      SH_TRY(sh_scratch_get(ctx, SCR_QUOT, 32ull * ptrs.size() * np, &d_cls));
      SH_TRY(sh_memset(ctx, d_cls, 0, 32ull * ptrs.size() * np));
      dim3 grid((unsigned)(((1ull << nh) + kEvalHiPerBlock - 1) / kEvalHiPerBlock), (unsigned)ptrs.size());
      auto kern = np == 1 ? k_eval_points<1> : np == 2 ? k_eval_points<2> : np == 3 ? k_eval_points<3>
                                                                              : k_eval_points<4>;
      kern<<<grid, 256, 0, ctx->stream>>>((const uint32_t* const*)d_ptrs, n, (const uint32_t*)d_lo,
                                          (const uint32_t*)d_hi, (unsigned long long*)d_cls);
      SH_TRY(sh_launch_check(ctx, "eval_at_points"));
      std::vector<unsigned long long> cls(4ull * ptrs.size() * np);
      SH_CHECK_HIP(ctx, hipMemcpyAsync(cls.data(), d_cls, 8 * cls.size(), hipMemcpyDeviceToHost, ctx->stream));
      SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
      for (size_t c = 0; c < idx.size(); c++)
        for (uint32_t pt = 0; pt < np; pt++)
          for (int q = 0; q < 4; q++)
            out[4 * (idx[c] * n_points + p0 + pt) + q] = (uint32_t)(cls[4 * (c * np + pt) + q] % sh::P);
    }
  }
  return SH_OK;
}

namespace {
// The factors of one point (slot 0) for polynomials of 2^n coefficients: the point doubled dbl
// times (CirclePoint::repeated_double), then f_0 = y, f_1 = x, f_k = pi^(k-1)(x).
__global__ void k_eval_factors(const uint32_t* __restrict__ point, uint32_t dbl, uint32_t n,
                               EvalFactors* __restrict__ f) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  sh::qm31 x = sh::qload(point), y = sh::qload(point + 4);
  for (uint32_t k = 0; k < dbl; k++) {  // (x, y) -> (2x^2 - 1, 2xy)
    const sh::qm31 xy = sh::qmul(x, y);
    x = sh::qdouble_x(x);
    y = sh::qadd(xy, xy);
  }
  for (uint32_t k = 0; k < n; k++) {
    if (k == 0) {
      sh::qstore(y, f->w[0][0]);
    } else {
      sh::qstore(x, f->w[0][k]);
      x = sh::qdouble_x(x);
    }
  }
}

// values[idx[c]] = the reduced accumulator of class column c (one point).
__global__ void k_eval_finalize(const unsigned long long* __restrict__ acc, uint32_t n, const uint32_t* __restrict__ idx,
                                uint32_t* __restrict__ values) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * n) return;
  values[4ull * idx[t >> 2] + (t & 3)] = lfinal(acc[t]);
}
}  // namespace

// eval_at_point of every column at one QM31 point held in device memory (8 words the stream has
// produced, e.g. drawn by a device channel), column i at the point doubled doublings[i] times
// (compute_samples, crates/stwo/src/prover/pcs/mod.rs:176-222); d_values (device) receives 4
// words per column.  Nothing returns to the host.
extern "C" int sh_eval_at_dpoint(sh_ctx* ctx, const uint32_t* const* d_coeffs, const uint32_t* log_sizes,
                                 const uint32_t* doublings, uint64_t n_cols, const uint32_t* d_point,
                                 uint32_t* d_values) {
  if (!ctx || (n_cols && (!d_coeffs || !log_sizes || !doublings || !d_point || !d_values))) return SH_ERR_INVALID;
  if (n_cols == 0) return SH_OK;
  std::vector<std::pair<uint32_t, uint32_t>> classes;  // (log size, doublings)
  for (uint64_t i = 0; i < n_cols; i++) {
    SH_REQUIRE(ctx, log_sizes[i] >= 1 && log_sizes[i] <= 30, "log size %u out of range", log_sizes[i]);
    SH_REQUIRE(ctx, doublings[i] <= 31, "doubling count %u out of range", doublings[i]);
    classes.push_back({log_sizes[i], doublings[i]});
  }
  std::sort(classes.begin(), classes.end());
  classes.erase(std::unique(classes.begin(), classes.end()), classes.end());
  SpanGuard span(ctx, "eval_at_point");
  for (auto& cl : classes) {
    const uint32_t n = cl.first;
    std::vector<const void*> ptrs;
    std::vector<uint32_t> idx;
    for (uint64_t i = 0; i < n_cols; i++)
      if (log_sizes[i] == n && doublings[i] == cl.second) {
        ptrs.push_back(d_coeffs[i]);
        idx.push_back((uint32_t)i);
      }
    const uint32_t sl = std::min(n, kEvalLoLog), nh = n - sl;
    void *d_ptrs, *d_idx, *d_tab, *d_cls;
    SH_TRY(sh_stage(ctx, ptrs.data(), ptrs.size() * sizeof(void*), &d_ptrs));
    SH_TRY(sh_stage(ctx, idx.data(), idx.size() * 4, &d_idx));
    const uint64_t lo_words = 4ull << sl, hi_words = 4ull << nh;
    const uint64_t fac_bytes = (sizeof(EvalFactors) + 255) & ~255ull;
    SH_TRY(sh_scratch_get(ctx, SCR_EVAL, fac_bytes + 4 * (lo_words + hi_words), &d_tab));
    EvalFactors* d_fac = (EvalFactors*)d_tab;
    uint32_t* d_lo = (uint32_t*)((uint8_t*)d_tab + fac_bytes);
    uint32_t* d_hi = d_lo + lo_words;
    k_eval_factors<<<1, 64, 0, ctx->stream>>>(d_point, cl.second, n, d_fac);
    SH_TRY(sh_launch_check(ctx, "eval_factors"));
    const uint32_t n_tab = (1u << sl) + (1u << nh);
    k_eval_tables_dev<<<(n_tab + 255) / 256, 256, 0, ctx->stream>>>(d_fac, 1, sl, nh, d_lo, d_hi);
    SH_TRY(sh_launch_check(ctx, "eval_tables"));
    SH_TRY(sh_scratch_get(ctx, SCR_QUOT, 32ull * ptrs.size(), &d_cls));
    SH_TRY(sh_memset(ctx, d_cls, 0, 32ull * ptrs.size()));
    dim3 grid((unsigned)(((1ull << nh) + kEvalHiPerBlock - 1) / kEvalHiPerBlock), (unsigned)ptrs.size());
    k_eval_points<1><<<grid, 256, 0, ctx->stream>>>((const uint32_t* const*)d_ptrs, n, (const uint32_t*)d_lo,
                                                    (const uint32_t*)d_hi, (unsigned long long*)d_cls);
    SH_TRY(sh_launch_check(ctx, "eval_at_points"));
    const uint32_t nw = 4 * (uint32_t)ptrs.size();
    k_eval_finalize<<<(nw + 255) / 256, 256, 0, ctx->stream>>>((const unsigned long long*)d_cls,
                                                                (uint32_t)ptrs.size(), (const uint32_t*)d_idx,
                                                                d_values);
    SH_TRY(sh_launch_check(ctx, "eval_finalize"));
  }
  return SH_OK;
}

extern "C" int sh_eval_at_point(sh_ctx* ctx, const uint32_t* d_coeffs, uint32_t log_size, const uint32_t point[8],
                                uint32_t out[4]) {
  if (!ctx || !d_coeffs || !point || !out) return SH_ERR_INVALID;
  const uint32_t* c = d_coeffs;
  return sh_eval_at_points(ctx, &c, &log_size, 1, point, 1, out);
}

extern "C" int sh_barycentric_weights(sh_ctx* ctx, uint32_t log_size, const uint32_t point[8], uint32_t* const d_w[4]) {
  if (!ctx || !point || !d_w) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, log_size >= 1 && log_size <= 30, "log size %u out of range", log_size);
  BaryConsts bc = {};
  bc.log_n = log_size;
  bc.p = {sh::qload(point), sh::qload(point + 4)};
  // V_n(p) on CanonicCoset(n).coset = Coset::odds(n)
  bc.vn_p = coset_vanishing_host(sh::subgroup_gen(log_size + 1), sh::subgroup_gen(log_size), log_size, bc.p);
  for (uint32_t j = 1; j < log_size; j++) {
    // q - initial + step/2 for CanonicCoset(j).coset: a single point shift
    sh::mpt ci = sh::index_to_point_host(sh::subgroup_gen(j + 1));
    sh::mpt hs = sh::index_to_point_host(sh::subgroup_gen(j) >> 1);
    bc.shift[j] = sh::padd({ci.x, sh::mneg(ci.y)}, hs);
  }
  bc.four_pow = sh::mpow(4, log_size - 1);
  sh_point_tab half;
  SH_TRY(sh_point_tab_make(ctx, sh::subgroup_gen(log_size + 1), log_size - 1, &half));
  void* d_wt;
  SH_TRY(sh_stage(ctx, d_w, 4 * sizeof(void*), &d_wt));
  SpanGuard span(ctx, "barycentric_weights");
  k_bary_weights<<<((1u << log_size) + 255) / 256, 256, 0, ctx->stream>>>(bc, half, (uint32_t* const*)d_wt);
  return sh_launch_check(ctx, "barycentric_weights");
}

extern "C" int sh_barycentric_eval_at_point(sh_ctx* ctx, const uint32_t* d_evals, uint32_t log_size,
                                            const uint32_t* const d_w[4], uint32_t out[4]) {
  if (!ctx || !d_evals || !d_w || !out) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, log_size <= 30, "log size %u too large", log_size);
  uint32_t n = 1u << log_size;
  uint32_t nblk = std::min<uint32_t>((n + 255) / 256, 1024);
  void* scr;
  SH_TRY(sh_scratch_get(ctx, SCR_MISC, 16ull * nblk + 16, &scr));
  uint32_t* partial = (uint32_t*)scr;
  uint32_t* res = partial + 4 * nblk;
  void* d_wt;
  SH_TRY(sh_stage(ctx, d_w, 4 * sizeof(void*), &d_wt));
  SpanGuard span(ctx, "barycentric_eval");
  k_bary_dot<<<nblk, 256, 0, ctx->stream>>>(d_evals, (const uint32_t* const*)d_wt, n, partial);
  k_sum_partials<<<1, 1, 0, ctx->stream>>>(partial, nblk, res);
  SH_TRY(sh_launch_check(ctx, "barycentric_eval"));
  SH_CHECK_HIP(ctx, hipMemcpyAsync(out, res, 16, hipMemcpyDeviceToHost, ctx->stream));
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SH_OK;
}
