// DEEP quotients and accumulation:
//   QuotientOps::accumulate_numerators, QuotientOps::compute_quotients_and_combine,
//   AccumulationOps::lift_and_accumulate / accumulate.
//
// accumulate_numerators (CpuBackend: crates/stwo/src/prover/backend/cpu/quotients.rs:23-49):
// for each sample batch (one point) and each row of the evaluation subdomain
// (the first 2^(n - log_blowup) rows in bit-reversed order),
//     partial[row] = sum_t (c_t * col_{i_t}[row] - b_t),   first_linear = sum_t a_t,
// with (a, b, c) the complex-conjugate line coefficients times the random power
// (crates/stwo/src/core/constraints.rs:119-134).  The sum is linear in the
// columns, so on the device it becomes  sum_col C[col][batch] * col[row] - B[batch]
// with C the per-(column, batch) sum of c and B the per-batch sum of b.  The SIMD
// backend re-reads every column once per batch (simd/quotients.rs:68-77); here all
// batches are accumulated in one pass over the columns, and when C has rank one
// (C[col][batch] = u[col] * w[batch], the case of every column sampled at the same
// points, because the random powers advance column by column) a single sum
// S = sum_col u[col] * col[row] is formed and each batch output is w[batch] * S - B[batch].
// Products are accumulated lazily in 64 bits (three 62-bit products between folds).
//
// compute_quotients_and_combine (cpu/quotients.rs:51-101): on the subdomain of
// log L - b, q[row] = sum_k (partial_k[lift(row)] - first_linear_k * p.y) * den_k(p)^-1
// with den (CM31) from crates/stwo/src/core/pcs/quotients.rs:233-256, then
// interpolate on the subdomain and evaluate on the full domain of log L.
#include <algorithm>

#include "common.h"

int sh_fft_interpolate_batch(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_src,
                             uint32_t* const* d_dst, uint32_t n_cols, uint32_t m, uint32_t Pl, uint32_t half_init);
int sh_fft_evaluate_batch(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_src,
                          uint32_t* const* d_dst, uint32_t n_cols, uint32_t n, uint32_t m, uint32_t Pl,
                          uint32_t half_init);

namespace {

constexpr int kMaxOut = 8;

struct OutSpec {
  uint32_t n_out;
  uint32_t src[kMaxOut];      // which accumulated sum feeds output b
  sh::qm31 w[kMaxOut];        // output b = w * S_src - B
  sh::qm31 B[kMaxOut];
  // outputs whose bit is set also subtract FL_b * p.y, p the row's point of the domain
  // CircleDomain(half) in bit-reversed order (the first-linear term of the combine, taken here
  // when the numerator is already at the combine's size, so the combine skips it)
  uint32_t fl_mask;
  sh::qm31 FL[kMaxOut];
  sh_point_tab half;
};

__device__ __forceinline__ void lazy_fold(uint64_t& v) { v = (v & sh::P) + (v >> 31); }
__device__ __forceinline__ uint32_t lazy_final(uint64_t v) {
  v = (v & sh::P) + (v >> 31);
  v = (v & sh::P) + (v >> 31);
  return sh::reduce_once((uint32_t)v);
}

// cols: device table of n_cols column pointers; coef: n_cols x NB x 4 words (canonical);
// outs: device table of 4 * n_out coordinate pointers.
template <int NB>
__global__ __launch_bounds__(256) void k_accum_num(const uint32_t* const* __restrict__ cols, uint32_t n_cols,
                                                   const uint32_t* __restrict__ coef, uint32_t rows,
                                                   uint32_t* const* __restrict__ outs, OutSpec spec,
                                                   const uint32_t* __restrict__ dev_B,
                                                   const uint32_t* __restrict__ dev_FL) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  uint64_t acc[NB][4];
#pragma unroll
  for (int j = 0; j < NB; j++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[j][c] = 0;
  auto mac = [&](uint32_t v, uint32_t col) {
    const uint32_t* cf = coef + (size_t)col * NB * 4;
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int c = 0; c < 4; c++) acc[j][c] += (uint64_t)cf[j * 4 + c] * v;
  };
  // three products of < 2^62 on a folded value stay below 2^64
  auto fold_all = [&]() {
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int c = 0; c < 4; c++) lazy_fold(acc[j][c]);
  };
  uint32_t col = 0;
  // six columns per step, their loads issued together
  for (; col + 6 <= n_cols; col += 6) {
    uint32_t v[6];
#pragma unroll
    for (int u = 0; u < 6; u++) v[u] = cols[col + u][r];
#pragma unroll
    for (int u = 0; u < 6; u++) {
      mac(v[u], col + u);
      if (u == 2 || u == 5) fold_all();
    }
  }
  for (; col < n_cols; col++) {
    mac(cols[col][r], col);
    fold_all();
  }
  sh::qm31 S[NB];
#pragma unroll
  for (int j = 0; j < NB; j++)
    S[j] = {{lazy_final(acc[j][0]), lazy_final(acc[j][1])}, {lazy_final(acc[j][2]), lazy_final(acc[j][3])}};
  uint32_t py = 0;
  if (spec.fl_mask) {
    const uint32_t sub_log = 31 - __builtin_clz(rows);
    py = sh_circle_domain_point(spec.half, sub_log - 1, sh::bitrev(r, sub_log)).y;
  }
  for (uint32_t b = 0; b < spec.n_out; b++) {
    sh::qm31 s = S[0];
#pragma unroll
    for (int j = 1; j < NB; j++)
      if (spec.src[b] == (uint32_t)j) s = S[j];
    // dev_B / dev_FL: the constant terms in device memory (a driver that computed them there)
    const sh::qm31 Bb = dev_B ? sh::qload(dev_B + 4 * b) : spec.B[b];
    sh::qm31 o = sh::qsub(sh::qmul(spec.w[b], s), Bb);
    if ((spec.fl_mask >> b) & 1) o = sh::qsub(o, sh::qmulm(dev_FL ? sh::qload(dev_FL + 4 * b) : spec.FL[b], py));
    outs[4 * b + 0][r] = o.a.a;
    outs[4 * b + 1][r] = o.a.b;
    outs[4 * b + 2][r] = o.b.a;
    outs[4 * b + 3][r] = o.b.b;
  }
}

// Per-accumulation constants of the combine kernel (host-computed, 16 words each):
//   den_k(p) = (Re px_k - p.x) * Im py_k - (Re py_k - p.y) * Im px_k
//            = c_k - p.x * Im py_k + p.y * Im px_k,   c_k = Re px_k * Im py_k - Re py_k * Im px_k
// (crates/stwo/src/core/pcs/quotients.rs:233-256), all CM31; and first_linear_k (f0 + f1 i,
// f2 + f3 i) as -f and the two imaginary parts, so that every product below is an addition.
struct CombineK {
  uint32_t c[2];      // c_k
  uint32_t npiy[2];   // -Im py_k
  uint32_t pix[2];    // Im px_k
  uint32_t nfl[4];    // -first_linear_k
  uint32_t flim[2];   // f1, f3
  uint32_t plog;      // log size of partial numerator k
  uint32_t pad[3];
};
static_assert(sizeof(CombineK) == 64, "CombineK layout");


// (u64) a * b + c reduced to canonical, for a, b, c < P: the sum is at most P^2 - P, whose two
// 31-bit halves add to at most 2P - 2.
__device__ __forceinline__ uint32_t mad_red(uint32_t a, uint32_t b, uint32_t c) {
  const uint64_t v = (uint64_t)a * b + c;
  return sh::reduce_once(((uint32_t)v & sh::P) + (uint32_t)(v >> 31));
}
// c + a * b + d * e reduced to canonical, for canonical a, b, c, d, e: the sum is below 2^63,
// so one fold and two conditional subtractions bring it to [0, P).
__device__ __forceinline__ uint32_t mad2_red(uint32_t a, uint32_t b, uint32_t d, uint32_t e, uint32_t c) {
  const uint64_t v = (uint64_t)a * b + (uint64_t)d * e + c;
  const uint64_t f = (v & sh::P) + (v >> 31);  // below 2^32 + 2^31
  const uint32_t g = (uint32_t)(f & sh::P) + (uint32_t)(f >> 31);  // below 2^31 + 2
  return sh::reduce_once(g);
}
// v += a * b with v kept below 2^64: at most four products of two canonical words between folds.
__device__ __forceinline__ void mac64(uint64_t& v, uint32_t a, uint32_t b) { v += (uint64_t)a * b; }

// Loads and stores through a base address held in scalar registers and a 32-bit byte offset
// (global_load/global_store with an SGPR base: no 64-bit address arithmetic per access).
typedef const __attribute__((address_space(1))) uint32_t* gcptr;
typedef __attribute__((address_space(1))) uint32_t* gptr;
__device__ __forceinline__ uint32_t gld(const uint32_t* base, uint32_t byte_off) {
  return *(gcptr)((const __attribute__((address_space(1))) char*)base + byte_off);
}
__device__ __forceinline__ void gst(uint32_t* base, uint32_t byte_off, uint32_t v) {
  *(gptr)((__attribute__((address_space(1))) char*)base + byte_off) = v;
}

// Accumulations per combine launch (their partial-numerator addresses travel by value).
constexpr int kCombineMaxAcc = 8;
constexpr int kCombineMaxRb = 8;
struct CombineArgs {
  const uint32_t* part[kCombineMaxAcc][4];  // partial numerator k, coordinate c
  uint32_t* out[4];
  sh::mpt rot[kCombineMaxRb];  // rotation of the half-coset point from a lane's row 0 to its row j
  uint32_t n_acc;              // accumulations of this launch (kc[0 .. n_acc))
  uint32_t accumulate;         // add to what out holds (a later launch of a long sum)
};

// compute_quotients_and_combine's row loop (crates/stwo/src/prover/backend/cpu/quotients.rs:
// 51-101): q[row] = sum_k (partial_k[lift(row)] - first_linear_k * p.y) / den_k(p).  Each lane
// takes RB rows (stride 256, so every load and store of a wave is coalesced) and inverts the
// denominators of its rows for NK accumulations at once: den^-1 = conj(den) / |den|^2 with the
// M31 norms batch-inverted (Montgomery's trick; the SIMD backend batch-inverts its denominators
// too, crates/stwo/src/prover/backend/simd/quotients.rs:255-283): one exponentiation per lane per
// RB x NK denominators.  A lane's rows differ only in bits that the bit reversal sends below the
// bits of its first row, so each row's domain point is the first row's point rotated by a fixed
// point (one point addition per row).  Products are summed lazily in 64 bits.
template <int NK, int RB, bool FL>
__global__ __launch_bounds__(256) void k_combine(const CombineArgs a, const CombineK* __restrict__ kc,
                                                 uint32_t sub_log, sh_point_tab half) {
  const uint32_t rows = 1u << sub_log;
  const uint32_t r0 = blockIdx.x * (256u * RB) + threadIdx.x;
  uint32_t px[RB], py[RB];
  {
    // CircleDomain(half).at(bitrev(r0)): the second half of the domain is the conjugate
    const uint32_t j = sub_log ? sh::bitrev(r0 & (rows - 1), sub_log) : 0;
    const uint32_t hs = rows >> 1;
    const bool conj = j >= hs;
    const sh::mpt p0 = sh_tab_point(half, conj ? j - hs : j);
#pragma unroll
    for (int k = 0; k < RB; k++) {
      const sh::mpt p = k == 0 ? p0 : sh::padd(p0, a.rot[k]);
      px[k] = p.x;
      py[k] = conj ? sh::mneg(p.y) : p.y;
    }
  }
  uint64_t acc[RB][4];
#pragma unroll
  for (int j = 0; j < RB; j++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[j][c] = 0;
  if (a.accumulate) {
#pragma unroll
    for (int j = 0; j < RB; j++) {
      const uint32_t r = r0 + 256u * j;
      if (r < rows)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[j][c] = gld(a.out[c], 4 * r);
    }
  }
  for (uint32_t k0 = 0; k0 < a.n_acc; k0 += NK) {
    // denominators and their norms; an absent accumulation (the last chunk may be short) takes
    // norm 1, and a zero norm is replaced by 1 with its inverse forced to 0
    uint32_t d0[NK][RB], d1[NK][RB], nrm[NK][RB], pre[NK * RB];
    uint32_t run = 1;
#pragma unroll
    for (int kk = 0; kk < NK; kk++) {
      const bool has = k0 + kk < a.n_acc;
      const CombineK K = kc[has ? k0 + kk : 0];
#pragma unroll
      for (int j = 0; j < RB; j++) {
        const uint32_t x = mad2_red(py[j], K.pix[0], px[j], K.npiy[0], K.c[0]);
        const uint32_t y = mad2_red(py[j], K.pix[1], px[j], K.npiy[1], K.c[1]);
        d0[kk][j] = x;
        d1[kk][j] = y;
        const uint32_t n = mad2_red(x, x, y, y, 0);
        nrm[kk][j] = (!has || n == 0) ? 1u : n;
        run = sh::mmul(run, nrm[kk][j]);
        pre[kk * RB + j] = run;
      }
    }
    uint32_t inv = sh::minv(run);
#pragma unroll
    for (int kk = NK - 1; kk >= 0; kk--) {
      const uint32_t k = k0 + kk;
      const bool has = k < a.n_acc;
      const CombineK K = kc[has ? k : 0];
      const uint32_t ratio = sub_log - K.plog;
#pragma unroll
      for (int j = RB - 1; j >= 0; j--) {
        const int e = kk * RB + j;
        const uint32_t x = d0[kk][j], y = d1[kk][j];
        uint32_t ninv = e > 0 ? sh::mmul(inv, pre[e - 1]) : inv;  // 1 / norm of element e
        inv = sh::mmul(inv, nrm[kk][j]);
        if (x == 0 && y == 0) ninv = 0;
        const uint32_t r = r0 + 256u * j;
        if (!has || r >= rows) continue;
        // den^-1 = (x - y i) / norm = (i0, -i1); g = p.y * den^-1 = (g0, g1)
        const uint32_t i0 = sh::mmul(x, ninv), i1 = sh::mmul(y, ninv);
        const uint32_t ni1 = sh::P - i1;  // -Im den^-1, in [1, P]
        // (FL false: every first-linear term is zero, taken by the numerators already)
        const uint32_t g0 = FL ? sh::mmul(py[j], i0) : 0, g1 = FL ? sh::mmul(py[j], ni1) : 0;
        const uint32_t lo = 4 * (((r >> (ratio + 1)) << 1) + (r & 1));
        uint32_t pv0, pv1, pv2, pv3;
        // the accumulation index is uniform: its addresses stay in scalar registers
        switch (k) {
#define SH_PART(K_)                   \
  case K_:                            \
    pv0 = gld(a.part[K_][0], lo);     \
    pv1 = gld(a.part[K_][1], lo);     \
    pv2 = gld(a.part[K_][2], lo);     \
    pv3 = gld(a.part[K_][3], lo);     \
    break;
          SH_PART(0) SH_PART(1) SH_PART(2) SH_PART(3) SH_PART(4) SH_PART(5) SH_PART(6)
          default:
            pv0 = gld(a.part[7][0], lo);
            pv1 = gld(a.part[7][1], lo);
            pv2 = gld(a.part[7][2], lo);
            pv3 = gld(a.part[7][3], lo);
            break;
#undef SH_PART
        }
        // + partial * den^-1: each CM31 half (u0 + u1 i)(i0 - i1 i) = (u0 i0 + u1 i1) + (u1 i0 - u0 i1) i
        mac64(acc[j][0], pv0, i0);
        mac64(acc[j][0], pv1, i1);
        mac64(acc[j][1], pv1, i0);
        mac64(acc[j][1], pv0, ni1);
        mac64(acc[j][2], pv2, i0);
        mac64(acc[j][2], pv3, i1);
        mac64(acc[j][3], pv3, i0);
        mac64(acc[j][3], pv2, ni1);
        if (FL) {
          // - first_linear * g: each CM31 half -(f0 + f1 i)(g0 + g1 i) = (-f0 g0 + f1 g1) + (-f0 g1 - f1 g0) i
          mac64(acc[j][0], K.nfl[0], g0);
          mac64(acc[j][0], K.flim[0], g1);
          mac64(acc[j][1], K.nfl[0], g1);
          mac64(acc[j][1], K.nfl[1], g0);
          mac64(acc[j][2], K.nfl[2], g0);
          mac64(acc[j][2], K.flim[1], g1);
          mac64(acc[j][3], K.nfl[2], g1);
          mac64(acc[j][3], K.nfl[3], g0);
        }
        // four products since the last fold with FL, two without: fold every accumulation, or
        // every second one
        if (FL || (kk & 1) || kk == 0)
#pragma unroll
          for (int c = 0; c < 4; c++) lazy_fold(acc[j][c]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < RB; j++) {
    const uint32_t r = r0 + 256u * j;
    if (r >= rows) continue;
#pragma unroll
    for (int c = 0; c < 4; c++) gst(a.out[c], 4 * r, lazy_final(acc[j][c]));
  }
}

// out[i] = sum_k col_k[lift_k(i)] for every coordinate.
__global__ __launch_bounds__(256) void k_lift_acc(const uint32_t* const* __restrict__ cols,
                                                  const uint32_t* __restrict__ logs, uint32_t n, uint32_t L,
                                                  uint32_t* const* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1u << L)) return;
  uint32_t s[4] = {0, 0, 0, 0};
  for (uint32_t k = 0; k < n; k++) {
    uint32_t ratio = L - logs[k];
    uint32_t li = ((i >> (ratio + 1)) << 1) + (i & 1);
#pragma unroll
    for (int c = 0; c < 4; c++) s[c] = sh::madd(s[c], cols[4 * k + c][li]);
  }
#pragma unroll
  for (int c = 0; c < 4; c++) out[c][i] = s[c];
}

__global__ __launch_bounds__(256) void k_accumulate(uint32_t* const* __restrict__ col,
                                                    const uint32_t* const* __restrict__ other, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 4; c++) col[c][i] = sh::madd(col[c][i], other[c][i]);
}

// complex_conjugate_line_coeffs (crates/stwo/src/core/constraints.rs:119-134), host side.
void line_coeffs(const sh::qpt& p, sh::qm31 value, sh::qm31 alpha, sh::qm31* a, sh::qm31* b, sh::qm31* c) {
  sh::qm31 av = sh::qsub(sh::qconj(value), value);
  sh::qm31 cv = sh::qsub(sh::qconj(p.y), p.y);
  sh::qm31 bv = sh::qsub(sh::qmul(value, cv), sh::qmul(av, p.y));
  *a = sh::qmul(alpha, av);
  *b = sh::qmul(alpha, bv);
  *c = sh::qmul(alpha, cv);
}

bool qis_zero(sh::qm31 x) { return sh::qeq(x, sh::qzero()); }

}  // namespace

// apply_fl (optional, one flag per batch): that batch's output also subtracts its first-linear
// term times the row's p.y (OutSpec::fl_mask); its first_linear still reports the term.
int sh_accumulate_numerators_fl(sh_ctx* ctx, const uint32_t* const* d_cols, uint64_t n_cols, uint32_t col_log,
                                uint32_t log_blowup, const uint32_t* batch_points, const uint32_t* batch_term_offsets,
                                const uint32_t* term_col_idx, const uint32_t* term_values,
                                const uint32_t* term_random_coeffs, uint64_t n_batches, uint32_t* const* d_outs,
                                uint32_t* first_linear, const uint8_t* apply_fl) {
  if (!ctx || (n_batches && (!batch_points || !batch_term_offsets || !d_outs || !first_linear)))
    return SH_ERR_INVALID;
  if (n_batches == 0) return SH_OK;
  SH_REQUIRE(ctx, n_cols > 0 && d_cols, "accumulate_numerators needs a non-empty set of columns");
  SH_REQUIRE(ctx, col_log >= log_blowup && col_log <= 30, "bad column log size %u", col_log);
  const uint64_t NB = n_batches;
  // dense coefficients C[col][batch], B[batch], first_linear[batch]
  std::vector<sh::qm31> C(n_cols * NB, sh::qzero()), B(NB, sh::qzero()), FL(NB, sh::qzero());
  for (uint64_t bi = 0; bi < NB; bi++) {
    sh::qpt p = {sh::qload(batch_points + 8 * bi), sh::qload(batch_points + 8 * bi + 4)};
    SH_REQUIRE(ctx, !sh::qeq(p.y, sh::qconj(p.y)), "Cannot evaluate a line with a single point");
    for (uint32_t t = batch_term_offsets[bi]; t < batch_term_offsets[bi + 1]; t++) {
      uint32_t ci = term_col_idx[t];
      SH_REQUIRE(ctx, ci < n_cols, "term column index %u out of range", ci);
      sh::qm31 a, b, c;
      line_coeffs(p, sh::qload(term_values + 4 * t), sh::qload(term_random_coeffs + 4 * t), &a, &b, &c);
      C[ci * NB + bi] = sh::qadd(C[ci * NB + bi], c);
      B[bi] = sh::qadd(B[bi], b);
      FL[bi] = sh::qadd(FL[bi], a);
    }
  }
  for (uint64_t bi = 0; bi < NB; bi++) sh::qstore(FL[bi], first_linear + 4 * bi);
  // rank-one detection: C[col][b] == w[b] * u[col] with u = C[.][pivot batch]
  uint64_t pivot_col = n_cols;
  for (uint64_t ci = 0; ci < n_cols && pivot_col == n_cols; ci++)
    for (uint64_t bi = 0; bi < NB; bi++)
      if (!qis_zero(C[ci * NB + bi])) {
        pivot_col = ci;
        break;
      }
  bool rank1 = false;
  std::vector<sh::qm31> w(NB, sh::qzero());
  uint64_t pivot_b = 0;
  if (pivot_col < n_cols) {
    while (qis_zero(C[pivot_col * NB + pivot_b])) pivot_b++;
    sh::qm31 inv = sh::qinv(C[pivot_col * NB + pivot_b]);
    for (uint64_t bi = 0; bi < NB; bi++) w[bi] = sh::qmul(C[pivot_col * NB + bi], inv);
    rank1 = true;
    for (uint64_t ci = 0; ci < n_cols && rank1; ci++)
      for (uint64_t bi = 0; bi < NB; bi++)
        if (!sh::qeq(C[ci * NB + bi], sh::qmul(w[bi], C[ci * NB + pivot_b]))) {
          rank1 = false;
          break;
        }
  }
  const uint32_t rows = 1u << (col_log - log_blowup);
  sh_point_tab half = {};
  bool any_fl = false;
  for (uint64_t bi = 0; bi < NB && apply_fl; bi++) any_fl |= apply_fl[bi] != 0;
  if (any_fl) {
    // the rows' domain: the subdomain of the combine, CircleDomain(half_odds(L - 1).initial,
    // log - 1) with L = col_log (domain.rs:72-82 split)
    SH_REQUIRE(ctx, col_log > log_blowup, "column too small for the blowup");
    SH_TRY(sh_point_tab_make(ctx, sh::subgroup_gen(col_log + 1), col_log - log_blowup - 1, &half));
  }
  auto set_fl = [&](OutSpec& spec, uint32_t k, uint64_t bi) {
    if (apply_fl && apply_fl[bi]) {
      spec.fl_mask |= 1u << k;
      spec.FL[k] = FL[bi];
      spec.half = half;
    }
  };
  void* d_cols_tab;
  SH_TRY(sh_stage(ctx, d_cols, n_cols * sizeof(void*), &d_cols_tab));
  auto launch = [&](int nb, const std::vector<uint32_t>& coef_words, OutSpec spec,
                    const std::vector<const void*>& out_ptrs) -> int {
    void* d_coef;
    void* d_outs_tab;
    SH_TRY(sh_stage(ctx, coef_words.data(), coef_words.size() * 4, &d_coef));
    SH_TRY(sh_stage(ctx, out_ptrs.data(), out_ptrs.size() * sizeof(void*), &d_outs_tab));
    unsigned grid = (rows + 255) / 256;
    const uint32_t* const* ct = (const uint32_t* const*)d_cols_tab;
    uint32_t* const* ot = (uint32_t* const*)d_outs_tab;
    SpanGuard span(ctx, "quotients_accumulate_numerators");
    switch (nb) {
      case 1: k_accum_num<1><<<grid, 256, 0, ctx->stream>>>(ct, (uint32_t)n_cols, (const uint32_t*)d_coef, rows, ot, spec, nullptr, nullptr); break;
      case 2: k_accum_num<2><<<grid, 256, 0, ctx->stream>>>(ct, (uint32_t)n_cols, (const uint32_t*)d_coef, rows, ot, spec, nullptr, nullptr); break;
      case 3: k_accum_num<3><<<grid, 256, 0, ctx->stream>>>(ct, (uint32_t)n_cols, (const uint32_t*)d_coef, rows, ot, spec, nullptr, nullptr); break;
      default: k_accum_num<4><<<grid, 256, 0, ctx->stream>>>(ct, (uint32_t)n_cols, (const uint32_t*)d_coef, rows, ot, spec, nullptr, nullptr); break;
    }
    return sh_launch_check(ctx, "accumulate_numerators");
  };
  if (rank1 || pivot_col == n_cols) {
    // one pass: S = sum u[col] * col, outputs in groups of kMaxOut
    std::vector<uint32_t> coef(n_cols * 4);
    for (uint64_t ci = 0; ci < n_cols; ci++)
      sh::qstore(pivot_col < n_cols ? C[ci * NB + pivot_b] : sh::qzero(), &coef[ci * 4]);
    for (uint64_t b0 = 0; b0 < NB; b0 += kMaxOut) {
      OutSpec spec = {};
      std::vector<const void*> outp;
      spec.n_out = (uint32_t)std::min<uint64_t>(kMaxOut, NB - b0);
      for (uint32_t k = 0; k < spec.n_out; k++) {
        spec.src[k] = 0;
        spec.w[k] = w[b0 + k];
        spec.B[k] = B[b0 + k];
        set_fl(spec, k, b0 + k);
        for (int c = 0; c < 4; c++) outp.push_back(d_outs[4 * (b0 + k) + c]);
      }
      SH_TRY(launch(1, coef, spec, outp));
    }
  } else {
    for (uint64_t b0 = 0; b0 < NB; b0 += 4) {
      int nb = (int)std::min<uint64_t>(4, NB - b0);
      std::vector<uint32_t> coef(n_cols * nb * 4);
      OutSpec spec = {};
      std::vector<const void*> outp;
      spec.n_out = nb;
      for (int j = 0; j < nb; j++) {
        for (uint64_t ci = 0; ci < n_cols; ci++) sh::qstore(C[ci * NB + b0 + j], &coef[(ci * nb + j) * 4]);
        spec.src[j] = j;
        spec.w[j] = sh::qone();
        spec.B[j] = B[b0 + j];
        set_fl(spec, j, b0 + j);
        for (int c = 0; c < 4; c++) outp.push_back(d_outs[4 * (b0 + j) + c]);
      }
      SH_TRY(launch(nb, coef, spec, outp));
    }
  }
  return SH_OK;
}

extern "C" int sh_accumulate_numerators(sh_ctx* ctx, const uint32_t* const* d_cols, uint64_t n_cols,
                                        uint32_t col_log, uint32_t log_blowup, const uint32_t* batch_points,
                                        const uint32_t* batch_term_offsets, const uint32_t* term_col_idx,
                                        const uint32_t* term_values, const uint32_t* term_random_coeffs,
                                        uint64_t n_batches, uint32_t* const* d_outs, uint32_t* first_linear) {
  return sh_accumulate_numerators_fl(ctx, d_cols, n_cols, col_log, log_blowup, batch_points, batch_term_offsets,
                                     term_col_idx, term_values, term_random_coeffs, n_batches, d_outs, first_linear,
                                     nullptr);
}

// The allocation-free variant's scratch: the subdomain quotient (4 coordinate arrays of up to
// 2^kStaticQuotLog words, 32 MiB), one copy per device for the life of the process.  Host code never
// names it directly: sh_static_scratch_init resolves each device's address once.
constexpr uint32_t kStaticQuotLog = 21;
static __device__ __align__(256) uint32_t scratch[4u << kStaticQuotLog];
constexpr int kMaxDevices = 64;
static uint32_t* g_scratch_ptr[kMaxDevices];  // per device: the address of `scratch` on it
static std::mutex g_scratch_mu;

// Called by sh_ctx_create with ctx->device current: resolves that device's copy of `scratch`.
int sh_static_scratch_init(sh_ctx* ctx) {
  SH_REQUIRE(ctx, ctx->device >= 0 && ctx->device < kMaxDevices, "device %d beyond the static scratch table",
             ctx->device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[ctx->device]) return SH_OK;
  void* ptr = nullptr;
  SH_CHECK_HIP(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  SH_REQUIRE(ctx, ptr != nullptr, "hipGetSymbolAddress returned no address for the static scratch");
  g_scratch_ptr[ctx->device] = (uint32_t*)ptr;
  return SH_OK;
}

static int combine_and_extend(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* h_parts,
                              const CombineK* d_kc, uint64_t n_acc, uint32_t L, uint32_t b, uint32_t* const* d_out,
                              uint32_t* q_scratch, bool with_fl);

// q_scratch: 4 << (L - b) words for the subdomain quotient, or nullptr to use the context's
// (lazily allocated) scratch slot.
static int quotients_and_combine(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_partials,
                                 const uint32_t* partial_logs, const uint32_t* points, const uint32_t* first_linear,
                                 uint64_t n_acc, uint32_t lifting_log_size, uint32_t log_blowup,
                                 uint32_t* const* d_out, uint32_t* q_scratch) {
  if (!ctx || !tw || !d_out || (n_acc && (!d_partials || !partial_logs || !points || !first_linear)))
    return SH_ERR_INVALID;
  const uint32_t L = lifting_log_size, b = log_blowup;
  SH_REQUIRE(ctx, L >= b + 1 && L <= 30, "bad lifting log size %u / blowup %u", L, b);
  const uint32_t s = L - b;
  SH_REQUIRE(ctx, sh::is_doubling_of(sh::subgroup_gen(L + 1), L - 1, tw->root_initial_index, tw->root_log_size),
             "domain of log size %u is not in the twiddle tower", L);
  std::vector<CombineK> kc(std::max<uint64_t>(n_acc, 1));
  for (uint64_t k = 0; k < n_acc; k++) {
    SH_REQUIRE(ctx, partial_logs[k] >= 1 && partial_logs[k] <= s, "partial numerator log size %u not in [1, %u]",
               partial_logs[k], s);
    const uint32_t* pt = points + 8 * k;  // x = (prx, pix), y = (pry, piy) as CM31 pairs
    const sh::cm31 prx = {pt[0], pt[1]}, pix = {pt[2], pt[3]}, pry = {pt[4], pt[5]}, piy = {pt[6], pt[7]};
    const sh::cm31 c = sh::csub(sh::cmul(prx, piy), sh::cmul(pry, pix));
    const sh::cm31 npiy = sh::cneg(piy);
    CombineK& K = kc[k];
    K.c[0] = c.a;
    K.c[1] = c.b;
    K.npiy[0] = npiy.a;
    K.npiy[1] = npiy.b;
    K.pix[0] = pix.a;
    K.pix[1] = pix.b;
    const uint32_t* fl = first_linear + 4 * k;
    for (int w = 0; w < 4; w++) K.nfl[w] = sh::mneg(fl[w]);
    K.flim[0] = fl[1];
    K.flim[1] = fl[3];
    K.plog = partial_logs[k];
    K.pad[0] = K.pad[1] = K.pad[2] = 0;
  }
  bool with_fl = false;
  for (uint64_t i = 0; i < 4 * n_acc; i++) with_fl |= first_linear[i] != 0;
  void* d_acc;
  SH_TRY(sh_stage(ctx, kc.data(), kc.size() * sizeof(CombineK), &d_acc));
  return combine_and_extend(ctx, tw, d_partials, (const CombineK*)d_acc, n_acc, L, b, d_out, q_scratch, with_fl);
}

// The combine over the subdomain of log L - b (k_combine, per-accumulation constants in device
// memory) and the quotient's extension: interpolate on the subdomain, evaluate on L.  with_fl
// false: every first-linear term is zero (the numerators took it), and the combine skips it.
static int combine_and_extend(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* h_parts,
                              const CombineK* d_kc, uint64_t n_acc, uint32_t L, uint32_t b, uint32_t* const* d_out,
                              uint32_t* q_scratch, bool with_fl) {
  const uint32_t s = L - b;
  // subdomain: CircleDomain(Coset(half_odds(L-1).initial, s - 1)) (domain.rs:72-82 split)
  const uint32_t half_init = sh::subgroup_gen(L + 1);
  sh_point_tab half;
  SH_TRY(sh_point_tab_make(ctx, half_init, s - 1, &half));
  uint32_t* q = q_scratch;
  if (!q) {
    void* scr;
    SH_TRY(sh_scratch_get(ctx, SCR_QUOT, 16ull << s, &scr));
    q = (uint32_t*)scr;
  }
  const void* qcols[8] = {q, q + (1u << s), q + (2u << s), q + (3u << s), d_out[0], d_out[1], d_out[2], d_out[3]};
  void* d_q;
  SH_TRY(sh_stage(ctx, qcols, sizeof(qcols), &d_q));
  uint32_t* const* d_qtab = (uint32_t* const*)d_q;
  {
    SpanGuard span(ctx, "quotients_combine");
    // accumulations per batch inversion: all of a launch's up to 4 (every inverted norm is used);
    // rows per lane: 8 for one or two accumulations, 4 otherwise (the register budget)
    const int nk = n_acc <= 1 ? 1 : n_acc == 2 ? 2 : n_acc == 3 ? 3 : 4;
    const int rb = nk <= 2 ? 8 : 4;
    CombineArgs a = {};
    for (int c = 0; c < 4; c++) a.out[c] = q + ((uint64_t)c << s);
    // row j of a lane is row 0 plus 256 j: its bit-reversed index adds bitrev(256 j), below the
    // first row's bits, i.e. a rotation by that many steps of the half coset
    const uint32_t step = sh::subgroup_gen(s - 1);
    for (int j = 0; j < rb; j++) {
      const uint64_t off = 256ull * j;
      const uint32_t d = off < (1ull << s) ? sh::bitrev((uint32_t)off, s) : 0;
      a.rot[j] = sh::index_to_point_host(sh::idx_reduce((uint64_t)d * step));
    }
    const unsigned grid = (unsigned)(((1ull << s) + 256ull * rb - 1) / (256ull * rb));
    for (uint64_t k0 = 0; k0 < std::max<uint64_t>(n_acc, 1); k0 += kCombineMaxAcc) {
      const uint64_t na = std::min<uint64_t>(kCombineMaxAcc, n_acc - std::min(n_acc, k0));
      a.n_acc = (uint32_t)na;
      a.accumulate = k0 > 0;
      for (uint64_t k = 0; k < kCombineMaxAcc; k++)
        for (int c = 0; c < 4; c++) a.part[k][c] = k < na ? h_parts[4 * (k0 + k) + c] : nullptr;
      const CombineK* kt = d_kc + k0;
#define SH_COMBINE(NK_, RB_)                                                            \
  (with_fl ? k_combine<NK_, RB_, true><<<grid, 256, 0, ctx->stream>>>(a, kt, s, half)  \
           : k_combine<NK_, RB_, false><<<grid, 256, 0, ctx->stream>>>(a, kt, s, half))
      if (nk == 1)
        SH_COMBINE(1, 8);
      else if (nk == 2)
        SH_COMBINE(2, 8);
      else if (nk == 3)
        SH_COMBINE(3, 4);
      else
        SH_COMBINE(4, 4);
#undef SH_COMBINE
      SH_TRY(sh_launch_check(ctx, "combine"));
    }
  }
  // interpolate on the subdomain (its twiddles are prefixes of the domain-L layers), evaluate on L
  SH_TRY(sh_fft_interpolate_batch(ctx, tw, (const uint32_t* const*)d_qtab, d_qtab, 4, s, L, half_init));
  SH_TRY(sh_fft_evaluate_batch(ctx, tw, (const uint32_t* const*)d_qtab, d_qtab + 4, 4, s, L, L, half_init));
  return SH_OK;
}

extern "C" int sh_compute_quotients_and_combine(sh_ctx* ctx, const sh_twiddles* tw,
                                                const uint32_t* const* d_partials, const uint32_t* partial_logs,
                                                const uint32_t* points, const uint32_t* first_linear,
                                                uint64_t n_acc, uint32_t lifting_log_size, uint32_t log_blowup,
                                                uint32_t* const* d_out) {
  return quotients_and_combine(ctx, tw, d_partials, partial_logs, points, first_linear, n_acc, lifting_log_size,
                               log_blowup, d_out, nullptr);
}

// The same with the subdomain quotient in the device's static scratch: no device allocation.
// Every call on one device shares that scratch (which keeps the last call's contents; nothing
// in it needs to start at zero, every word is written before it is read), so calls that overlap
// on one device race.
extern "C" int sh_compute_quotients_and_combine_noalloc(sh_ctx* ctx, const sh_twiddles* tw,
                                                        const uint32_t* const* d_partials,
                                                        const uint32_t* partial_logs, const uint32_t* points,
                                                        const uint32_t* first_linear, uint64_t n_acc,
                                                        uint32_t lifting_log_size, uint32_t log_blowup,
                                                        uint32_t* const* d_out) {
  if (!ctx) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, lifting_log_size >= log_blowup + 1 && lifting_log_size - log_blowup <= kStaticQuotLog,
             "quotient subdomain of log %u exceeds the static scratch (at most 2^%u rows per coordinate)",
             lifting_log_size - log_blowup, kStaticQuotLog);
  uint32_t* q = ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch_ptr[ctx->device] : nullptr;
  SH_REQUIRE(ctx, q != nullptr, "the static scratch of device %d was not resolved at context creation",
             ctx->device);
  // the four coordinate arrays packed at the subdomain's size (4 << s words in all)
  return quotients_and_combine(ctx, tw, d_partials, partial_logs, points, first_linear, n_acc, lifting_log_size,
                               log_blowup, d_out, q);
}

extern "C" int sh_lift_and_accumulate(sh_ctx* ctx, const uint32_t* const* d_cols, const uint32_t* log_sizes,
                                      uint64_t n, uint32_t* const* d_out) {
  if (!ctx || !d_out || (n && (!d_cols || !log_sizes))) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, n > 0, "lift_and_accumulate of no columns");
  SH_REQUIRE(ctx, log_sizes[0] >= 1, "A column must be of length at least 2.");
  for (uint64_t k = 1; k < n; k++)
    SH_REQUIRE(ctx, log_sizes[k] > log_sizes[k - 1], "columns must be strictly ascending in size");
  const uint32_t L = log_sizes[n - 1];
  SH_REQUIRE(ctx, L <= 30, "log size %u too large", L);
  void *d_tab, *d_logs, *d_o;
  SH_TRY(sh_stage(ctx, d_cols, 4 * n * sizeof(void*), &d_tab));
  SH_TRY(sh_stage(ctx, log_sizes, n * 4, &d_logs));
  SH_TRY(sh_stage(ctx, d_out, 4 * sizeof(void*), &d_o));
  SpanGuard span(ctx, "lift_and_accumulate");
  k_lift_acc<<<((1u << L) + 255) / 256, 256, 0, ctx->stream>>>((const uint32_t* const*)d_tab,
                                                               (const uint32_t*)d_logs, (uint32_t)n, L,
                                                               (uint32_t* const*)d_o);
  return sh_launch_check(ctx, "lift_and_accumulate");
}

extern "C" int sh_accumulate(sh_ctx* ctx, uint32_t* const* d_col, const uint32_t* const* d_other, uint32_t log_size) {
  if (!ctx || !d_col || !d_other) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, log_size <= 30, "log size %u too large", log_size);
  void *d_a, *d_b;
  SH_TRY(sh_stage(ctx, d_col, 4 * sizeof(void*), &d_a));
  SH_TRY(sh_stage(ctx, d_other, 4 * sizeof(void*), &d_b));
  SpanGuard span(ctx, "accumulate");
  k_accumulate<<<((1u << log_size) + 255) / 256, 256, 0, ctx->stream>>>((uint32_t* const*)d_a,
                                                                        (const uint32_t* const*)d_b, 1u << log_size);
  return sh_launch_check(ctx, "accumulate");
}

// ---------------------------------------------------------------- one device point
// compute_fri_quotients (crates/stwo/src/prover/pcs/quotient_ops.rs:86-160) when every sampled
// column has exactly one sample, all at one point p held in device memory with the sampled values
// and the random coefficient (as a device channel leaves them): the line coefficients, the
// per-size sums and the combine constants are computed on the device, so the whole driver runs
// without the host reading p or the values.
namespace {

// Sorted position i (columns ascending by size, stable) takes column perm[i]: its sample gets the
// random power of its place in the column order (build_samples_with_randomness_and_periodicity,
// crates/stwo/src/core/pcs/quotients.rs:274-319) and complex_conjugate_line_coeffs (crates/stwo/
// src/core/constraints.rs:119-134): a = conj(v) - v, c = conj(p.y) - p.y, b = v c - a p.y, each
// times the power.  coef[i] = c; ba[2i] = b, ba[2i + 1] = a.
__global__ void k_dpoint_terms(const uint32_t* __restrict__ point, const uint32_t* __restrict__ values,
                               const uint32_t* __restrict__ powers, const uint32_t* __restrict__ perm, uint32_t n,
                               uint32_t* __restrict__ coef, uint32_t* __restrict__ ba) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t col = perm[i];
  const sh::qm31 py = sh::qload(point + 4), v = sh::qload(values + 4 * col), pw = sh::qload(powers + 4 * col);
  const sh::qm31 av = sh::qsub(sh::qconj(v), v);
  const sh::qm31 cv = sh::qsub(sh::qconj(py), py);
  const sh::qm31 bv = sh::qsub(sh::qmul(v, cv), sh::qmul(av, py));
  sh::qstore(sh::qmul(pw, cv), coef + 4 * i);
  sh::qstore(sh::qmul(pw, bv), ba + 8 * i);
  sh::qstore(sh::qmul(pw, av), ba + 8 * i + 4);
}

// One workgroup per size class k (sorted positions [off[k], off[k+1])): B[k] = sum of b,
// FL[k] = sum of a.
__global__ __launch_bounds__(256) void k_dpoint_sums(const uint32_t* __restrict__ ba, const uint32_t* __restrict__ off,
                                                     uint32_t* __restrict__ B, uint32_t* __restrict__ FL) {
  __shared__ uint32_t red[256][8];
  const uint32_t k = blockIdx.x, t = threadIdx.x;
  sh::qm31 sb = sh::qzero(), sa = sh::qzero();
  for (uint32_t i = off[k] + t; i < off[k + 1]; i += 256) {
    sb = sh::qadd(sb, sh::qload(ba + 8 * i));
    sa = sh::qadd(sa, sh::qload(ba + 8 * i + 4));
  }
  sh::qstore(sb, red[t]);
  sh::qstore(sa, red[t] + 4);
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (t < w) {
      sh::qstore(sh::qadd(sh::qload(red[t]), sh::qload(red[t + w])), red[t]);
      sh::qstore(sh::qadd(sh::qload(red[t] + 4), sh::qload(red[t + w] + 4)), red[t] + 4);
    }
    __syncthreads();
  }
  if (t < 4) {
    B[4 * k + t] = red[0][t];
    FL[4 * k + t] = red[0][4 + t];
  }
}

// The combine constants of the one accumulation (the point, the first linear terms of every
// size summed), as the host path builds them.
__global__ void k_dpoint_combine_const(const uint32_t* __restrict__ point, const uint32_t* __restrict__ FL,
                                       uint32_t n_fl, uint32_t plog, CombineK* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const sh::cm31 prx = {point[0], point[1]}, pix = {point[2], point[3]}, pry = {point[4], point[5]},
                 piy = {point[6], point[7]};
  const sh::cm31 c = sh::csub(sh::cmul(prx, piy), sh::cmul(pry, pix));
  const sh::cm31 npiy = sh::cneg(piy);
  sh::qm31 fl = sh::qzero();
  for (uint32_t k = 0; k < n_fl; k++) fl = sh::qadd(fl, sh::qload(FL + 4 * k));
  CombineK K;
  K.c[0] = c.a;
  K.c[1] = c.b;
  K.npiy[0] = npiy.a;
  K.npiy[1] = npiy.b;
  K.pix[0] = pix.a;
  K.pix[1] = pix.b;
  uint32_t f[4];
  sh::qstore(fl, f);
  for (int w = 0; w < 4; w++) K.nfl[w] = sh::mneg(f[w]);
  K.flim[0] = f[1];
  K.flim[1] = f[3];
  K.plog = plog;
  K.pad[0] = K.pad[1] = K.pad[2] = 0;
  *out = K;
}

}  // namespace

extern "C" int sh_secure_powers_dev(sh_ctx* ctx, const uint32_t* d_felt, uint64_t n, int reversed, uint32_t* d_out);

extern "C" int sh_fri_quotients_dpoint(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_cols,
                                       const uint32_t* col_log_sizes, uint64_t n_cols, const uint32_t* d_point,
                                       const uint32_t* d_values, const uint32_t* d_random_coeff,
                                       uint32_t lifting_log_size, uint32_t log_blowup, uint32_t* const d_out[4]) {
  if (!ctx || !tw || !d_out || !d_point || !d_random_coeff || (n_cols && (!d_cols || !col_log_sizes || !d_values)))
    return SH_ERR_INVALID;
  const uint32_t L = lifting_log_size, b = log_blowup;
  SH_REQUIRE(ctx, n_cols > 0, "no sampled columns");
  SH_REQUIRE(ctx, L >= b + 1 && L <= 30, "bad lifting log size %u / blowup %u", L, b);
  SH_REQUIRE(ctx, sh::is_doubling_of(sh::subgroup_gen(L + 1), L - 1, tw->root_initial_index, tw->root_log_size),
             "domain of log size %u is not in the twiddle tower", L);
  for (uint64_t i = 0; i < n_cols; i++)
    SH_REQUIRE(ctx, col_log_sizes[i] >= b + 1 && col_log_sizes[i] <= L, "column log size %u not in [%u, %u]",
               col_log_sizes[i], b + 1, L);
  // columns by size, ascending (quotient_ops.rs:108-122), stable
  std::vector<uint32_t> perm(n_cols);
  for (uint64_t i = 0; i < n_cols; i++) perm[i] = (uint32_t)i;
  std::stable_sort(perm.begin(), perm.end(), [&](uint32_t x, uint32_t y) { return col_log_sizes[x] < col_log_sizes[y]; });
  std::vector<uint32_t> off{0}, logs;
  std::vector<const void*> sorted(n_cols);
  for (uint64_t i = 0; i < n_cols; i++) {
    sorted[i] = d_cols[perm[i]];
    if (i > 0 && col_log_sizes[perm[i]] != col_log_sizes[perm[i - 1]]) off.push_back((uint32_t)i);
    if (i == 0 || col_log_sizes[perm[i]] != col_log_sizes[perm[i - 1]]) logs.push_back(col_log_sizes[perm[i]]);
  }
  off.push_back((uint32_t)n_cols);
  const uint32_t n_cls = (uint32_t)logs.size();
  // device work areas (pooled): powers, coefficients, (b, a) terms, B and FL per size, the combine
  // constants, then one partial numerator per size (and a lifted sum when there are several)
  const uint64_t w_pw = 4 * n_cols, w_coef = 4 * n_cols, w_ba = 8 * n_cols, w_B = 4 * n_cls, w_FL = 4 * n_cls;
  const uint64_t w_kc = (sizeof(CombineK) + 3) / 4;
  const uint64_t small_words = w_pw + w_coef + w_ba + w_B + w_FL + w_kc + 64;
  uint32_t* work = nullptr;
  SH_TRY(sh_dev_alloc(ctx, 4 * small_words, (void**)&work));
  uint32_t* d_pw = work;
  uint32_t* d_coef = d_pw + w_pw;
  uint32_t* d_ba = d_coef + w_coef;
  uint32_t* d_B = d_ba + w_ba;
  uint32_t* d_FL = d_B + w_B;
  CombineK* d_kc = (CombineK*)(((uintptr_t)(d_FL + w_FL) + 63) & ~(uintptr_t)63);
  std::vector<uint32_t*> partial(n_cls, nullptr);
  uint32_t* lifted = nullptr;
  auto release = [&]() {
    for (auto p : partial)
      if (p) sh_dev_free(ctx, p);
    if (lifted) sh_dev_free(ctx, lifted);
    sh_dev_free(ctx, work);
  };
  int rc = SH_OK;
  auto run = [&]() -> int {
    void *d_perm, *d_off;
    SH_TRY(sh_stage(ctx, perm.data(), 4 * n_cols, &d_perm));
    SH_TRY(sh_stage(ctx, off.data(), 4 * off.size(), &d_off));
    SH_TRY(sh_secure_powers_dev(ctx, d_random_coeff, n_cols, 0, d_pw));
    {
      SpanGuard span(ctx, "quotients_terms");
      k_dpoint_terms<<<(unsigned)((n_cols + 255) / 256), 256, 0, ctx->stream>>>(
          d_point, d_values, d_pw, (const uint32_t*)d_perm, (uint32_t)n_cols, d_coef, d_ba);
      SH_TRY(sh_launch_check(ctx, "quotient_terms"));
      k_dpoint_sums<<<n_cls, 256, 0, ctx->stream>>>(d_ba, (const uint32_t*)d_off, d_B, d_FL);
      SH_TRY(sh_launch_check(ctx, "quotient_sums"));
    }
    // accumulate_numerators per size: out = S - B with S = sum_col c_col * col (one batch)
    const bool fl_in_numerator = n_cls == 1 && logs[0] == L;
    void* d_sorted;
    SH_TRY(sh_stage(ctx, sorted.data(), sorted.size() * sizeof(void*), &d_sorted));
    for (uint32_t k = 0; k < n_cls; k++) {
      const uint32_t plog = logs[k] - b;
      SH_TRY(sh_dev_alloc(ctx, 16ull << plog, (void**)&partial[k]));
      const void* outp[4];
      for (int c = 0; c < 4; c++) outp[c] = partial[k] + ((uint64_t)c << plog);
      void* d_outs;
      SH_TRY(sh_stage(ctx, outp, sizeof(outp), &d_outs));
      OutSpec spec = {};
      spec.n_out = 1;
      spec.src[0] = 0;
      spec.w[0] = sh::qone();
      if (fl_in_numerator) {
        // the only size, at the combine's size: the numerator takes the first-linear term
        spec.fl_mask = 1;
        SH_TRY(sh_point_tab_make(ctx, sh::subgroup_gen(L + 1), L - b - 1, &spec.half));
      }
      const uint32_t rows = 1u << plog;
      SpanGuard span(ctx, "quotients_accumulate_numerators");
      k_accum_num<1><<<(rows + 255) / 256, 256, 0, ctx->stream>>>(
          (const uint32_t* const*)d_sorted + off[k], off[k + 1] - off[k], d_coef + 4ull * off[k], rows,
          (uint32_t* const*)d_outs, spec, d_B + 4 * k, d_FL + 4 * k);
      SH_TRY(sh_launch_check(ctx, "accumulate_numerators"));
    }
    // one point: every size's partial lifted and summed (lift_and_accumulate, ascending sizes)
    const uint32_t top = logs.back() - b;
    const void* parts[4];
    if (n_cls == 1) {
      for (int c = 0; c < 4; c++) parts[c] = partial[0] + ((uint64_t)c << top);
    } else {
      SH_TRY(sh_dev_alloc(ctx, 16ull << top, (void**)&lifted));
      std::vector<const uint32_t*> cols;
      std::vector<uint32_t> plogs;
      for (uint32_t k = 0; k < n_cls; k++) {
        for (int c = 0; c < 4; c++) cols.push_back(partial[k] + ((uint64_t)c << (logs[k] - b)));
        plogs.push_back(logs[k] - b);
      }
      uint32_t* o[4];
      for (int c = 0; c < 4; c++) o[c] = lifted + ((uint64_t)c << top);
      SH_TRY(sh_lift_and_accumulate(ctx, cols.data(), plogs.data(), n_cls, o));
      for (int c = 0; c < 4; c++) parts[c] = o[c];
    }
    k_dpoint_combine_const<<<1, 64, 0, ctx->stream>>>(d_point, d_FL, fl_in_numerator ? 0 : n_cls, top, d_kc);
    SH_TRY(sh_launch_check(ctx, "combine_const"));
    return combine_and_extend(ctx, tw, (const uint32_t* const*)parts, d_kc, 1, L, b, d_out, nullptr,
                              !fl_in_numerator);
  };
  rc = run();
  release();  // stream-ordered: later users of these buffers run after the kernels above
  return rc;
}
