/* CPU baseline for the trace commitment: a C restatement of the reference's SimdBackend commit
 * path (interpolate -> LDE -> lifted Blake2s Merkle tree), 16 M31 lanes per AVX-512 register,
 * POSIX threads over columns and over row chunks.
 *
 * TEST AND BENCHMARK INFRASTRUCTURE ONLY.  Nothing in the product library links or calls this
 * file; the benchmark times it beside the GPU commit, and the tests check that it produces the
 * same Merkle root as the scalar oracle (oracle.c).
 *
 * What it restates (all paths relative to the reference repository):
 *   - M31 lanes: crates/stwo/src/prover/backend/simd/m31.rs:147-297 (add/sub by min(x, x -+ P);
 *     multiply as 32x32 -> 64-bit products of even and odd lanes, then reduce)
 *   - interpolate: crates/stwo/src/prover/backend/simd/circle.rs:135-163 and
 *     simd/fft/ifft.rs:35-317 (circle layer, line layers ascending, then the 2^-n scale)
 *   - evaluate (LDE): simd/circle.rs:389-425 and simd/fft/rfft.rs:121-260 (coefficients read
 *     once and replicated into the 2^b subdomains, then line layers descending, circle layer last)
 *   - the layer arithmetic: butterfly / ibutterfly, crates/stwo/src/core/fft.rs:5-21; twiddles of
 *     the circle layer from the first line layer, [x, y] -> [y, -y, -x, x]
 *     (crates/stwo/src/prover/backend/cpu/circle.rs:306-323)
 *   - lifted leaves: simd/blake2s_lifted.rs:46-229 (16 rows per vector, columns ascending by size,
 *     the partial hash states of one size group lifted to the next by
 *     row -> ((row >> (d + 1)) << 1) + (row & 1))
 *   - inner layers: simd/blake2s_lifted.rs:232-263 (one compression per node)
 *   - Blake2s: RFC 7693 (simd/blake2s.rs:148-299 is the reference's 16-lane version)
 *
 * The FFT is cache-blocked rather than a copy of the reference's pass structure: the four
 * lowest layers run inside one register (lane permutations), the layers below 2^12 run inside
 * an L1-resident block, and the layers above run as radix-8 passes over the column.  Columns
 * must have at least 2^4 rows (the reference's small-size special cases are not restated here;
 * the commit workloads this times start at 2^12).
 */
#include <immintrin.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define P 0x7fffffffu
typedef __m512i v16;

/* ---------------------------------------------------------------- M31, 16 lanes */
static inline v16 vp(void) { return _mm512_set1_epi32((int)P); }
static inline v16 vadd(v16 a, v16 b) {
  v16 s = _mm512_add_epi32(a, b);
  return _mm512_min_epu32(s, _mm512_sub_epi32(s, vp()));
}
static inline v16 vsub(v16 a, v16 b) {
  v16 d = _mm512_sub_epi32(a, b);
  return _mm512_min_epu32(d, _mm512_add_epi32(d, vp()));
}
/* a*b mod P for canonical inputs: the 62-bit product x reduces to (x & P) + (x >> 31) < 2P */
static inline v16 vmul(v16 a, v16 b) {
  const v16 p64 = _mm512_set1_epi64((long long)P);
  v16 pe = _mm512_mul_epu32(a, b);
  v16 po = _mm512_mul_epu32(_mm512_srli_epi64(a, 32), _mm512_srli_epi64(b, 32));
  v16 re = _mm512_add_epi64(_mm512_and_si512(pe, p64), _mm512_srli_epi64(pe, 31));
  v16 ro = _mm512_add_epi64(_mm512_and_si512(po, p64), _mm512_srli_epi64(po, 31));
  v16 r = _mm512_mask_blend_epi32((__mmask16)0xAAAA, re, _mm512_slli_epi64(ro, 32));
  return _mm512_min_epu32(r, _mm512_sub_epi32(r, vp()));
}
static inline uint32_t smul(uint32_t a, uint32_t b) {
  uint64_t x = (uint64_t)a * b;
  uint32_t r = (uint32_t)((x & P) + (x >> 31));
  return r >= P ? r - P : r;
}
static inline uint32_t sneg(uint32_t a) { return a == 0 ? 0 : P - a; }
static uint32_t sinv(uint32_t a) {
  uint32_t r = 1, b = a;
  for (uint32_t e = P - 2; e; e >>= 1, b = smul(b, b))
    if (e & 1) r = smul(r, b);
  return r;
}

/* forward butterfly (fft.rs:5-12): v0 + v1*t, v0 - v1*t */
static inline void bf(v16 *a, v16 *b, v16 t) {
  v16 m = vmul(*b, t);
  *b = vsub(*a, m);
  *a = vadd(*a, m);
}
/* inverse butterfly (fft.rs:14-21): v0 + v1, (v0 - v1)*it */
static inline void ibf(v16 *a, v16 *b, v16 t) {
  v16 s = vadd(*a, *b);
  *b = vmul(vsub(*a, *b), t);
  *a = s;
}

/* ---------------------------------------------------------------- twiddles of one domain */
/* Line layer j (j = 0 largest) of a domain whose line coset has log size m, inside the twiddle
 * tree of the largest domain: the tree keeps smaller domains' layers at its tail
 * (crates/stwo/src/core/poly/utils.rs:98-111). */
typedef struct {
  const uint32_t *line[32]; /* line[j]: twiddles of butterfly span j + 1 */
  uint32_t *circle;         /* 2^(log - 1) twiddles of span 0 */
  uint32_t log;
} dom_tw;

static void dom_tw_init(dom_tw *d, const uint32_t *tree, uint64_t tree_len, uint32_t log) {
  const uint32_t m = log - 1;
  d->log = log;
  for (uint32_t j = 0; j < m; j++) d->line[j] = tree + tree_len - 2 * (1ull << (m - 1 - j));
  d->circle = (uint32_t *)aligned_alloc(64, 64 * (((1ull << (log - 1)) + 15) / 16 + 1));
  const uint32_t *l0 = d->line[0];
  for (uint64_t h = 0; h < (1ull << (log - 1)); h++) {
    uint32_t x = l0[(h >> 2) * 2], y = l0[(h >> 2) * 2 + 1];
    uint32_t v[4] = {y, sneg(y), sneg(x), x};
    d->circle[h] = v[h & 3];
  }
}

/* per-lane twiddle of span s (0..3) for the 16 values at index base */
static inline v16 lane_twiddles(const dom_tw *d, uint32_t s, uint64_t base) {
  const uint32_t *t = s == 0 ? d->circle : d->line[s - 1];
  const uint32_t cnt = 16u >> (s + 1);
  v16 raw = _mm512_maskz_loadu_epi32((__mmask16)((1u << cnt) - 1), t + (base >> (s + 1)));
  v16 idx = _mm512_srli_epi32(_mm512_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15),
                              (int)(s + 1));
  return _mm512_permutexvar_epi32(idx, raw);
}
/* butterflies of span s (< 16) between lanes of one register */
static inline v16 in_reg_layer(v16 x, v16 t, uint32_t s, int inverse) {
  const v16 lanes = _mm512_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
  v16 partner = _mm512_permutexvar_epi32(_mm512_xor_si512(lanes, _mm512_set1_epi32(1 << s)), x);
  __mmask16 hi = _mm512_test_epi32_mask(lanes, _mm512_set1_epi32(1 << s));
  if (inverse) { /* low lane: v0 + v1; high lane: (v0 - v1) * it */
    v16 sum = vadd(x, partner);
    v16 dif = vmul(vsub(partner, x), t);
    return _mm512_mask_blend_epi32(hi, sum, dif);
  }
  /* low lane: v0 + v1*t; high lane: v0 - v1*t */
  v16 m = vmul(_mm512_mask_blend_epi32(hi, partner, x), t);
  v16 base = _mm512_mask_blend_epi32(hi, x, partner);
  return _mm512_mask_blend_epi32(hi, vadd(base, m), vsub(base, m));
}

/* Radix-2^r pass (r = 1..3) over spans lo .. lo + r - 1 (lo >= 4) of x[start, start + len). */
static inline __attribute__((always_inline)) void pass_r(uint32_t *x, uint64_t start, uint64_t len,
                                                         uint32_t lo, const uint32_t r, const dom_tw *d,
                                                         const int inverse, uint32_t scale) {
  const uint64_t stride = 1ull << lo, group = stride << r;
  const v16 vs = _mm512_set1_epi32((int)scale);
  for (uint64_t g0 = start; g0 < start + len; g0 += group) {
    for (uint64_t l = 0; l < stride; l += 16) {
      v16 v[8];
      for (uint32_t k = 0; k < (1u << r); k++) v[k] = _mm512_load_si512(x + g0 + l + k * stride);
      for (uint32_t qq = 0; qq < r; qq++) {
        const uint32_t q = inverse ? qq : r - 1 - qq;
        const uint32_t *tw = d->line[lo + q - 1];
        for (uint32_t k = 0; k < (1u << r); k++) {
          if (k & (1u << q)) continue;
          v16 t = _mm512_set1_epi32((int)tw[(g0 >> (lo + q + 1)) + (k >> (q + 1))]);
          if (inverse)
            ibf(&v[k], &v[k + (1u << q)], t);
          else
            bf(&v[k], &v[k + (1u << q)], t);
        }
      }
      for (uint32_t k = 0; k < (1u << r); k++) {
        if (scale != 1) v[k] = vmul(v[k], vs);
        _mm512_store_si512(x + g0 + l + k * stride, v[k]);
      }
    }
  }
}

static void pass(uint32_t *x, uint64_t start, uint64_t len, uint32_t lo, uint32_t r, const dom_tw *d,
                 int inverse, uint32_t scale) {
  if (inverse) {
    if (r == 3) pass_r(x, start, len, lo, 3, d, 1, scale);
    else if (r == 2) pass_r(x, start, len, lo, 2, d, 1, scale);
    else pass_r(x, start, len, lo, 1, d, 1, scale);
  } else {
    if (r == 3) pass_r(x, start, len, lo, 3, d, 0, scale);
    else if (r == 2) pass_r(x, start, len, lo, 2, d, 0, scale);
    else pass_r(x, start, len, lo, 1, d, 0, scale);
  }
}

#ifndef BLOCK_LOG
#define BLOCK_LOG 12
#endif

/* interpolate (simd/circle.rs:135-163): bit-reversed evaluations src (2^n) -> coefficients dst */
static void interpolate(const uint32_t *src, uint32_t *dst, uint32_t n, const dom_tw *d) {
  const uint32_t K = n < BLOCK_LOG ? n : BLOCK_LOG;
  const uint32_t scale = sinv((uint32_t)(1u << n));
  for (uint64_t B = 0; B < (1ull << n); B += 1ull << K) {
    for (uint64_t v = B; v < B + (1ull << K); v += 16) {
      v16 x = _mm512_load_si512(src + v);
      for (uint32_t s = 0; s < 4; s++) x = in_reg_layer(x, lane_twiddles(d, s, v), s, 1);
      _mm512_store_si512(dst + v, x);
    }
    for (uint32_t lo = 4; lo < K;) {
      uint32_t r = K - lo < 3 ? K - lo : 3;
      pass(dst, B, 1ull << K, lo, r, d, 1, (lo + r == n) ? scale : 1);
      lo += r;
    }
    if (K == 4 && n == 4)
      for (uint64_t v = B; v < B + 16; v += 16)
        _mm512_store_si512(dst + v, vmul(_mm512_load_si512(dst + v), _mm512_set1_epi32((int)scale)));
  }
  for (uint32_t lo = K; lo < n;) {
    uint32_t r = n - lo < 3 ? n - lo : 3;
    pass(dst, 0, 1ull << n, lo, r, d, 1, (lo + r == n) ? scale : 1);
    lo += r;
  }
}

/* evaluate (simd/circle.rs:389-425): coefficients (2^n) -> bit-reversed evaluations on the
 * domain of size 2^(n+b).  The top b layers see a zero upper half, so they copy. */
static void evaluate(const uint32_t *coef, uint32_t *dst, uint32_t n, uint32_t b, const dom_tw *d) {
  const uint64_t N = 1ull << (n + b);
  for (uint64_t j = 0; j < N; j += 1ull << n) memcpy(dst + j, coef, 4ull << n);
  const uint32_t K = n < BLOCK_LOG ? n : BLOCK_LOG;
  for (uint32_t hi = n; hi > K;) {
    uint32_t r = hi - K < 3 ? hi - K : 3;
    pass(dst, 0, N, hi - r, r, d, 0, 1);
    hi -= r;
  }
  for (uint64_t B = 0; B < N; B += 1ull << K) {
    for (uint32_t hi = K; hi > 4;) {
      uint32_t r = hi - 4 < 3 ? hi - 4 : 3;
      pass(dst, B, 1ull << K, hi - r, r, d, 0, 1);
      hi -= r;
    }
    for (uint64_t v = B; v < B + (1ull << K); v += 16) {
      v16 x = _mm512_load_si512(dst + v);
      for (int s = 3; s >= 0; s--) x = in_reg_layer(x, lane_twiddles(d, (uint32_t)s, v), (uint32_t)s, 0);
      _mm512_store_si512(dst + v, x);
    }
  }
}

/* ---------------------------------------------------------------- Blake2s, 16 lanes */
static const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                               0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t SIGMA[10][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15},
    {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4},
    {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13},
    {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11},
    {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5},
    {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0}};

#define G(a, b, c, d, x, y)                                                 \
  do {                                                                      \
    v[a] = _mm512_add_epi32(_mm512_add_epi32(v[a], v[b]), (x));             \
    v[d] = _mm512_ror_epi32(_mm512_xor_si512(v[d], v[a]), 16);              \
    v[c] = _mm512_add_epi32(v[c], v[d]);                                    \
    v[b] = _mm512_ror_epi32(_mm512_xor_si512(v[b], v[c]), 12);              \
    v[a] = _mm512_add_epi32(_mm512_add_epi32(v[a], v[b]), (y));             \
    v[d] = _mm512_ror_epi32(_mm512_xor_si512(v[d], v[a]), 8);               \
    v[c] = _mm512_add_epi32(v[c], v[d]);                                    \
    v[b] = _mm512_ror_epi32(_mm512_xor_si512(v[b], v[c]), 7);               \
  } while (0)

/* one compression of 16 independent states; t = bytes absorbed including this block */
static void compress16(v16 h[8], const v16 m[16], uint32_t t, int last) {
  v16 v[16];
  for (int i = 0; i < 8; i++) {
    v[i] = h[i];
    v[i + 8] = _mm512_set1_epi32((int)IV[i]);
  }
  v[12] = _mm512_set1_epi32((int)(IV[4] ^ t));
  if (last) v[14] = _mm512_set1_epi32((int)~IV[6]);
  for (int r = 0; r < 10; r++) {
    const uint8_t *s = SIGMA[r];
    G(0, 4, 8, 12, m[s[0]], m[s[1]]);
    G(1, 5, 9, 13, m[s[2]], m[s[3]]);
    G(2, 6, 10, 14, m[s[4]], m[s[5]]);
    G(3, 7, 11, 15, m[s[6]], m[s[7]]);
    G(0, 5, 10, 15, m[s[8]], m[s[9]]);
    G(1, 6, 11, 12, m[s[10]], m[s[11]]);
    G(2, 7, 8, 13, m[s[12]], m[s[13]]);
    G(3, 4, 9, 14, m[s[14]], m[s[15]]);
  }
  for (int i = 0; i < 8; i++) h[i] = _mm512_xor_si512(h[i], _mm512_xor_si512(v[i], v[i + 8]));
}
#undef G

/* ---------------------------------------------------------------- parallel for */
typedef void (*range_fn)(void *arg, uint64_t begin, uint64_t end);
typedef struct {
  range_fn fn;
  void *arg;
  uint64_t n, grain;
  atomic_ullong next;
} pfor;
static void *pfor_worker(void *p) {
  pfor *f = (pfor *)p;
  for (;;) {
    uint64_t i = atomic_fetch_add(&f->next, f->grain);
    if (i >= f->n) break;
    f->fn(f->arg, i, i + f->grain < f->n ? i + f->grain : f->n);
  }
  return NULL;
}
static void parallel_for(int nthreads, uint64_t n, uint64_t grain, range_fn fn, void *arg) {
  pfor f = {fn, arg, n, grain ? grain : 1, 0};
  if (nthreads <= 1 || n <= f.grain) {
    fn(arg, 0, n);
    return;
  }
  pthread_t th[256];
  int extra = nthreads - 1 < 255 ? nthreads - 1 : 255;
  int started = 0;
  for (int i = 0; i < extra; i++)
    if (pthread_create(&th[started], NULL, pfor_worker, &f) == 0) started++;
  pfor_worker(&f);
  for (int i = 0; i < started; i++) pthread_join(th[i], NULL);
}

/* ---------------------------------------------------------------- the commit */
typedef struct {
  const uint32_t *const *cols;
  const uint32_t *logs;
  uint32_t **coef, **lde;
  const uint32_t *order; /* columns, largest first, for load balance */
  uint32_t b;
  dom_tw *itw_by_log, *tw_by_log;
} fft_job;
static void fft_range(void *arg, uint64_t begin, uint64_t end) {
  fft_job *j = (fft_job *)arg;
  for (uint64_t k = begin; k < end; k++) {
    const uint32_t c = j->order[k], n = j->logs[c];
    interpolate(j->cols[c], j->coef[c], n, &j->itw_by_log[n]);
    evaluate(j->coef[c], j->lde[c], n, j->b, &j->tw_by_log[n + j->b]);
  }
}

/* state of 16 rows: 8 chaining words + 16 pending message words, lane = row */
enum { STATE_WORDS = 24 };
typedef struct {
  const uint32_t *const *cols; /* this group's LDE columns */
  uint32_t ncols, log, prev_log;
  const uint32_t *prev; /* previous group's states (NULL: first group) */
  uint32_t *next;       /* this group's states (NULL: finalize into leaves) */
  uint32_t *leaves;     /* 8 words per row */
  uint32_t npend0, count0;
} leaf_job;

static void load_states(const leaf_job *j, uint64_t r0, v16 h[8], v16 pend[16]) {
  if (!j->prev) {
    for (int i = 0; i < 8; i++) h[i] = _mm512_set1_epi32((int)(IV[i] ^ (i == 0 ? 0x01010020u : 0)));
    for (int i = 0; i < 16; i++) pend[i] = _mm512_setzero_si512();
    return;
  }
  const uint32_t d = j->log - j->prev_log;
  v16 rows = _mm512_add_epi32(_mm512_set1_epi32((int)r0),
                              _mm512_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15));
  /* row -> ((row >> (d + 1)) << 1) + (row & 1) at the previous size */
  v16 src = _mm512_or_si512(_mm512_slli_epi32(_mm512_srli_epi32(rows, (int)(d + 1)), 1),
                            _mm512_and_si512(rows, _mm512_set1_epi32(1)));
  v16 base = _mm512_mullo_epi32(_mm512_srli_epi32(src, 4), _mm512_set1_epi32(STATE_WORDS * 16));
  v16 lane = _mm512_and_si512(src, _mm512_set1_epi32(15));
  v16 idx = _mm512_add_epi32(base, lane);
  for (int w = 0; w < STATE_WORDS; w++) {
    v16 x = _mm512_i32gather_epi32(_mm512_add_epi32(idx, _mm512_set1_epi32(16 * w)), (const void *)j->prev, 4);
    if (w < 8)
      h[w] = x;
    else
      pend[w - 8] = x;
  }
}

static void leaves_range(void *arg, uint64_t begin, uint64_t end) {
  const leaf_job *j = (const leaf_job *)arg;
  for (uint64_t chunk = begin; chunk < end; chunk++) {
    const uint64_t r0 = chunk * 16;
    v16 h[8], pend[16];
    load_states(j, r0, h, pend);
    uint32_t npend = j->npend0, count = j->count0;
    for (uint32_t c = 0; c < j->ncols; c++) {
      if (npend == 16) {
        compress16(h, pend, 4 * count, 0);
        npend = 0;
      }
      pend[npend++] = _mm512_loadu_si512(j->cols[c] + r0);
      count++;
    }
    if (j->next) {
      uint32_t *st = j->next + chunk * STATE_WORDS * 16;
      for (int w = 0; w < 8; w++) _mm512_store_si512(st + 16 * w, h[w]);
      for (int w = 0; w < 16; w++) _mm512_store_si512(st + 16 * (8 + w), pend[w]);
    } else {
      for (uint32_t w = npend; w < 16; w++) pend[w] = _mm512_setzero_si512();
      compress16(h, pend, 4 * count, 1);
      v16 out_idx = _mm512_mullo_epi32(_mm512_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15),
                                       _mm512_set1_epi32(8));
      for (int w = 0; w < 8; w++)
        _mm512_i32scatter_epi32(j->leaves + r0 * 8, _mm512_add_epi32(out_idx, _mm512_set1_epi32(w)), h[w], 4);
    }
  }
}

typedef struct {
  const uint32_t *prev; /* 2^(k+1) nodes of 8 words */
  uint32_t *out;        /* 2^k nodes */
} layer_job;
static void layer_range(void *arg, uint64_t begin, uint64_t end) {
  const layer_job *j = (const layer_job *)arg;
  const v16 idx16 = _mm512_mullo_epi32(_mm512_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15),
                                       _mm512_set1_epi32(16));
  const v16 idx8 = _mm512_srli_epi32(idx16, 1);
  for (uint64_t chunk = begin; chunk < end; chunk++) {
    const uint64_t i0 = chunk * 16;
    v16 m[16], h[8];
    for (int w = 0; w < 16; w++)
      m[w] = _mm512_i32gather_epi32(_mm512_add_epi32(idx16, _mm512_set1_epi32(w)), (const void *)(j->prev + i0 * 16), 4);
    for (int i = 0; i < 8; i++) h[i] = _mm512_set1_epi32((int)(IV[i] ^ (i == 0 ? 0x01010020u : 0)));
    compress16(h, m, 64, 1);
    for (int w = 0; w < 8; w++)
      _mm512_i32scatter_epi32(j->out + i0 * 8, _mm512_add_epi32(idx8, _mm512_set1_epi32(w)), h[w], 4);
  }
}
/* nodes of a layer below 16 nodes: the same compression with lanes 0..n-1 used */
static void small_layer(const uint32_t *prev, uint32_t *out, uint64_t n) {
  uint32_t buf[16 * 16] __attribute__((aligned(64)));
  memset(buf, 0, sizeof buf);
  memcpy(buf, prev, 64 * n);
  uint32_t tmp[16 * 8] __attribute__((aligned(64)));
  layer_job j = {buf, tmp};
  layer_range(&j, 0, 1);
  memcpy(out, tmp, 32 * n);
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int sp_cpu_has_avx512(void) {
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx512f") != 0;
}

/* The whole commit of ncols columns (bit-reversed evaluations on CanonicCoset(logs[c]), logs
 * ascending and each >= 4): interpolate, LDE with blowup 2^b, lifted tree of height L (>= every
 * logs[c] + b).  tw / itw: the twiddle tree of CanonicCoset(L) and its inverses (tw_len words).
 * Buffers are allocated and touched before the clock starts; reps commits are timed and the
 * fastest is reported in times[0..2] (seconds: FFT, leaves, inner layers).  root: 32 bytes. */
int sp_commit_lifted(const uint32_t *const *cols, const uint32_t *logs, uint32_t ncols, uint32_t b,
                     uint32_t L, const uint32_t *tw, const uint32_t *itw, uint64_t tw_len, int nthreads,
                     int reps, uint8_t *root, double *times) {
  if (ncols == 0 || L < 5 || L > 28) return -1;
  for (uint32_t c = 0; c < ncols; c++)
    if (logs[c] < 4 || logs[c] + b > L || (c && logs[c] < logs[c - 1])) return -1;
  for (uint32_t c = 0; c < ncols; c++)
    if (((uintptr_t)cols[c]) & 63) return -2;
  dom_tw itwd[32], twd[32];
  memset(itwd, 0, sizeof itwd);
  memset(twd, 0, sizeof twd);
  for (uint32_t c = 0; c < ncols; c++) {
    const uint32_t n = logs[c];
    if (!itwd[n].circle) dom_tw_init(&itwd[n], itw, tw_len, n);
    if (!twd[n + b].circle) dom_tw_init(&twd[n + b], tw, tw_len, n + b);
  }
  uint32_t **coef = (uint32_t **)calloc(ncols, sizeof(uint32_t *));
  uint32_t **lde = (uint32_t **)calloc(ncols, sizeof(uint32_t *));
  uint32_t *order = (uint32_t *)malloc(4ull * ncols);
  for (uint32_t c = 0; c < ncols; c++) {
    coef[c] = (uint32_t *)aligned_alloc(64, 4ull << logs[c]);
    lde[c] = (uint32_t *)aligned_alloc(64, 4ull << (logs[c] + b));
    memset(coef[c], 0, 4ull << logs[c]);
    memset(lde[c], 0, 4ull << (logs[c] + b));
    order[c] = ncols - 1 - c;
  }
  /* size groups and their state buffers */
  uint32_t ngroups = 0, gstart[64], glog[64];
  for (uint32_t c = 0; c < ncols; c++)
    if (c == 0 || logs[c] != logs[c - 1]) {
      if (ngroups == 64) return -1;
      gstart[ngroups] = c;
      glog[ngroups++] = logs[c] + b;
    }
  gstart[ngroups] = ncols;
  const int lift_at_end = glog[ngroups - 1] < L;
  uint32_t *states[65];
  for (uint32_t g = 0; g < ngroups; g++) {
    states[g] = NULL;
    if (g + 1 < ngroups || lift_at_end) {
      states[g] = (uint32_t *)aligned_alloc(64, 4ull * STATE_WORDS << glog[g]);
      memset(states[g], 0, 4ull * STATE_WORDS << glog[g]);
    }
  }
  uint32_t **layers = (uint32_t **)calloc(L + 1, sizeof(uint32_t *));
  for (uint32_t k = 0; k <= L; k++) {
    layers[k] = (uint32_t *)aligned_alloc(64, k ? 32ull << k : 64);
    memset(layers[k], 0, 32ull << k);
  }
  const uint32_t **gcols = (const uint32_t **)malloc(sizeof(uint32_t *) * ncols);
  for (uint32_t c = 0; c < ncols; c++) gcols[c] = lde[c];

  double best[3] = {1e30, 1e30, 1e30}, best_total = 1e30;
  for (int rep = 0; rep < (reps > 0 ? reps : 1); rep++) {
    const double t0 = now_s();
    fft_job fj = {cols, logs, coef, lde, order, b, itwd, twd};
    parallel_for(nthreads, ncols, 1, fft_range, &fj);
    const double t1 = now_s();
    uint32_t count = 0, npend = 0;
    for (uint32_t g = 0; g < ngroups; g++) {
      const uint32_t k = gstart[g + 1] - gstart[g];
      leaf_job lj = {gcols + gstart[g], k, glog[g], g ? glog[g - 1] : 1, g ? states[g - 1] : NULL,
                     states[g], layers[L], npend, count};
      parallel_for(nthreads, (1ull << glog[g]) / 16, 64, leaves_range, &lj);
      /* the uniform pending count after the group's columns */
      for (uint32_t c = 0; c < k; c++) {
        if (npend == 16) npend = 0;
        npend++;
        count++;
      }
    }
    if (lift_at_end) { /* lift the last states to height L and finalize (no further columns) */
      leaf_job lj = {NULL, 0, L, glog[ngroups - 1], states[ngroups - 1], NULL, layers[L], npend, count};
      parallel_for(nthreads, (1ull << L) / 16, 64, leaves_range, &lj);
    }
    const double t2 = now_s();
    for (int k = (int)L - 1; k >= 0; k--) {
      if (k >= 4) {
        layer_job lj = {layers[k + 1], layers[k]};
        parallel_for(nthreads, (1ull << k) / 16, 64, layer_range, &lj);
      } else {
        small_layer(layers[k + 1], layers[k], 1ull << k);
      }
    }
    const double t3 = now_s();
    if (t3 - t0 < best_total) {
      best_total = t3 - t0;
      best[0] = t1 - t0;
      best[1] = t2 - t1;
      best[2] = t3 - t2;
    }
  }
  memcpy(root, layers[0], 32);
  for (int i = 0; i < 3; i++) times[i] = best[i];

  for (uint32_t k = 0; k <= L; k++) free(layers[k]);
  free(layers);
  for (uint32_t g = 0; g < ngroups; g++) free(states[g]);
  for (uint32_t c = 0; c < ncols; c++) {
    free(coef[c]);
    free(lde[c]);
  }
  for (uint32_t n = 0; n < 32; n++) {
    free(itwd[n].circle);
    free(twd[n].circle);
  }
  free(coef);
  free(lde);
  free(order);
  free(gcols);
  return 0;
}
