/*
 * oracle.c — CPU restatement of the Stwo CpuBackend algorithms for the
 * backend hot path.  TEST INFRASTRUCTURE ONLY: the checker the HIP backend
 * is compared against (see oracle.h).  Each function cites the reference
 * file:line it restates; the code follows the reference's loop structure so
 * that a reader can check it line by line, and is deliberately scalar and
 * unoptimised.
 *
 * Every function here is pinned by tests/test_oracle_*.py against the
 * reference's own known-answer vectors (tests/golden/) or against closed
 * forms and invariants that the mathematics fixes.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#define P OR_P

/* ======================================================================
 * M31  (crates/stwo/src/core/fields/m31.rs)
 * ====================================================================== */

/* m31.rs:46-48 partial_reduce: val % P for val in [0, 2P) */
static uint32_t partial_reduce(uint32_t v) { return v >= P ? v - P : v; }
/* m31.rs:58-60 reduce: val % P for val in [0, P^2) */
static uint32_t reduce64(uint64_t v) {
  return (uint32_t)(((((v >> 31) + v + 1) >> 31) + v) & (uint64_t)P);
}
static uint32_t m_add(uint32_t a, uint32_t b) { return partial_reduce(a + b); } /* m31.rs:81-87 */
static uint32_t m_neg(uint32_t a) { return partial_reduce(P - a); }             /* m31.rs:89-95 */
static uint32_t m_sub(uint32_t a, uint32_t b) { return partial_reduce(a + P - b); } /* :97-103 */
static uint32_t m_mul(uint32_t a, uint32_t b) { return reduce64((uint64_t)a * b); } /* :105-111 */
static uint32_t m_pow(uint32_t a, uint64_t e) {
  uint32_t r = 1;
  while (e) {
    if (e & 1) r = m_mul(r, a);
    a = m_mul(a, a);
    e >>= 1;
  }
  return r;
}
/* m31.rs:66-69: inverse = a^(P-2) */
static uint32_t m_inv(uint32_t a) {
  if (a == 0) abort(); /* "0 has no inverse" */
  return m_pow(a, P - 2);
}
uint32_t or_m31_mul(uint32_t a, uint32_t b) { return m_mul(a, b); }
uint32_t or_m31_add(uint32_t a, uint32_t b) { return m_add(a, b); }
uint32_t or_m31_reduce(uint64_t v) { return reduce64(v); }
uint32_t or_m31_inv(uint32_t a) { return m_inv(a); }

/* ======================================================================
 * CM31 (crates/stwo/src/core/fields/cm31.rs) : a + b i, i^2 = -1
 * ====================================================================== */
typedef struct { uint32_t a, b; } cm31;
static cm31 c_add(cm31 x, cm31 y) { return (cm31){m_add(x.a, y.a), m_add(x.b, y.b)}; }
static cm31 c_sub(cm31 x, cm31 y) { return (cm31){m_sub(x.a, y.a), m_sub(x.b, y.b)}; }
static cm31 c_neg(cm31 x) { return (cm31){m_neg(x.a), m_neg(x.b)}; }
/* cm31.rs:44-50 */
static cm31 c_mul(cm31 x, cm31 y) {
  return (cm31){m_sub(m_mul(x.a, y.a), m_mul(x.b, y.b)), m_add(m_mul(x.a, y.b), m_mul(x.b, y.a))};
}
static cm31 c_mul_m(cm31 x, uint32_t m) { return (cm31){m_mul(x.a, m), m_mul(x.b, m)}; }
/* cm31.rs:63-70: 1/(a+bi) = (a-bi)/(a^2+b^2) */
static cm31 c_inv(cm31 x) {
  uint32_t n = m_add(m_mul(x.a, x.a), m_mul(x.b, x.b));
  return c_mul_m((cm31){x.a, m_neg(x.b)}, m_inv(n));
}
static int c_is_zero(cm31 x) { return x.a == 0 && x.b == 0; }

/* ======================================================================
 * QM31 (crates/stwo/src/core/fields/qm31.rs) : a + b u, u^2 = 2 + i
 * ====================================================================== */
typedef struct { cm31 a, b; } qm31;
static const cm31 QM31_R = {2, 1}; /* qm31.rs:14 */
static qm31 q_from4(const uint32_t w[4]) { return (qm31){{w[0], w[1]}, {w[2], w[3]}}; }
static void q_to4(qm31 x, uint32_t w[4]) { w[0] = x.a.a; w[1] = x.a.b; w[2] = x.b.a; w[3] = x.b.b; }
static qm31 q_m(uint32_t m) { return (qm31){{m, 0}, {0, 0}}; }
static qm31 q_zero(void) { return q_m(0); }
static qm31 q_one(void) { return q_m(1); }
static qm31 q_add(qm31 x, qm31 y) { return (qm31){c_add(x.a, y.a), c_add(x.b, y.b)}; }
static qm31 q_sub(qm31 x, qm31 y) { return (qm31){c_sub(x.a, y.a), c_sub(x.b, y.b)}; }
static qm31 q_neg(qm31 x) { return (qm31){c_neg(x.a), c_neg(x.b)}; }
/* qm31.rs:75-82: (a + bu)(c + du) = (ac + R bd) + (ad + bc) u */
static qm31 q_mul(qm31 x, qm31 y) {
  return (qm31){c_add(c_mul(x.a, y.a), c_mul(QM31_R, c_mul(x.b, y.b))),
                c_add(c_mul(x.a, y.b), c_mul(x.b, y.a))};
}
static qm31 q_mul_m(qm31 x, uint32_t m) { return (qm31){c_mul_m(x.a, m), c_mul_m(x.b, m)}; }
/* qm31.rs:55-57 mul_cm31 */
static qm31 q_mul_c(qm31 x, cm31 c) { return (qm31){c_mul(x.a, c), c_mul(x.b, c)}; }
static int q_is_zero(qm31 x) { return c_is_zero(x.a) && c_is_zero(x.b); }
static int q_eq(qm31 x, qm31 y) {
  return x.a.a == y.a.a && x.a.b == y.a.b && x.b.a == y.b.a && x.b.b == y.b.b;
}
/* qm31.rs:108-117: (a + bu)^-1 = (a - bu) / (a^2 - (2+i) b^2) */
static qm31 q_inv(qm31 x) {
  if (q_is_zero(x)) abort();
  cm31 b2 = c_mul(x.b, x.b);
  cm31 ib2 = {m_neg(b2.b), b2.a};
  cm31 denom = c_sub(c_mul(x.a, x.a), c_add(c_add(b2, b2), ib2));
  cm31 di = c_inv(denom);
  return (qm31){c_mul(x.a, di), c_neg(c_mul(x.b, di))};
}
/* fields/mod.rs:410-412 complex_conjugate: (a, b) -> (a, -b) */
static qm31 q_conj(qm31 x) { return (qm31){x.a, c_neg(x.b)}; }
void or_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]) {
  q_to4(q_mul(q_from4(a), q_from4(b)), out);
}
void or_qm31_inv(const uint32_t a[4], uint32_t out[4]) { q_to4(q_inv(q_from4(a)), out); }

/* ======================================================================
 * Circle group (crates/stwo/src/core/circle.rs)
 * ====================================================================== */
typedef struct { uint32_t x, y; } mpt;
typedef struct { qm31 x, y; } qpt;
static const mpt M31_CIRCLE_GEN = {2, 1268011823}; /* circle.rs:181-182 */
static mpt mp_zero(void) { return (mpt){1, 0}; }
/* circle.rs:111-117 */
static mpt mp_add(mpt p, mpt q) {
  return (mpt){m_sub(m_mul(p.x, q.x), m_mul(p.y, q.y)), m_add(m_mul(p.x, q.y), m_mul(p.y, q.x))};
}
/* circle.rs:67-78 */
static mpt mp_mul(mpt p, uint64_t scalar) {
  mpt res = mp_zero(), cur = p;
  while (scalar > 0) {
    if (scalar & 1) res = mp_add(res, cur);
    cur = mp_add(cur, cur);
    scalar >>= 1;
  }
  return res;
}
#define LOG_ORDER 31u
static uint32_t idx_reduce(uint64_t i) { return (uint32_t)(i & ((1ull << LOG_ORDER) - 1)); }
/* circle.rs:229-231 CirclePointIndex::to_point */
static mpt index_to_point(uint32_t idx) { return mp_mul(M31_CIRCLE_GEN, idx); }
void or_index_to_point(uint32_t index, uint32_t out[2]) {
  mpt p = index_to_point(idx_reduce(index));
  out[0] = p.x;
  out[1] = p.y;
}
static uint32_t subgroup_gen(uint32_t log_size) { return 1u << (LOG_ORDER - log_size); } /* :214-217 */
static uint32_t idx_neg(uint32_t i) { return idx_reduce((1ull << LOG_ORDER) - i); }    /* :256-262 */

/* Coset (circle.rs:264-376) */
typedef struct { uint32_t initial_index, step_size, log_size; } coset;
static coset coset_new(uint32_t initial_index, uint32_t log_size) {
  return (coset){idx_reduce(initial_index), subgroup_gen(log_size), log_size};
}
static uint32_t coset_index_at(coset c, uint64_t i) {
  return idx_reduce((uint64_t)c.initial_index + (uint64_t)c.step_size * i);
}
static mpt coset_at(coset c, uint64_t i) { return index_to_point(coset_index_at(c, i)); }
/* circle.rs:328-337 */
static coset coset_double(coset c) {
  return (coset){idx_reduce((uint64_t)c.initial_index * 2), idx_reduce((uint64_t)c.step_size * 2),
                 c.log_size ? c.log_size - 1 : 0};
}
/* poly/circle/domain.rs:50-60 CircleDomain::index_at */
static uint32_t circle_domain_index_at(coset half, uint64_t i) {
  uint64_t hs = 1ull << half.log_size;
  if (i < hs) return coset_index_at(half, i);
  return idx_neg(coset_index_at(half, i - hs));
}
static mpt circle_domain_at(coset half, uint64_t i) {
  return index_to_point(circle_domain_index_at(half, i));
}
void or_circle_domain_at(uint32_t half_initial_index, uint32_t half_log_size, uint64_t i,
                         uint32_t out[2]) {
  mpt p = circle_domain_at(coset_new(half_initial_index, half_log_size), i);
  out[0] = p.x;
  out[1] = p.y;
}
/* crates/stwo/src/core/utils.rs:106-111 */
static uint64_t bit_reverse_index(uint64_t i, uint32_t log_size) {
  if (log_size == 0) return i;
  uint64_t r = 0;
  for (uint32_t b = 0; b < log_size; b++) r |= ((i >> b) & 1ull) << (log_size - 1 - b);
  return r;
}
/* crates/stwo/src/core/utils.rs:118-128 */
static void bit_reverse_u32(uint32_t *v, uint64_t n, uint32_t log_n) {
  for (uint64_t i = 0; i < n; i++) {
    uint64_t j = bit_reverse_index(i, log_n);
    if (j > i) {
      uint32_t t = v[i];
      v[i] = v[j];
      v[j] = t;
    }
  }
}

/* QM31 circle points */
static qpt qp_from8(const uint32_t w[8]) { return (qpt){q_from4(w), q_from4(w + 4)}; }
static qpt qp_from_m(mpt p) { return (qpt){q_m(p.x), q_m(p.y)}; }
static qpt qp_add(qpt p, qpt q) {
  return (qpt){q_sub(q_mul(p.x, q.x), q_mul(p.y, q.y)), q_add(q_mul(p.x, q.y), q_mul(p.y, q.x))};
}
static qpt qp_conj(qpt p) { return (qpt){p.x, q_neg(p.y)}; }
static qpt qp_sub(qpt p, qpt q) { return qp_add(p, qp_conj(q)); } /* circle.rs:131-139 */
/* circle.rs:37-40 double_x = 2x^2 - 1 */
static qm31 q_double_x(qm31 x) {
  qm31 sx = q_mul(x, x);
  return q_sub(q_add(sx, sx), q_one());
}

/* ======================================================================
 * Twiddle tree (crates/stwo/src/prover/backend/cpu/circle.rs:240-285)
 * ====================================================================== */
/* cpu/circle.rs:274-285 slow_precompute_twiddles */
static void slow_precompute_twiddles(coset c, uint32_t *tw) {
  uint64_t len = 0;
  uint32_t levels = c.log_size;
  for (uint32_t l = 0; l < levels; l++) {
    uint64_t i0 = len;
    uint64_t half = (1ull << c.log_size) / 2;
    for (uint64_t i = 0; i < half; i++) tw[len++] = coset_at(c, i).x;
    bit_reverse_u32(tw + i0, half, c.log_size - 1);
    c = coset_double(c);
  }
  tw[len++] = 1; /* padding to a power of two */
}
void or_precompute_twiddles(uint32_t initial_index, uint32_t log_size, uint32_t *tw,
                            uint32_t *itw) {
  coset c = coset_new(initial_index, log_size);
  slow_precompute_twiddles(c, tw);
  uint64_t n = 1ull << log_size;
  for (uint64_t i = 0; i < n; i++) itw[i] = m_inv(tw[i]); /* cpu/circle.rs:240-262 */
}
/* crates/stwo/src/prover/poly/twiddles.rs:40-76 */
void or_extract_subdomain_twiddles(const uint32_t *buf, uint64_t buf_len, uint32_t domain_log,
                                   uint32_t sub_log, uint32_t *out) {
  uint32_t dh = domain_log - 1, sh = sub_log - 1;
  uint32_t bh = 0;
  while ((1ull << (bh + 1)) <= buf_len) bh++;
  uint32_t skip = bh - dh;
  uint64_t buf_size = 1ull << bh;
  uint64_t k = 0;
  for (uint32_t layer = 0; layer < sh; layer++) {
    uint32_t root_layer = skip + layer;
    uint64_t start = buf_size - (buf_size >> root_layer);
    uint64_t sz = 1ull << (sh - 1 - layer);
    for (uint64_t j = 0; j < sz; j++) out[k++] = buf[start + j];
  }
  out[k++] = buf[buf_len - 1];
}

/* ======================================================================
 * Circle FFT (crates/stwo/src/prover/backend/cpu/circle.rs)
 * ====================================================================== */
/* crates/stwo/src/core/fft.rs:5-12 */
static void butterfly(uint32_t *v0, uint32_t *v1, uint32_t t) {
  uint32_t tmp = m_mul(*v1, t);
  *v1 = m_sub(*v0, tmp);
  *v0 = m_add(*v0, tmp);
}
/* crates/stwo/src/core/fft.rs:14-21 */
static void ibutterfly(uint32_t *v0, uint32_t *v1, uint32_t it) {
  uint32_t tmp = *v0;
  *v0 = m_add(tmp, *v1);
  *v1 = m_mul(m_sub(tmp, *v1), it);
}
/* QM31 variant of fft.rs:14-21 (used by the folds) */
static void q_ibutterfly(qm31 *v0, qm31 *v1, uint32_t it) {
  qm31 tmp = *v0;
  *v0 = q_add(tmp, *v1);
  *v1 = q_mul_m(q_sub(tmp, *v1), it);
}
/* cpu/circle.rs:287-301 fft_layer_loop */
static void fft_layer_loop(uint32_t *values, uint32_t i, uint64_t h, uint32_t t, int inverse) {
  for (uint64_t l = 0; l < (1ull << i); l++) {
    uint64_t idx0 = (h << (i + 1)) + l;
    uint64_t idx1 = idx0 + (1ull << i);
    if (inverse)
      ibutterfly(&values[idx0], &values[idx1], t);
    else
      butterfly(&values[idx0], &values[idx1], t);
  }
}
/* crates/stwo/src/core/poly/utils.rs:98-111: line layer j (j = 0 largest) of a domain whose
 * line coset has log size m; returns pointer into buf, layer length 2^(m-1-j). */
static const uint32_t *line_layer(const uint32_t *buf, uint64_t buf_len, uint32_t m, uint32_t j) {
  uint64_t len = 1ull << (m - 1 - j);
  return buf + buf_len - 2 * len;
}
/* cpu/circle.rs:306-323 circle_twiddles_from_line_twiddles: [x, y] -> [y, -y, -x, x] */
static uint32_t circle_twiddle(const uint32_t *l0, uint64_t h) {
  uint32_t x = l0[(h >> 2) * 2], y = l0[(h >> 2) * 2 + 1];
  switch (h & 3) {
    case 0: return y;
    case 1: return m_neg(y);
    case 2: return m_neg(x);
    default: return x;
  }
}

/* cpu/circle.rs:26-79 PolyOps::interpolate */
void or_interpolate(const uint32_t *itw, uint64_t tw_len, uint32_t half_initial,
                    uint32_t half_log, uint32_t *values) {
  coset half = coset_new(half_initial, half_log);
  uint32_t log_size = half_log + 1;
  if (log_size == 1) { /* :33-42 */
    uint32_t y = index_to_point(half.initial_index).y;
    uint32_t n = 2;
    uint32_t yn_inv = m_inv(m_mul(y, n));
    uint32_t y_inv = m_mul(yn_inv, n);
    uint32_t n_inv = m_mul(yn_inv, y);
    uint32_t v0 = values[0], v1 = values[1];
    ibutterfly(&v0, &v1, y_inv);
    values[0] = m_mul(v0, n_inv);
    values[1] = m_mul(v1, n_inv);
    return;
  }
  if (log_size == 2) { /* :44-58 */
    mpt p = index_to_point(half.initial_index);
    uint32_t x = p.x, y = p.y, n = 4;
    uint32_t xyn_inv = m_inv(m_mul(m_mul(x, y), n));
    uint32_t x_inv = m_mul(m_mul(xyn_inv, y), n);
    uint32_t y_inv = m_mul(m_mul(xyn_inv, x), n);
    uint32_t n_inv = m_mul(m_mul(xyn_inv, x), y);
    uint32_t v0 = values[0], v1 = values[1], v2 = values[2], v3 = values[3];
    ibutterfly(&v0, &v1, y_inv);
    ibutterfly(&v2, &v3, m_neg(y_inv));
    ibutterfly(&v0, &v2, x_inv);
    ibutterfly(&v1, &v3, x_inv);
    values[0] = m_mul(v0, n_inv);
    values[1] = m_mul(v1, n_inv);
    values[2] = m_mul(v2, n_inv);
    values[3] = m_mul(v3, n_inv);
    return;
  }
  uint32_t m = half_log; /* line domain log size */
  const uint32_t *l0 = line_layer(itw, tw_len, m, 0);
  for (uint64_t h = 0; h < (1ull << (log_size - 1)); h++)
    fft_layer_loop(values, 0, h, circle_twiddle(l0, h), 1);
  for (uint32_t layer = 0; layer < m; layer++) {
    const uint32_t *lt = line_layer(itw, tw_len, m, layer);
    for (uint64_t h = 0; h < (1ull << (m - 1 - layer)); h++)
      fft_layer_loop(values, layer + 1, h, lt[h], 1);
  }
  uint32_t inv = m_inv((uint32_t)(1ull << log_size)); /* :73-77 */
  for (uint64_t i = 0; i < (1ull << log_size); i++) values[i] = m_mul(values[i], inv);
}

/* cpu/circle.rs:178-238 PolyOps::evaluate / evaluate_into */
void or_evaluate(const uint32_t *tw, uint64_t tw_len, uint32_t half_initial, uint32_t half_log,
                 const uint32_t *coeffs, uint32_t coeff_log, uint32_t *buffer) {
  coset half = coset_new(half_initial, half_log);
  uint32_t log_size = half_log + 1;
  uint64_t size = 1ull << log_size, poly_len = 1ull << coeff_log;
  memcpy(buffer, coeffs, poly_len * 4); /* :196-201 */
  for (uint64_t i = poly_len; i < size; i++) buffer[i] = 0;
  if (log_size == 1) { /* :203-209 */
    butterfly(&buffer[0], &buffer[1], index_to_point(half.initial_index).y);
    return;
  }
  if (log_size == 2) { /* :211-223 */
    mpt p = index_to_point(half.initial_index);
    butterfly(&buffer[0], &buffer[2], p.x);
    butterfly(&buffer[1], &buffer[3], p.x);
    butterfly(&buffer[0], &buffer[1], p.y);
    butterfly(&buffer[2], &buffer[3], m_neg(p.y));
    return;
  }
  uint32_t m = half_log;
  for (int layer = (int)m - 1; layer >= 0; layer--) { /* :228-232 */
    const uint32_t *lt = line_layer(tw, tw_len, m, (uint32_t)layer);
    for (uint64_t h = 0; h < (1ull << (m - 1 - layer)); h++)
      fft_layer_loop(buffer, (uint32_t)layer + 1, h, lt[h], 0);
  }
  const uint32_t *l0 = line_layer(tw, tw_len, m, 0); /* :233-235 */
  for (uint64_t h = 0; h < (1ull << (log_size - 1)); h++)
    fft_layer_loop(buffer, 0, h, circle_twiddle(l0, h), 0);
}

/* crates/stwo/src/core/poly/utils.rs:27-38 fold, bottom-up */
static qm31 fold_values(const uint32_t *values, uint32_t k, const qm31 *factors) {
  uint64_t n = 1ull << k;
  qm31 *cur = (qm31 *)malloc(sizeof(qm31) * (n ? n : 1));
  for (uint64_t i = 0; i < n; i++) cur[i] = q_m(values[i]);
  for (int j = (int)k - 1; j >= 0; j--) {
    n >>= 1;
    for (uint64_t t = 0; t < n; t++) cur[t] = q_add(cur[2 * t], q_mul(cur[2 * t + 1], factors[j]));
  }
  qm31 r = cur[0];
  free(cur);
  return r;
}
/* cpu/circle.rs:81-98 eval_at_point */
void or_eval_at_point(const uint32_t *coeffs, uint32_t log_size, const uint32_t point[8],
                      uint32_t out[4]) {
  qpt p = qp_from8(point);
  if (log_size == 0) {
    q_to4(q_m(coeffs[0]), out);
    return;
  }
  qm31 mappings[64];
  uint32_t k = 0;
  mappings[k++] = p.y;
  qm31 x = p.x;
  for (uint32_t i = 1; i < log_size; i++) {
    mappings[k++] = x;
    x = q_double_x(x);
  }
  qm31 rev[64];
  for (uint32_t i = 0; i < k; i++) rev[i] = mappings[k - 1 - i];
  q_to4(fold_values(coeffs, log_size, rev), out);
}

/* ======================================================================
 * Vanishing polynomials (crates/stwo/src/core/constraints.rs:12-103)
 * ====================================================================== */
/* constraints.rs:12-35 coset_vanishing */
static qm31 coset_vanishing(coset c, qpt p) {
  p = qp_add(qp_sub(p, qp_from_m(index_to_point(c.initial_index))),
             qp_from_m(index_to_point(c.step_size >> 1)));
  qm31 x = p.x;
  for (uint32_t i = 1; i < c.log_size; i++) x = q_double_x(x);
  return x;
}
/* poly/circle/canonic.rs:41-44: CanonicCoset(n).coset = Coset::odds(n) */
static coset canonic_coset(uint32_t log_size) {
  return coset_new(subgroup_gen(log_size + 1), log_size); /* circle.rs:283-285 odds */
}
/* constraints.rs:37-55 coset_vanishing_derivative */
static qm31 coset_vanishing_derivative(uint32_t log_size, qpt p) {
  qm31 exp = q_one();
  for (uint32_t i = 1; i < log_size; i++) exp = q_mul(exp, q_m(4));
  qm31 van = q_one();
  for (uint32_t i = 1; i < log_size; i++) van = q_mul(van, coset_vanishing(canonic_coset(i), p));
  return q_mul(exp, van);
}
/* constraints.rs:83-91 point_vanishing */
static qm31 point_vanishing(qpt vanish_point, qpt p) {
  qpt h = qp_sub(p, vanish_point);
  return q_mul(h.y, q_inv(q_add(q_one(), h.x)));
}

/* cpu/circle.rs:100-131 barycentric_weights_into */
void or_barycentric_weights(uint32_t log_size, const uint32_t point[8], uint32_t *w0,
                            uint32_t *w1, uint32_t *w2, uint32_t *w3) {
  qpt p = qp_from8(point);
  coset half = coset_new(subgroup_gen(log_size + 1), log_size - 1); /* half_odds(log-1) */
  qm31 vn_p = coset_vanishing(canonic_coset(log_size), p);
  uint32_t *w[4] = {w0, w1, w2, w3};
  for (uint64_t i = 0; i < (1ull << log_size); i++) {
    qpt cp = qp_from_m(circle_domain_at(half, bit_reverse_index(i, log_size)));
    qm31 m2y = q_mul(cp.y, q_m(m_neg(2)));
    qm31 si = q_mul(m2y, coset_vanishing_derivative(log_size, cp));
    qm31 vi = point_vanishing(cp, p);
    uint32_t o[4];
    q_to4(q_mul(vn_p, q_inv(q_mul(si, vi))), o);
    for (int c = 0; c < 4; c++) w[c][i] = o[c];
  }
}
/* cpu/circle.rs:133-140 barycentric_eval_at_point */
void or_barycentric_eval(const uint32_t *evals, uint32_t log_size, const uint32_t *w0,
                         const uint32_t *w1, const uint32_t *w2, const uint32_t *w3,
                         uint32_t out[4]) {
  qm31 acc = q_zero();
  for (uint64_t i = 0; i < (1ull << log_size); i++) {
    qm31 wi = {{w0[i], w1[i]}, {w2[i], w3[i]}};
    acc = q_add(acc, q_mul_m(wi, evals[i]));
  }
  q_to4(acc, out);
}

/* ======================================================================
 * Blake2s-256 (RFC 7693), the function the `blake2` crate 0.10.6 computes
 * for crates/stwo/src/core/vcs/blake2_hash.rs:71-99.
 * ====================================================================== */
static const uint32_t B2_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                  0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t B2_SIGMA[10][16] = {
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
static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
static uint32_t ld32(const uint8_t *p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static void st32(uint8_t *p, uint32_t v) {
  p[0] = (uint8_t)v;
  p[1] = (uint8_t)(v >> 8);
  p[2] = (uint8_t)(v >> 16);
  p[3] = (uint8_t)(v >> 24);
}
static void b2_compress(uint32_t h[8], const uint8_t block[64], uint64_t t, int last) {
  uint32_t v[16], m[16];
  for (int i = 0; i < 16; i++) m[i] = ld32(block + 4 * i);
  for (int i = 0; i < 8; i++) {
    v[i] = h[i];
    v[i + 8] = B2_IV[i];
  }
  v[12] ^= (uint32_t)t;
  v[13] ^= (uint32_t)(t >> 32);
  if (last) v[14] = ~v[14];
#define G(a, b, c, d, x, y)        \
  do {                             \
    v[a] = v[a] + v[b] + (x);      \
    v[d] = rotr(v[d] ^ v[a], 16);  \
    v[c] = v[c] + v[d];            \
    v[b] = rotr(v[b] ^ v[c], 12);  \
    v[a] = v[a] + v[b] + (y);      \
    v[d] = rotr(v[d] ^ v[a], 8);   \
    v[c] = v[c] + v[d];            \
    v[b] = rotr(v[b] ^ v[c], 7);   \
  } while (0)
  for (int r = 0; r < 10; r++) {
    const uint8_t *s = B2_SIGMA[r];
    G(0, 4, 8, 12, m[s[0]], m[s[1]]);
    G(1, 5, 9, 13, m[s[2]], m[s[3]]);
    G(2, 6, 10, 14, m[s[4]], m[s[5]]);
    G(3, 7, 11, 15, m[s[6]], m[s[7]]);
    G(0, 5, 10, 15, m[s[8]], m[s[9]]);
    G(1, 6, 11, 12, m[s[10]], m[s[11]]);
    G(2, 7, 8, 13, m[s[12]], m[s[13]]);
    G(3, 4, 9, 14, m[s[14]], m[s[15]]);
  }
#undef G
  for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[i + 8];
}
typedef struct {
  uint32_t h[8];
  uint8_t buf[64];
  uint32_t buflen;
  uint64_t t;
} b2state;
static void b2_init(b2state *s) {
  for (int i = 0; i < 8; i++) s->h[i] = B2_IV[i];
  s->h[0] ^= 0x01010020u; /* digest length 32, no key, fanout 1, depth 1 */
  s->buflen = 0;
  s->t = 0;
  memset(s->buf, 0, 64);
}
/* streaming update: the last (possibly full) block is kept until finalize */
static void b2_update(b2state *s, const uint8_t *in, uint64_t len) {
  while (len > 0) {
    if (s->buflen == 64) {
      s->t += 64;
      b2_compress(s->h, s->buf, s->t, 0);
      s->buflen = 0;
    }
    uint64_t take = 64 - s->buflen;
    if (take > len) take = len;
    memcpy(s->buf + s->buflen, in, take);
    s->buflen += (uint32_t)take;
    in += take;
    len -= take;
  }
}
static void b2_final(b2state *s, uint8_t out[32]) {
  s->t += s->buflen;
  memset(s->buf + s->buflen, 0, 64 - s->buflen);
  b2_compress(s->h, s->buf, s->t, 1);
  for (int i = 0; i < 8; i++) st32(out + 4 * i, s->h[i]);
}
void or_blake2s(const uint8_t *data, uint64_t len, uint8_t out[32]) {
  b2state s;
  b2_init(&s);
  b2_update(&s, data, len);
  b2_final(&s, out);
}
static void b2_update_u32(b2state *s, uint32_t w) {
  uint8_t b[4];
  st32(b, w);
  b2_update(s, b, 4);
}

/* ======================================================================
 * Merkle trees
 * ====================================================================== */
/* cpu/merkle_lifted.rs:46-90 build_leaves; columns sorted ascending by length */
void or_merkle_build_leaves(const uint32_t *const *cols, const uint32_t *log_sizes, uint64_t n,
                            uint32_t lifting_log_size, uint8_t *out) {
  if (n == 0) { /* :48-50 */
    or_blake2s(NULL, 0, out);
    return;
  }
  b2state *prev = (b2state *)malloc(sizeof(b2state) * 2);
  b2_init(&prev[0]);
  prev[1] = prev[0];
  uint32_t prev_log = 1;
  uint64_t c = 0;
  while (c < n) {
    uint32_t log_size = log_sizes[c];
    uint64_t g_end = c;
    while (g_end < n && log_sizes[g_end] == log_size) g_end++;
    uint32_t log_ratio = log_size - prev_log; /* :56-60 */
    b2state *next = (b2state *)malloc(sizeof(b2state) << log_size);
    for (uint64_t idx = 0; idx < (1ull << log_size); idx++)
      next[idx] = prev[((idx >> (log_ratio + 1)) << 1) + (idx & 1)];
    free(prev);
    prev = next;
    /* :62-80: chunks of 16 columns, each row absorbs the chunk's values */
    for (uint64_t ch = c; ch < g_end; ch += 16) {
      uint64_t ch_end = ch + 16 < g_end ? ch + 16 : g_end;
      for (uint64_t i = 0; i < (1ull << log_size); i++)
        for (uint64_t k = ch; k < ch_end; k++) b2_update_u32(&prev[i], cols[k][i]);
    }
    prev_log = log_size;
    c = g_end;
  }
  uint32_t log_ratio = lifting_log_size - prev_log; /* :82-87 */
  if (log_ratio > 0) {
    b2state *next = (b2state *)malloc(sizeof(b2state) << lifting_log_size);
    for (uint64_t idx = 0; idx < (1ull << lifting_log_size); idx++)
      next[idx] = prev[((idx >> (log_ratio + 1)) << 1) + (idx & 1)];
    free(prev);
    prev = next;
    prev_log = lifting_log_size;
  }
  for (uint64_t i = 0; i < (1ull << prev_log); i++) b2_final(&prev[i], out + 32 * i);
  free(prev);
}
/* cpu/merkle_lifted.rs:92-97 build_next_layer; hash_children
 * (crates/stwo/src/core/vcs_lifted/blake2_merkle.rs:29-36) */
void or_merkle_next_layer(const uint8_t *prev, uint32_t prev_log_size, uint8_t *out) {
  uint64_t n = 1ull << (prev_log_size - 1);
  for (uint64_t i = 0; i < n; i++) or_blake2s(prev + 64 * i, 64, out + 32 * i);
}
/* cpu/blake2s.rs:11-24 commit_on_layer; hash_node (crates/stwo/src/core/vcs/blake2_merkle.rs:13-29) */
void or_commit_on_layer(uint32_t log_size, const uint8_t *prev, const uint32_t *const *cols,
                        uint64_t n, uint8_t *out) {
  for (uint64_t i = 0; i < (1ull << log_size); i++) {
    b2state s;
    b2_init(&s);
    if (prev) b2_update(&s, prev + 64 * i, 64);
    for (uint64_t k = 0; k < n; k++) b2_update_u32(&s, cols[k][i]);
    b2_final(&s, out + 32 * i);
  }
}

/* ======================================================================
 * FRI folds (crates/stwo/src/prover/backend/cpu/fri.rs)
 * ====================================================================== */
/* cpu/fri.rs:66-91 fold_line_cpu; LineDomain::at = coset.at(i).x (core/poly/line.rs:51-53) */
void or_fold_line(const uint32_t *const src[4], uint32_t line_initial, uint32_t log_size,
                  const uint32_t alpha_w[4], uint32_t *const dst[4]) {
  coset dom = coset_new(line_initial, log_size);
  qm31 alpha = q_from4(alpha_w);
  for (uint64_t i = 0; i < (1ull << (log_size - 1)); i++) {
    uint32_t x = coset_at(dom, bit_reverse_index(i << 1, log_size)).x;
    qm31 f0 = {{src[0][2 * i], src[1][2 * i]}, {src[2][2 * i], src[3][2 * i]}};
    qm31 f1 = {{src[0][2 * i + 1], src[1][2 * i + 1]}, {src[2][2 * i + 1], src[3][2 * i + 1]}};
    q_ibutterfly(&f0, &f1, m_inv(x));
    uint32_t o[4];
    q_to4(q_add(f0, q_mul(alpha, f1)), o);
    for (int c = 0; c < 4; c++) dst[c][i] = o[c];
  }
}
/* cpu/fri.rs:94-116 fold_circle_into_line_cpu */
void or_fold_circle_into_line(const uint32_t *const src[4], uint32_t half_initial,
                              uint32_t half_log, const uint32_t alpha_w[4],
                              uint32_t *const dst[4]) {
  coset half = coset_new(half_initial, half_log);
  uint32_t log_size = half_log + 1;
  qm31 alpha = q_from4(alpha_w);
  for (uint64_t i = 0; i < (1ull << (log_size - 1)); i++) {
    mpt p = circle_domain_at(half, bit_reverse_index(i << 1, log_size));
    qm31 f0 = {{src[0][2 * i], src[1][2 * i]}, {src[2][2 * i], src[3][2 * i]}};
    qm31 f1 = {{src[0][2 * i + 1], src[1][2 * i + 1]}, {src[2][2 * i + 1], src[3][2 * i + 1]}};
    q_ibutterfly(&f0, &f1, m_inv(p.y));
    uint32_t o[4];
    q_to4(q_add(q_mul(alpha, f1), f0), o);
    for (int c = 0; c < 4; c++) dst[c][i] = o[c];
  }
}

/* ======================================================================
 * DEEP quotients
 * ====================================================================== */
/* crates/stwo/src/core/constraints.rs:119-134 complex_conjugate_line_coeffs */
void or_line_coeffs(const uint32_t point_w[8], const uint32_t value_w[4], const uint32_t alpha_w[4],
                    uint32_t out[12]) {
  qpt point = qp_from8(point_w);
  qm31 value = q_from4(value_w), alpha = q_from4(alpha_w);
  if (q_eq(point.y, q_conj(point.y))) abort(); /* :125-130 */
  qm31 a = q_sub(q_conj(value), value);
  qm31 c = q_sub(q_conj(point.y), point.y);
  qm31 b = q_sub(q_mul(value, c), q_mul(a, point.y));
  q_to4(q_mul(alpha, a), out);
  q_to4(q_mul(alpha, b), out + 4);
  q_to4(q_mul(alpha, c), out + 8);
}
/* cpu/quotients.rs:23-49 accumulate_numerators (one sample batch);
 * accumulate_row_partial_numerators: crates/stwo/src/core/pcs/quotients.rs:178-193 */
void or_accumulate_numerators(const uint32_t *const *cols, uint32_t col_log, uint32_t log_blowup,
                              const uint32_t *col_idx, const uint32_t *line_coeffs,
                              uint64_t n_terms, uint32_t *const out[4],
                              uint32_t first_linear[4]) {
  uint64_t sub = 1ull << (col_log - log_blowup);
  for (uint64_t row = 0; row < sub; row++) {
    qm31 num = q_zero();
    for (uint64_t k = 0; k < n_terms; k++) {
      qm31 b = q_from4(line_coeffs + 12 * k + 4), c = q_from4(line_coeffs + 12 * k + 8);
      num = q_add(num, q_sub(q_mul_m(c, cols[col_idx[k]][row]), b));
    }
    uint32_t o[4];
    q_to4(num, o);
    for (int cc = 0; cc < 4; cc++) out[cc][row] = o[cc];
  }
  qm31 fl = q_zero();
  for (uint64_t k = 0; k < n_terms; k++) fl = q_add(fl, q_from4(line_coeffs + 12 * k));
  q_to4(fl, first_linear);
}
/* crates/stwo/src/core/pcs/quotients.rs:233-256 denominator_inverses (for one point) */
static cm31 denominator_inverse(qpt sp, mpt dp) {
  cm31 prx = sp.x.a, pry = sp.y.a, pix = sp.x.b, piy = sp.y.b;
  cm31 d = c_sub(c_mul(c_sub(prx, (cm31){dp.x, 0}), piy), c_mul(c_sub(pry, (cm31){dp.y, 0}), pix));
  return c_inv(d);
}
/* cpu/quotients.rs:51-101 compute_quotients_and_combine */
void or_compute_quotients_and_combine(const uint32_t *const *partials, const uint32_t *logs,
                                      const uint32_t *points, const uint32_t *first_linear,
                                      uint64_t n_acc, uint32_t L, uint32_t b,
                                      const uint32_t *tw, const uint32_t *itw, uint64_t tw_len,
                                      uint32_t *const out[4]) {
  /* eval_domain = CanonicCoset(L).circle_domain(): half = half_odds(L-1) */
  uint32_t half_init = subgroup_gen(L + 1);
  coset sub_half = coset_new(half_init, L - 1 - b); /* domain.rs:72-82 split */
  uint32_t sl = L - b;
  uint64_t sub_n = 1ull << sl;
  uint32_t *q[4];
  for (int c = 0; c < 4; c++) q[c] = (uint32_t *)malloc(4 * sub_n);
  for (uint64_t row = 0; row < sub_n; row++) {
    mpt dp = circle_domain_at(sub_half, bit_reverse_index(row, sl));
    qm31 quot = q_zero();
    for (uint64_t k = 0; k < n_acc; k++) {
      qpt sp = qp_from8(points + 8 * k);
      cm31 den_inv = denominator_inverse(sp, dp);
      uint32_t log_ratio = sl - logs[k];
      uint64_t li = ((row >> (log_ratio + 1)) << 1) + (row & 1);
      const uint32_t *const *pk = partials + 4 * k;
      qm31 pv = {{pk[0][li], pk[1][li]}, {pk[2][li], pk[3][li]}};
      qm31 full = q_sub(pv, q_mul_m(q_from4(first_linear + 4 * k), dp.y));
      quot = q_add(quot, q_mul_c(full, den_inv));
    }
    uint32_t o[4];
    q_to4(quot, o);
    for (int c = 0; c < 4; c++) q[c][row] = o[c];
  }
  /* :80-99 interpolate on the subdomain, evaluate on the full domain */
  uint64_t sub_tw_len = 1ull << (sl - 1);
  uint32_t *sub_itw = (uint32_t *)malloc(4 * sub_tw_len);
  or_extract_subdomain_twiddles(itw, tw_len, L, sl, sub_itw);
  for (int c = 0; c < 4; c++) {
    or_interpolate(sub_itw, sub_tw_len, sub_half.initial_index, sub_half.log_size, q[c]);
    or_evaluate(tw, tw_len, half_init, L - 1, q[c], sl, out[c]);
    free(q[c]);
  }
  free(sub_itw);
}

/* cpu/accumulation.rs:26-47 lift_and_accumulate */
void or_lift_and_accumulate(const uint32_t *const *cols, const uint32_t *logs, uint64_t n,
                            uint32_t *const out[4]) {
  uint32_t curr_log = 1;
  uint32_t *curr[4];
  for (int c = 0; c < 4; c++) curr[c] = (uint32_t *)calloc(2, 4);
  for (uint64_t k = 0; k < n; k++) {
    uint32_t lg = logs[k];
    uint32_t log_ratio = lg - curr_log;
    uint32_t *next[4];
    for (int c = 0; c < 4; c++) next[c] = (uint32_t *)malloc(4ull << lg);
    for (uint64_t i = 0; i < (1ull << lg); i++) {
      uint64_t li = ((i >> (log_ratio + 1)) << 1) + (i & 1);
      for (int c = 0; c < 4; c++) next[c][i] = m_add(cols[4 * k + c][i], curr[c][li]);
    }
    for (int c = 0; c < 4; c++) {
      free(curr[c]);
      curr[c] = next[c];
    }
    curr_log = lg;
  }
  for (int c = 0; c < 4; c++) {
    memcpy(out[c], curr[c], 4ull << curr_log);
    free(curr[c]);
  }
}

/* ======================================================================
 * Proof of work (crates/stwo/src/core/channel/blake2s.rs:102-117,
 * crates/stwo/src/prover/backend/cpu/grind.rs:5-15)
 * ====================================================================== */
#define POW_PREFIX 0x12345678u
static void pow_prefixed_digest(const uint8_t digest[32], uint32_t n_bits, uint8_t out[32]) {
  b2state s;
  b2_init(&s);
  b2_update_u32(&s, POW_PREFIX);
  b2_update(&s, digest, 32);
  b2_update_u32(&s, n_bits);
  b2_final(&s, out);
}
static int pow_check(const uint8_t prefixed[32], uint32_t n_bits, uint64_t nonce) {
  uint8_t buf[40], res[32];
  memcpy(buf, prefixed, 32);
  for (int i = 0; i < 8; i++) buf[32 + i] = (uint8_t)(nonce >> (8 * i));
  or_blake2s(buf, 40, res);
  /* trailing zeros of the little-endian u128 res[0..16] */
  uint32_t tz = 0;
  for (int i = 0; i < 16; i++) {
    if (res[i] == 0) {
      tz += 8;
      continue;
    }
    uint8_t v = res[i];
    while (!(v & 1)) {
      tz++;
      v >>= 1;
    }
    break;
  }
  return tz >= n_bits;
}
int or_verify_pow_nonce(const uint8_t digest[32], uint32_t n_bits, uint64_t nonce) {
  uint8_t pre[32];
  pow_prefixed_digest(digest, n_bits, pre);
  return pow_check(pre, n_bits, nonce);
}
uint64_t or_grind(const uint8_t digest[32], uint32_t n_bits) {
  uint8_t pre[32];
  pow_prefixed_digest(digest, n_bits, pre);
  uint64_t nonce = 0;
  while (!pow_check(pre, n_bits, nonce)) nonce++;
  return nonce;
}

/* ======================================================================
 * Blake2s channel (crates/stwo/src/core/channel/blake2s.rs:16-100)
 * ====================================================================== */
static void ch_update(or_channel *ch, const uint8_t d[32]) { /* :27-30 */
  memcpy(ch->digest, d, 32);
  ch->n_draws = 0;
}
void or_channel_mix_u32s(or_channel *ch, const uint32_t *words, uint64_t n) { /* :55-63 */
  b2state s;
  uint8_t d[32];
  b2_init(&s);
  b2_update(&s, ch->digest, 32);
  for (uint64_t i = 0; i < n; i++) b2_update_u32(&s, words[i]);
  b2_final(&s, d);
  ch_update(ch, d);
}
void or_channel_mix_u64(or_channel *ch, uint64_t v) { /* :65-67 */
  uint32_t w[2] = {(uint32_t)v, (uint32_t)(v >> 32)};
  or_channel_mix_u32s(ch, w, 2);
}
void or_channel_mix_felts(or_channel *ch, const uint32_t *felts, uint64_t n_felts) { /* :42-53 */
  or_channel_mix_u32s(ch, felts, 4 * n_felts);
}
/* crates/stwo/src/core/vcs_lifted/blake2_merkle.rs:56-58 */
void or_channel_mix_hash(or_channel *ch, const uint8_t root[32]) {
  uint8_t buf[64], d[32];
  memcpy(buf, ch->digest, 32);
  memcpy(buf + 32, root, 32);
  or_blake2s(buf, 64, d);
  ch_update(ch, d);
}
void or_channel_draw_u32s(or_channel *ch, uint32_t out[8]) { /* :80-99 */
  uint8_t buf[37], d[32];
  memcpy(buf, ch->digest, 32);
  st32(buf + 32, ch->n_draws);
  buf[36] = 0;
  ch->n_draws++;
  or_blake2s(buf, 37, d);
  for (int i = 0; i < 8; i++) out[i] = ld32(d + 4 * i);
}
void or_channel_draw_secure_felt(or_channel *ch, uint32_t out[4]) { /* :32-36, :69-74 */
  uint32_t u[8], f[8];
  or_channel_draw_u32s(ch, u);
  for (int i = 0; i < 8; i++) f[i] = reduce64(u[i]);
  for (int i = 0; i < 4; i++) out[i] = m_add(m_add(f[2 * i], f[2 * i + 1]), f[2 * i + 1]);
}
/* crates/stwo/src/core/circle.rs:150-160 get_random_point */
void or_channel_random_point(or_channel *ch, uint32_t out[8]) {
  uint32_t tw[4];
  or_channel_draw_secure_felt(ch, tw);
  qm31 t = q_from4(tw);
  qm31 t2 = q_mul(t, t);
  qm31 inv = q_inv(q_add(t2, q_one()));
  qm31 x = q_mul(q_add(q_one(), q_neg(t2)), inv);
  qm31 y = q_mul(q_add(t, t), inv);
  q_to4(x, out);
  q_to4(y, out + 4);
}
