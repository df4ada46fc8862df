// M31 / CM31 / QM31 arithmetic usable on host and device (gfx950).
//
// Every function maps canonical inputs in [0, P) to canonical outputs, which is
// the representation CpuBackend uses (crates/stwo/src/core/fields/m31.rs:46-108)
// and the one every HIP kernel in this library reads and writes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SH_HD __host__ __device__ __forceinline__

namespace sh {

constexpr uint32_t P = 0x7fffffffu;

// min(v, v - P) as unsigned: v - P wraps above v when v < P.
SH_HD uint32_t reduce_once(uint32_t v) {
  uint32_t w = v - P;
  return w < v ? w : v;
}
SH_HD uint32_t madd(uint32_t a, uint32_t b) { return reduce_once(a + b); }
// a - b wraps below zero to 2^32 - (b - a), where adding P lands in [0, P):
// the smaller of d and d + P is the canonical difference (3 ops)
SH_HD uint32_t msub(uint32_t a, uint32_t b) {
  uint32_t d = a - b, e = d + P;
  return e < d ? e : d;
}
SH_HD uint32_t mneg(uint32_t a) { return reduce_once(P - a); }
// a*b mod P for a, b < 2^31: the 62-bit product splits as hi*2^31 + lo and
// 2^31 = 1 (mod P), so the residue is hi + lo < 2P.
SH_HD uint32_t mmul(uint32_t a, uint32_t b) {
  uint64_t p = (uint64_t)a * b;
  uint32_t lo = (uint32_t)p & P;
  uint32_t hi = (uint32_t)(p >> 31);
  return reduce_once(lo + hi);
}
// a*b mod P with the second factor given doubled (b2 = 2b < 2^32, b < P), the SIMD backend's
// twiddle form (crates/stwo/src/prover/backend/simd/circle.rs:452-459): the 64-bit product 2ab
// has ab >> 31 as its high word and (ab mod 2^31) << 1 as its low word, so the split costs one
// shift instead of a mask and a funnel shift.
SH_HD uint32_t mmul_dbl(uint32_t a, uint32_t b2) {
  const uint64_t p = (uint64_t)a * b2;
  return reduce_once((uint32_t)(p >> 32) + ((uint32_t)p >> 1));
}
SH_HD uint32_t mpow(uint32_t a, uint32_t e) {
  uint32_t r = 1;
  while (e) {
    if (e & 1) r = mmul(r, a);
    a = mmul(a, a);
    e >>= 1;
  }
  return r;
}
// a^(P-2) (Fermat; 0 maps to 0) by an addition chain: P - 2 = 2^31 - 3 = 4 (2^29 - 1) + 1, with
// a^(2^k - 1) built from a^(2^j - 1) as (a^(2^j - 1))^(2^(k-j)) * a^(2^(k-j) - 1): 30 squarings and
// 8 multiplications instead of the 60 of square-and-multiply.
SH_HD uint32_t msqn(uint32_t a, int n) {
  for (int i = 0; i < n; i++) a = mmul(a, a);
  return a;
}
SH_HD uint32_t minv(uint32_t a) {
  const uint32_t e2 = mmul(mmul(a, a), a);    // a^(2^2 - 1)
  const uint32_t e4 = mmul(msqn(e2, 2), e2);  // a^(2^4 - 1)
  const uint32_t e8 = mmul(msqn(e4, 4), e4);  // a^(2^8 - 1)
  const uint32_t e16 = mmul(msqn(e8, 8), e8);
  const uint32_t e24 = mmul(msqn(e16, 8), e8);
  const uint32_t e28 = mmul(msqn(e24, 4), e4);
  const uint32_t e29 = mmul(mmul(e28, e28), a);
  return mmul(msqn(e29, 2), a);               // a^(4 (2^29 - 1) + 1)
}

struct cm31 {
  uint32_t a, b;
};
SH_HD cm31 cadd(cm31 x, cm31 y) { return {madd(x.a, y.a), madd(x.b, y.b)}; }
SH_HD cm31 csub(cm31 x, cm31 y) { return {msub(x.a, y.a), msub(x.b, y.b)}; }
SH_HD cm31 cneg(cm31 x) { return {mneg(x.a), mneg(x.b)}; }
SH_HD cm31 cmul(cm31 x, cm31 y) {
  return {msub(mmul(x.a, y.a), mmul(x.b, y.b)), madd(mmul(x.a, y.b), mmul(x.b, y.a))};
}
SH_HD cm31 cmulm(cm31 x, uint32_t m) { return {mmul(x.a, m), mmul(x.b, m)}; }
SH_HD cm31 cinv(cm31 x) {
  uint32_t n = madd(mmul(x.a, x.a), mmul(x.b, x.b));
  return cmulm({x.a, mneg(x.b)}, minv(n));
}

struct qm31 {
  cm31 a, b;  // a + b*u, u^2 = 2 + i; words [a.a, a.b, b.a, b.b]
};
SH_HD qm31 qzero() { return {{0, 0}, {0, 0}}; }
SH_HD qm31 qone() { return {{1, 0}, {0, 0}}; }
SH_HD qm31 qfromm(uint32_t m) { return {{m, 0}, {0, 0}}; }
SH_HD qm31 qload(const uint32_t* w) { return {{w[0], w[1]}, {w[2], w[3]}}; }
SH_HD void qstore(qm31 x, uint32_t* w) {
  w[0] = x.a.a;
  w[1] = x.a.b;
  w[2] = x.b.a;
  w[3] = x.b.b;
}
SH_HD qm31 qadd(qm31 x, qm31 y) { return {cadd(x.a, y.a), cadd(x.b, y.b)}; }
SH_HD qm31 qsub(qm31 x, qm31 y) { return {csub(x.a, y.a), csub(x.b, y.b)}; }
SH_HD qm31 qneg(qm31 x) { return {cneg(x.a), cneg(x.b)}; }
// (a + bu)(c + du) = (ac + (2+i) bd) + (ad + bc) u
SH_HD qm31 qmul(qm31 x, qm31 y) {
  cm31 bd = cmul(x.b, y.b);
  // (2 + i) * (p + qi) = (2p - q) + (p + 2q) i
  cm31 rbd = {msub(madd(bd.a, bd.a), bd.b), madd(bd.a, madd(bd.b, bd.b))};
  return {cadd(cmul(x.a, y.a), rbd), cadd(cmul(x.a, y.b), cmul(x.b, y.a))};
}
SH_HD qm31 qmulm(qm31 x, uint32_t m) { return {cmulm(x.a, m), cmulm(x.b, m)}; }
SH_HD qm31 qmulc(qm31 x, cm31 c) { return {cmul(x.a, c), cmul(x.b, c)}; }
SH_HD qm31 qinv(qm31 x) {
  cm31 b2 = cmul(x.b, x.b);
  cm31 ib2 = {mneg(b2.b), b2.a};
  cm31 denom = csub(cmul(x.a, x.a), cadd(cadd(b2, b2), ib2));
  cm31 di = cinv(denom);
  return {cmul(x.a, di), cneg(cmul(x.b, di))};
}
SH_HD qm31 qconj(qm31 x) { return {x.a, cneg(x.b)}; }
SH_HD bool qeq(qm31 x, qm31 y) {
  return x.a.a == y.a.a && x.a.b == y.a.b && x.b.a == y.b.a && x.b.b == y.b.b;
}

// Circle points over M31 and QM31 (crates/stwo/src/core/circle.rs:111-117).
struct mpt {
  uint32_t x, y;
};
SH_HD mpt padd(mpt p, mpt q) {
  return {msub(mmul(p.x, q.x), mmul(p.y, q.y)), madd(mmul(p.x, q.y), mmul(p.y, q.x))};
}
struct qpt {
  qm31 x, y;
};
SH_HD qpt qpadd(qpt p, qpt q) {
  return {qsub(qmul(p.x, q.x), qmul(p.y, q.y)), qadd(qmul(p.x, q.y), qmul(p.y, q.x))};
}
SH_HD qm31 qdouble_x(qm31 x) {
  qm31 s = qmul(x, x);
  return qsub(qadd(s, s), qone());
}

// Index arithmetic of the circle group of order 2^31 (circle.rs:197-262).
SH_HD uint32_t idx_reduce(uint64_t i) { return (uint32_t)(i & 0x7fffffffull); }
SH_HD uint32_t subgroup_gen(uint32_t log_size) { return 1u << (31 - log_size); }

SH_HD uint32_t bitrev(uint32_t i, uint32_t log_n) {
#ifdef __HIP_DEVICE_COMPILE__
  return log_n == 0 ? i : (__brev(i) >> (32 - log_n));
#else
  if (log_n == 0) return i;
  uint32_t r = 0;
  for (uint32_t b = 0; b < log_n; b++) r |= ((i >> b) & 1u) << (log_n - 1 - b);
  return r;
#endif
}

}  // namespace sh
