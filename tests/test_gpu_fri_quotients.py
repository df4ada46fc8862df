"""GPU parity: FRI folds, DEEP quotients, accumulation, OOD sampling and grinding against the
oracle, bit-exact.  Mirrors crates/stwo/src/prover/backend/simd/fri.rs:316-390,
simd/quotients.rs:308-381, prover/air/accumulation.rs:230-311 and simd/circle.rs:786+.
"""
import numpy as np
import pytest

import oracle as O
from proving_amd.backend import BaseColumn, SecureColumnByCoords, canonic_half_initial

pytestmark = pytest.mark.gpu
P = O.P


def rnd(rng, shape):
    return rng.integers(0, P, shape, dtype=np.uint32)


@pytest.mark.parametrize("log", [1, 2, 3, 5, 10, 16])
def test_fold_circle_into_line(ctx, backend, log):
    rng = np.random.default_rng(log)
    tree = backend.precompute_canonic_twiddles(max(log, 1))
    src = rnd(rng, (4, 1 << log))
    alpha = rnd(rng, 4)
    out = backend.fold_circle_into_line(SecureColumnByCoords.from_cpu(ctx, src), alpha, tree)
    np.testing.assert_array_equal(out.to_cpu(), O.fold_circle_into_line(src, log, alpha))


@pytest.mark.parametrize("log,fold_step", [(1, 1), (2, 1), (4, 1), (9, 1), (15, 1), (6, 2), (8, 3), (10, 4)])
def test_fold_line(ctx, backend, log, fold_step):
    rng = np.random.default_rng(10 * log + fold_step)
    # the first FRI line domain of a circle domain of log (log + 1), inside the tree of log + 2
    tree = backend.precompute_canonic_twiddles(log + 2)
    line_init = canonic_half_initial(log + 1)
    src = rnd(rng, (4, 1 << log))
    alphas = [rnd(rng, 4) for _ in range(fold_step)]
    out = backend.fold_line(SecureColumnByCoords.from_cpu(ctx, src), line_init, alphas, tree)
    exp, init, lg = src, line_init, log
    for a in alphas:
        exp = O.fold_line(exp, init, lg, a)
        init, lg = (init * 2) % (1 << 31), lg - 1
    np.testing.assert_array_equal(out.to_cpu(), exp)


def test_eval_by_folding_on_gpu(ctx, backend, golden):
    """fold_circle_into_line + fold_line with the folding alphas gives 2^n f(p) == eval_at_point
    (crates/stwo/src/prover/backend/cpu/circle.rs:395-423), entirely on the device."""
    k = golden["reference_kats"]
    poly = np.array(k["eval_by_folding_poly"], np.uint32)
    log = 10
    tree = backend.precompute_canonic_twiddles(11)
    coeffs = BaseColumn.from_cpu(ctx, poly)
    ev = backend.evaluate(coeffs, log, tree).to_cpu()
    ch = O.Channel()
    pt = ch.random_point()
    x, y = list(pt[:4]), list(pt[4:])
    alphas = [None] * log
    alphas[log - 1] = y
    for i in range(log - 2, -1, -1):
        alphas[i] = x
        s = list(O.qm31_mul(x, x))
        x = [(2 * int(a) - (1 if j == 0 else 0)) % P for j, a in enumerate(s)]
    src = np.zeros((4, 1 << log), np.uint32)
    src[0] = ev
    cur = backend.fold_circle_into_line(SecureColumnByCoords.from_cpu(ctx, src), alphas.pop(), tree)
    init = canonic_half_initial(log)
    while cur.log_size > 0:
        cur = backend.fold_line(cur, init, [alphas.pop()], tree)
        init = (init * 2) % (1 << 31)
    inv = O.m31_inv(1 << log)
    got = [int(c[0]) * inv % P for c in cur.to_cpu()]
    assert got == list(O.eval_at_point(poly, pt))
    assert got == list(backend.eval_at_point(coeffs, pt))


@pytest.mark.parametrize("log", [5, 9, 16])
def test_decompose(ctx, backend, log):
    rng = np.random.default_rng(2 + log)
    src = rnd(rng, (4, 1 << log))
    g, lam = backend.decompose(SecureColumnByCoords.from_cpu(ctx, src))
    exp_g, exp_lam = O.fri_decompose(src)
    assert list(lam) == exp_lam
    np.testing.assert_array_equal(g.to_cpu(), exp_g)


def _samples(rng, polys, pts, alpha, periodic_first=False):
    """Sample batches as build_samples_with_randomness assigns powers: column by column, sample
    by sample (crates/stwo/src/core/pcs/quotients.rs:274-319)."""
    pow_ = [1, 0, 0, 0]
    batches = {i: [] for i in range(len(pts))}
    for ci, poly in enumerate(polys):
        for pi, pt in enumerate(pts):
            v = O.eval_at_point(poly, pt)
            batches[pi].append((ci, v, np.array(pow_, np.uint32)))
            pow_ = list(O.qm31_mul(pow_, alpha))
    return [(pts[i], batches[i]) for i in range(len(pts))]


@pytest.mark.parametrize("log_n,n_cols,n_pts", [(5, 3, 2), (8, 20, 3), (10, 7, 1), (12, 40, 4)])
def test_accumulate_numerators_and_combine(ctx, backend, log_n, n_cols, n_pts):
    rng = np.random.default_rng(log_n + n_cols)
    b = 1
    L = log_n + b
    tree = backend.precompute_canonic_twiddles(L)
    tw, itw = O.canonic_twiddles(L)
    polys = [rnd(rng, 1 << log_n) for _ in range(n_cols)]
    ldes_h = [O.evaluate(tw, c, L) for c in polys]
    ldes = [BaseColumn.from_cpu(ctx, e) for e in ldes_h]
    ch = O.Channel()
    ch.mix_u64(log_n)
    alpha = ch.draw_secure_felt()
    pts = [ch.random_point() for _ in range(n_pts)]
    batches = _samples(rng, polys, pts, alpha)
    accs = backend.accumulate_numerators(ldes, batches, b)
    exp_parts, exp_fls = [], []
    for (pt, terms), (gpt, part, fl) in zip(batches, accs):
        coeffs = np.stack([O.line_coeffs(pt, v, r) for _, v, r in terms])
        ep, efl = O.accumulate_numerators(ldes_h, [ci for ci, _, _ in terms], coeffs, b)
        np.testing.assert_array_equal(part.to_cpu(), ep)
        assert list(fl) == list(efl)
        exp_parts.append(ep)
        exp_fls.append(efl)
    out = backend.compute_quotients_and_combine(accs, L, b, tree)
    exp = O.compute_quotients_and_combine(exp_parts, pts, exp_fls, L, b, tw, itw)
    np.testing.assert_array_equal(out.to_cpu(), exp)
    # quotient is low degree: its high coefficients vanish
    for c in range(4):
        assert not O.interpolate(itw, exp[c], L)[1 << (L - b):].any()


def test_accumulate_numerators_general_coefficients(ctx, backend):
    """Batches whose coefficient matrix is not rank one (different column subsets per point)."""
    rng = np.random.default_rng(77)
    log_n, b = 7, 1
    L = log_n + b
    tw, _ = O.canonic_twiddles(L)
    polys = [rnd(rng, 1 << log_n) for _ in range(6)]
    ldes_h = [O.evaluate(tw, c, L) for c in polys]
    ldes = [BaseColumn.from_cpu(ctx, e) for e in ldes_h]
    ch = O.Channel()
    pts = [ch.random_point() for _ in range(3)]
    batches = []
    for pi, cols in enumerate([[0, 1, 2], [2, 3], [0, 4, 5, 5]]):
        terms = [(ci, O.eval_at_point(polys[ci], pts[pi]), rnd(rng, 4)) for ci in cols]
        batches.append((pts[pi], terms))
    accs = backend.accumulate_numerators(ldes, batches, b)
    for (pt, terms), (_, part, fl) in zip(batches, accs):
        coeffs = np.stack([O.line_coeffs(pt, v, r) for _, v, r in terms])
        ep, efl = O.accumulate_numerators(ldes_h, [ci for ci, _, _ in terms], coeffs, b)
        np.testing.assert_array_equal(part.to_cpu(), ep)
        assert list(fl) == list(efl)


def test_lift_and_accumulate(ctx, backend):
    rng = np.random.default_rng(4)
    cols_h = [rnd(rng, (4, 1 << lg)) for lg in (1, 3, 4, 9, 12)]
    out = backend.lift_and_accumulate([SecureColumnByCoords.from_cpu(ctx, c) for c in cols_h])
    np.testing.assert_array_equal(out.to_cpu(), O.lift_and_accumulate(cols_h))
    one = backend.lift_and_accumulate([SecureColumnByCoords.from_cpu(ctx, cols_h[3])])
    np.testing.assert_array_equal(one.to_cpu(), cols_h[3])


def test_accumulate(ctx, backend):
    rng = np.random.default_rng(6)
    a, b = rnd(rng, (4, 256)), rnd(rng, (4, 256))
    ca = SecureColumnByCoords.from_cpu(ctx, a)
    backend.accumulate(ca, SecureColumnByCoords.from_cpu(ctx, b))
    np.testing.assert_array_equal(ca.to_cpu(), ((a.astype(np.int64) + b) % P).astype(np.uint32))


@pytest.mark.parametrize("log", [0, 1, 3, 4, 5, 9, 14])
def test_eval_at_point(ctx, backend, log):
    rng = np.random.default_rng(log)
    coeffs = rnd(rng, 1 << log)
    ch = O.Channel()
    ch.mix_u64(log)
    pt = ch.random_point()
    got = backend.eval_at_point(BaseColumn.from_cpu(ctx, coeffs), pt)
    assert list(got) == list(O.eval_at_point(coeffs, pt))


@pytest.mark.parametrize("log", [2, 3, 6, 11])
def test_barycentric(ctx, backend, log):
    rng = np.random.default_rng(log)
    ch = O.Channel()
    ch.mix_u64(100 + log)
    pt = ch.random_point()
    w = backend.barycentric_weights(log, pt)
    np.testing.assert_array_equal(w.to_cpu(), O.barycentric_weights(log, pt))
    ev = rnd(rng, 1 << log)
    got = backend.barycentric_eval_at_point(BaseColumn.from_cpu(ctx, ev), w)
    assert list(got) == list(O.barycentric_eval(ev, O.barycentric_weights(log, pt)))
    tw, itw = O.canonic_twiddles(log)
    assert list(got) == list(O.eval_at_point(O.interpolate(itw, ev, log), pt))


def test_grind(ctx, backend, golden):
    for g in golden["hashlib"]["grind"]:
        assert backend.grind(bytes.fromhex(g["digest"]), g["bits"]) == g["nonce"]
    k = golden["reference_kats"]["verify_pow_nonce"]
    d = b"".join(int(w).to_bytes(4, "little") for w in k["digest_words"])
    n10 = backend.grind(d, 10)
    assert n10 <= 649 and O.verify_pow_nonce(d, 10, n10)
    assert all(not O.verify_pow_nonce(d, 10, j) for j in range(n10))
    n20 = backend.grind(d, 20)
    assert O.verify_pow_nonce(d, 20, n20)
    assert backend.grind(d, 0) == 0


@pytest.mark.parametrize("bits", [1, 4, 8, 12, 16])
def test_grind_smallest_nonce_vs_brute_force(ctx, backend, bits):
    """The chunked search returns the smallest valid nonce (the oracle's linear search), also
    when many chunks hold hits at once (small bits) and when the first hits lie past the first
    chunks (16 bits)."""
    rng = np.random.default_rng(500 + bits)
    for _ in range(4):
        d = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        want = O.grind(d, bits)
        assert backend.grind(d, bits) == want
        assert O.verify_pow_nonce(d, bits, want)


@pytest.mark.parametrize("n_pts", [1, 3, 4, 6])
def test_eval_at_points_batched(ctx, backend, n_pts):
    """Many polynomials of mixed sizes at many points in one call == eval_at_point of each."""
    rng = np.random.default_rng(70 + n_pts)
    logs = [0, 1, 3, 8, 9, 12, 9, 16, 3]
    polys = [rng.integers(0, P, 1 << lg, dtype=np.uint32) for lg in logs]
    ch = O.Channel()
    ch.mix_u64(n_pts)
    pts = [[int(v) for v in ch.random_point()] for _ in range(n_pts)]
    cols = [BaseColumn.from_cpu(ctx, p) for p in polys]
    got = backend.eval_at_points(cols, pts)
    for i, p in enumerate(polys):
        for j, pt in enumerate(pts):
            assert got[i][j] == [int(v) for v in O.eval_at_point(p, pt)], (logs[i], j)
