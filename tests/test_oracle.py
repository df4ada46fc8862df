"""Pins the CPU oracle (tests/golden + closed forms + invariants), no GPU needed.

Each check is chosen so that a plausible slip in the oracle (a dropped term, a
wrong sign or index, a transposed operand) fails it.
"""
import hashlib
import random

import numpy as np
import pytest

import oracle as O

P = O.P


def bitrev(i, n):
    return int(bin(i)[2:].zfill(n)[::-1], 2) if n else i


def q(m):
    return [m % P, 0, 0, 0]


# ---------------------------------------------------------------- fields
def test_field_kats(golden):
    k = golden["reference_kats"]
    for a, b, c in k["qm31_mul"]:
        assert list(O.qm31_mul(a, b)) == c
        assert list(O.qm31_mul(O.qm31_inv(b), c)) == a  # (qm0*qm1)/qm1 == qm0
    # CM31 embeds in QM31 as (a, b, 0, 0)
    for a, b, c in k["cm31_mul"]:
        assert list(O.qm31_mul(a + [0, 0], b + [0, 0])) == c + [0, 0]
    for v, r in k["m31_reduce"]:
        assert O.m31_reduce(v) == r
    for v, r in k["m31_partial_reduce"]:
        assert O.m31_add(v - (P - 1), P - 1) == r
    assert O.m31_mul(P - 1, P - 1) == 1


def test_m31_brute_force():
    rng = random.Random(0)
    for _ in range(2000):
        a, b = rng.randrange(P), rng.randrange(P)
        assert O.m31_mul(a, b) == a * b % P
    for a in [1, 2, 3, P - 1, 123456789]:
        assert O.m31_inv(a) * a % P == 1


def test_qm31_against_big_integer_model():
    """QM31 = CM31[u]/(u^2 - 2 - i), CM31 = M31[i]/(i^2 + 1): model with Python ints."""
    rng = random.Random(1)

    def cmul(x, y):
        return ((x[0] * y[0] - x[1] * y[1]) % P, (x[0] * y[1] + x[1] * y[0]) % P)

    def qmul(x, y):
        a, b, c, d = (x[0], x[1]), (x[2], x[3]), (y[0], y[1]), (y[2], y[3])
        ac, bd, ad, bc = cmul(a, c), cmul(b, d), cmul(a, d), cmul(b, c)
        rbd = cmul((2, 1), bd)
        return [(ac[0] + rbd[0]) % P, (ac[1] + rbd[1]) % P, (ad[0] + bc[0]) % P, (ad[1] + bc[1]) % P]

    for _ in range(200):
        x = [rng.randrange(P) for _ in range(4)]
        y = [rng.randrange(P) for _ in range(4)]
        assert list(O.qm31_mul(x, y)) == qmul(x, y)
        assert list(O.qm31_mul(x, O.qm31_inv(x))) == [1, 0, 0, 0]


# ---------------------------------------------------------------- circle
def test_circle_generator(golden):
    k = golden["reference_kats"]
    g = O.index_to_point(1)
    assert list(g) == k["m31_circle_gen"]
    # order 2^31: G * 2^30 != identity, G * 2^31 == identity (index arithmetic mod 2^31)
    assert list(O.index_to_point(1 << 30)) != [1, 0]
    assert list(O.index_to_point(0)) == [1, 0]
    x, y = (int(v) for v in O.index_to_point(12345))
    assert (x * x + y * y) % P == 1


def test_twiddle_layout(golden):
    """Layer l = x of the first half of C.double()^l, bit-reversed; pad word 1; itw = 1/tw."""
    for log in (1, 2, 3, 5, 8):
        init = O.canonic_half_initial(log + 1)
        tw, itw = O.precompute_twiddles(init, log)
        assert len(tw) == 1 << log and tw[-1] == 1
        off = 0
        c_init, c_log = init, log
        for layer in range(log):
            half = 1 << (c_log - 1)
            step = 1 << (31 - c_log)
            for j in range(half):
                idx = (c_init + step * bitrev(j, c_log - 1)) % (1 << 31)
                assert tw[off + j] == O.index_to_point(idx)[0]
            off += half
            c_init, c_log = (c_init * 2) % (1 << 31), c_log - 1
        assert all(int(a) * int(b) % P == 1 for a, b in zip(tw, itw))


def test_subdomain_twiddles_equal_direct():
    """crates/stwo/src/prover/poly/twiddles.rs:85-134: extracted == slow_precompute of subdomain."""
    for root_log, dom_log, sub_log in [(6, 6, 4), (8, 6, 4), (9, 9, 3)]:
        tw, _ = O.canonic_twiddles(root_log)
        ext = O.extract_subdomain_twiddles(tw, dom_log, sub_log)
        # subdomain half coset: Coset(half_odds(dom_log-1).initial, sub_log - 1)
        direct, _ = O.precompute_twiddles(O.canonic_half_initial(dom_log), sub_log - 1)
        assert list(ext[:-1]) == list(direct[:-1])


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 9])
def test_fft_closed_form_kat(n):
    """crates/stwo/src/prover/backend/simd/circle.rs:643-659: interpolate values 0..2^n (natural
    order) then eval_at_point(domain.at(i)) == i."""
    tw, itw = O.canonic_twiddles(n + 1)
    evals = np.array([bitrev(j, n) for j in range(1 << n)], np.uint32)  # stored bit-reversed
    coeffs = O.interpolate(itw, evals, n)
    half = O.canonic_half_initial(n)
    for i in range(0, 1 << n, max(1, (1 << n) // 16)):
        x, y = O.circle_domain_at(half, n - 1, i)
        assert list(O.eval_at_point(coeffs, [x, 0, 0, 0, y, 0, 0, 0])) == [i, 0, 0, 0]


def test_eval_at_point_basis(golden):
    k = golden["reference_kats"]["eval_at_point_4coeffs"]
    v = O.eval_at_point(k["coeffs"], q(k["x"]) + q(k["y"]))
    assert list(v) == [k["value"], 0, 0, 0]
    assert list(O.eval_at_point([1, 2], q(5) + q(8))) == [(1 + 2 * 8) % P, 0, 0, 0]
    assert list(O.eval_at_point([1], q(5) + q(8))) == [1, 0, 0, 0]


@pytest.mark.parametrize("n,m", [(3, 3), (3, 5), (4, 4), (2, 4), (1, 3), (0, 3), (5, 6)])
def test_evaluate_brute_force(n, m):
    """evaluate on CanonicCoset(m) == eval_at_point at every domain point (O(n^2) brute force)."""
    rng = np.random.default_rng(n * 10 + m)
    coeffs = rng.integers(0, P, 1 << n, dtype=np.uint32)
    tw, itw = O.canonic_twiddles(max(m, 2) + 1)
    ev = O.evaluate(tw, coeffs, m)
    half = O.canonic_half_initial(m)
    for j in range(1 << m):
        x, y = O.circle_domain_at(half, m - 1, bitrev(j, m))
        assert list(O.eval_at_point(coeffs, [x, 0, 0, 0, y, 0, 0, 0])) == [ev[j], 0, 0, 0]
    back = O.interpolate(itw, ev, m)
    assert list(back[: 1 << n]) == list(coeffs) and not back[1 << n:].any()


def test_eval_by_folding_equals_eval_at_point(golden):
    """crates/stwo/src/prover/backend/cpu/circle.rs:395-423: folding the evaluation with the
    folding alphas (fold_circle_into_line, then fold_line) gives 2^n * f(p); pins the fold
    formulas and their factor-2 convention."""
    k = golden["reference_kats"]
    poly = k["eval_by_folding_poly"]
    log = 10
    tw, itw = O.canonic_twiddles(11)
    ev = O.evaluate(tw, poly, log)
    gx, gy = k["secure_field_circle_gen"]
    for idx in k["eval_by_folding_point_indices"]:
        # CirclePoint::get_point(idx) = SECURE_FIELD_CIRCLE_GEN * idx (double-and-add over QM31)
        pt = _qpoint_mul(gx, gy, idx)
        expected = O.eval_at_point(poly, pt[0] + pt[1])
        # get_folding_alphas (crates/stwo/src/core/poly/utils.rs:45-68): [pi^(n-2)(x) .. x, y]
        alphas = [None] * log
        alphas[log - 1] = pt[1]
        x = pt[0]
        for i in range(log - 2, -1, -1):
            alphas[i] = x
            x = _qdouble_x(x)
        src = np.zeros((4, 1 << log), np.uint32)
        src[0] = ev
        cur = O.fold_circle_into_line(src, log, alphas.pop())
        line_init, line_log = O.canonic_half_initial(log), log - 1
        while cur.shape[1] > 1:
            cur = O.fold_line(cur, line_init, line_log, alphas.pop())
            line_init, line_log = (line_init * 2) % (1 << 31), line_log - 1
        inv = O.m31_inv(1 << log)
        got = [int(c[0]) * inv % P for c in cur]
        assert got == list(expected)


def _qadd(a, b):
    return [(int(x) + int(y)) % P for x, y in zip(a, b)]


def _qsub(a, b):
    return [(int(x) - int(y)) % P for x, y in zip(a, b)]


def _qdouble_x(x):
    s = list(O.qm31_mul(x, x))
    return _qsub(_qadd(s, s), [1, 0, 0, 0])


def _qpoint_mul(gx, gy, k):
    rx, ry = [1, 0, 0, 0], [0, 0, 0, 0]
    cx, cy = list(gx), list(gy)
    while k:
        if k & 1:
            rx, ry = (_qsub(list(O.qm31_mul(rx, cx)), list(O.qm31_mul(ry, cy))),
                      _qadd(list(O.qm31_mul(rx, cy)), list(O.qm31_mul(ry, cx))))
        cx, cy = (_qsub(list(O.qm31_mul(cx, cx)), list(O.qm31_mul(cy, cy))),
                  _qadd(list(O.qm31_mul(cx, cy)), list(O.qm31_mul(cy, cx))))
        k >>= 1
    return rx, ry


# ---------------------------------------------------------------- Blake2s / Merkle
def test_blake2s_kats(golden):
    k = golden["reference_kats"]
    assert O.blake2s(bytes.fromhex(k["blake2s_a"]["input_hex"])).hex() == k["blake2s_a"]["hash"]
    w = k["hash_u32s"]
    data = b"".join(int(x).to_bytes(4, "little") for x in w["words"])
    assert O.blake2s(data).hex() == w["hash"]
    assert O.blake2s(data + bytes.fromhex(w["hash"])).hex() == w["followed_by_digest"]
    for v in golden["hashlib"]["blake2s"]:
        d = bytes((i * 7 + 3) & 0xFF for i in range(v["len"]))
        assert O.blake2s(d).hex() == v["hash"]


def test_lifted_leaves_reference_kat(golden):
    k = golden["reference_kats"]["lifted_leaves_small"]
    cols = [np.arange(1 << lg, dtype=np.uint32) for lg in k["logs"]]
    leaves = O.merkle_build_leaves(cols, 4)
    assert leaves[0].tobytes().hex() == k["leaf0"]
    assert leaves[4].tobytes().hex() == k["leaf4"]
    assert leaves[15].tobytes().hex() == k["leaf15"]
    layers = O.merkle_commit_lifted(cols, 4)
    assert layers[0][0].tobytes().hex() == k["root"]


def _lcg_cols(seed, logs):
    x = seed
    cols = []
    for lg in logs:
        c = []
        for _ in range(1 << lg):
            x = (x * 6364136223846793005 + 1442695040888963407) & ((1 << 64) - 1)
            c.append((x >> 33) % P)
        cols.append(np.array(c, np.uint32))
    return cols


def test_trees_against_hashlib_definition(golden):
    for case in golden["hashlib"]["trees"]:
        cols = _lcg_cols(case["seed"], case["logs"])
        L = case["lifting_log"]
        layers = O.merkle_commit_lifted(cols, L)
        assert layers[0][0].tobytes().hex() == case["lifted_root"]
        for i, h in case["leaves_sample"].items():
            assert layers[-1][int(i)].tobytes().hex() == h
        mixed = O.merkle_commit_mixed(cols)
        assert mixed[0][0].tobytes().hex() == case["mixed_root"]


def test_lifted_root_equals_mixed_root_after_explicit_lift():
    """crates/stwo/src/prover/vcs_lifted/prover.rs:267-296: a column lifted explicitly to the
    tree height commits to the same root in both trees."""
    L = 6
    cols = _lcg_cols(9, [3, 6])
    lifted = np.array([cols[0][((i >> (L - 3 + 1)) << 1) + (i & 1)] for i in range(1 << L)], np.uint32)
    a = O.merkle_commit_lifted([lifted, cols[1]], L)[0][0]
    b = O.merkle_commit_lifted(cols, L)[0][0]
    c = O.merkle_commit_mixed([lifted, cols[1]])[0][0]
    assert a.tobytes() == b.tobytes() == c.tobytes()


def test_empty_tree():
    assert O.merkle_build_leaves([], 0)[0].tobytes() == hashlib.blake2s(b"").digest()


# ---------------------------------------------------------------- channel / PoW
def test_channel_kats(golden):
    k = golden["reference_kats"]
    ch = O.Channel()
    ch.mix_u64(k["mix_u64"]["value"])
    assert ch.digest.hex() == k["mix_u64"]["digest"]
    ch = O.Channel()
    ch.mix_u32s(k["mix_u32s"]["words"])
    assert ch.digest.hex() == k["mix_u32s"]["digest"]
    ch = O.Channel()
    ch.draw_u32s()
    assert ch.n_draws == 1
    for _ in range(9):
        ch.draw_secure_felt()
    assert ch.n_draws == 10
    tr = golden["hashlib"]["channel"]
    ch = O.Channel()
    assert list(ch.draw_u32s()) == tr["draw_u32s_0"]
    assert list(ch.draw_secure_felt()) == tr["draw_secure_felt_1"]
    ch.mix_u32s([7, 8, 9])
    assert ch.digest.hex() == tr["after_mix_digest"] and ch.n_draws == 0
    assert list(ch.draw_secure_felt()) == tr["draw_secure_felt_after_mix"]


def test_random_point_on_circle():
    ch = O.Channel()
    for _ in range(5):
        p = ch.random_point()
        x, y = list(p[:4]), list(p[4:])
        s = _qadd(list(O.qm31_mul(x, x)), list(O.qm31_mul(y, y)))
        assert s == [1, 0, 0, 0]


def test_pow_kats(golden):
    k = golden["reference_kats"]["verify_pow_nonce"]
    d = b"".join(int(w).to_bytes(4, "little") for w in k["digest_words"])
    for bits, nonce in k["true"]:
        assert O.verify_pow_nonce(d, bits, nonce)
    for bits, nonce in k["false"]:
        assert not O.verify_pow_nonce(d, bits, nonce)
    for g in golden["hashlib"]["grind"]:
        assert O.grind(bytes.fromhex(g["digest"]), g["bits"]) == g["nonce"]


# ---------------------------------------------------------------- quotients / accumulation
def test_lift_and_accumulate_definition():
    rng = np.random.default_rng(3)
    cols = [rng.integers(0, P, (4, 1 << lg), dtype=np.uint32) for lg in (2, 4, 5)]
    out = O.lift_and_accumulate(cols)
    L = 5
    for i in range(1 << L):
        exp = np.zeros(4, np.int64)
        for c in cols:
            n = c.shape[1].bit_length() - 1
            exp += c[:, ((i >> (L - n + 1)) << 1) + (i & 1)]
        assert list(out[:, i]) == list(exp % P)


def test_lift_is_evaluation_at_doubled_points():
    """What lifting means geometrically (crates/stwo/src/core/vcs_lifted/ and
    crates/stwo/src/prover/air/accumulation.rs:26-47): a column of the evaluations of p on
    CanonicCoset(n), lifted to log size L, holds p at the (L - n)-fold doubled points of
    CanonicCoset(L) (bit-reversed order).  Pins the index map ((i >> (L-n+1)) << 1) + (i & 1) of
    lift_and_accumulate and of the lifted Merkle leaves against the circle group, not against
    its own formula (the maps i >> d, (i >> (d+1)) << 1 and a shifted parity bit all fail it)."""
    from oracle import protocol as PR
    rng = np.random.default_rng(11)
    L = 7

    def dbl(x, y):  # the doubling map pi on circle points
        return (2 * x * x - 1) % P, (2 * x * y) % P

    half = O.canonic_half_initial(L)
    for n in (2, 3, 5, 7):
        coeffs = rng.integers(0, P, 1 << n, dtype=np.uint32)
        tw, _ = O.canonic_twiddles(n)
        col = O.evaluate(tw, coeffs, n)
        cols = [np.vstack([col[None, :], np.zeros((3, 1 << n), np.uint32)])]
        if n < L:
            cols.append(np.zeros((4, 1 << L), np.uint32))
        lifted = O.lift_and_accumulate(cols)[0]
        for i in range(1 << L):
            x, y = (int(v) for v in O.index_to_point(PR._circle_domain_index(half, L - 1, PR.bitrev(i, L))))
            for _ in range(L - n):
                x, y = dbl(x, y)
            assert int(lifted[i]) == int(O.eval_at_point(coeffs, [x, 0, 0, 0, y, 0, 0, 0])[0]), (n, i)


def test_quotients_low_degree_and_match_row_formula():
    """crates/stwo/src/prover/pcs/quotient_ops.rs:188-225: with exact sample values the combined
    quotient is a polynomial of degree < 2^(L - b) (high coefficients vanish).  The pointwise
    values also match the verifier's row formula (crates/stwo/src/core/pcs/quotients.rs:126-168)."""
    log_n, b = 5, 1
    L = log_n + b
    rng = np.random.default_rng(5)
    tw, itw = O.canonic_twiddles(L)
    polys = [rng.integers(0, P, 1 << log_n, dtype=np.uint32) for _ in range(3)]
    ldes = [O.evaluate(tw, c, L) for c in polys]
    ch = O.Channel()
    ch.mix_u64(99)
    alpha = list(ch.draw_secure_felt())
    pts = [ch.random_point() for _ in range(2)]
    partials, fls, pow_ = [], [], [1, 0, 0, 0]
    all_terms = []
    for pt in pts:
        coeffs, idx = [], []
        for ci, poly in enumerate(polys):
            v = O.eval_at_point(poly, pt)
            coeffs.append(O.line_coeffs(pt, v, pow_))
            all_terms.append((pt, ci, coeffs[-1]))
            idx.append(ci)
            pow_ = list(O.qm31_mul(pow_, alpha))
        part, fl = O.accumulate_numerators(ldes, idx, np.stack(coeffs), b)
        partials.append(part)
        fls.append(fl)
    out = O.compute_quotients_and_combine(partials, pts, fls, L, b, tw, itw)
    for c in range(4):
        coeffs = O.interpolate(itw, out[c], L)
        assert not coeffs[1 << (L - b):].any()
    # row formula at a few rows
    half = O.canonic_half_initial(L)
    for row in (0, 3, 17, 63):
        x, y = (int(v) for v in O.circle_domain_at(half, L - 1, bitrev(row, L)))
        acc = [0, 0, 0, 0]
        for pt in pts:
            num = [0, 0, 0, 0]
            for (tp, ci, cf) in all_terms:
                if list(tp) != list(pt):
                    continue
                a, bb, cc = list(cf[:4]), list(cf[4:8]), list(cf[8:])
                val = [int(t) * int(ldes[ci][row]) % P for t in cc]
                lin = _qadd([int(t) * y % P for t in a], bb)
                num = _qadd(num, _qsub(val, lin))
            prx, pix = [int(pt[0]), int(pt[1])], [int(pt[2]), int(pt[3])]
            pry, piy = [int(pt[4]), int(pt[5])], [int(pt[6]), int(pt[7])]
            d0 = _cmul([(prx[0] - x) % P, prx[1]], piy)
            d1 = _cmul([(pry[0] - y) % P, pry[1]], pix)
            den = [(d0[0] - d1[0]) % P, (d0[1] - d1[1]) % P]
            dinv = list(O.qm31_inv(den + [0, 0]))
            acc = _qadd(acc, list(O.qm31_mul(num, dinv)))
        assert list(out[:, row]) == acc


def _cmul(x, y):
    x, y = [int(v) for v in x], [int(v) for v in y]
    return [(x[0] * y[0] - x[1] * y[1]) % P, (x[0] * y[1] + x[1] * y[0]) % P]


def test_barycentric_equals_eval_at_point():
    """crates/stwo/src/prover/backend/simd/circle.rs:786+: barycentric evaluation of the
    evaluations == eval_at_point of the interpolated polynomial."""
    for log in (3, 5):
        rng = np.random.default_rng(log)
        tw, itw = O.canonic_twiddles(log)
        ev = rng.integers(0, P, 1 << log, dtype=np.uint32)
        coeffs = O.interpolate(itw, ev, log)
        ch = O.Channel()
        pt = ch.random_point()
        w = O.barycentric_weights(log, pt)
        assert list(O.barycentric_eval(ev, w)) == list(O.eval_at_point(coeffs, pt))


def _pi_q(x):
    """pi(x) = 2x^2 - 1 over QM31 (crates/stwo/src/core/circle.rs:31-36, double_x)."""
    sq = [int(v) for v in O.qm31_mul(x, x)]
    return [(2 * sq[0] - 1) % P] + [2 * v % P for v in sq[1:]]


@pytest.mark.parametrize("L", [1, 2, 3, 6, 10])
def test_split_at_mid_identity(L):
    """test_secure_circle_poly_split_at_mid (crates/stwo/src/prover/poly/circle/secure_poly.rs:
    152-170): p(z) = p_left(z) + pi^(L-2)(z.x) * p_right(z) at a random QM31 point, with the
    halves of the coefficient vector (L = 1: the factor is y, the basis element of bit 0)."""
    rng = np.random.default_rng(L)
    coeffs = rng.integers(0, P, 1 << L, dtype=np.uint32)
    ch = O.Channel()
    ch.mix_u64(L)
    z = [int(v) for v in ch.random_point()]
    left, right = coeffs[: 1 << (L - 1)], coeffs[1 << (L - 1):]
    if L == 1:
        f = z[4:]
    else:
        f = z[:4]
        for _ in range(L - 2):
            f = _pi_q(f)
    lhs = [int(v) for v in O.eval_at_point(coeffs, z)]
    rv = [int(v) for v in O.eval_at_point(right, z)]
    lv = [int(v) for v in O.eval_at_point(left, z)]
    fr = [int(v) for v in O.qm31_mul(f, rv)]
    assert lhs == [(a + b) % P for a, b in zip(lv, fr)]
    # a dropped factor fails: p(z) != p_left(z) + p_right(z) in general
    assert lhs != [(a + b) % P for a, b in zip(lv, rv)]


@pytest.mark.parametrize("log", [5, 6, 8, 11])
def test_decompose_splits_off_the_component_outside_the_fft_space(log):
    """FriOps::decompose pinned by what it is for (the reference's
    decompose_coeff_out_fft_space_test, crates/stwo/src/prover/backend/cpu/fri.rs:163-210): the
    circle basis element at index 2^(log-1) lies outside the FFT space of log - 1 and takes the
    value +c on the first half of the bit-reversed CanonicCoset(log) and -c on the second, so
    p + kappa * that element (p in the FFT space) decomposes into exactly the evaluations of p
    and lambda = kappa * c.  A wrong sign or divisor in lambda breaks both equalities."""
    rng = np.random.default_rng(log)
    n, h = 1 << log, 1 << (log - 1)
    tw, itw = O.canonic_twiddles(log)
    e = np.zeros(n, np.uint32)
    e[h] = 1
    ev = O.evaluate(tw, e, log)
    c = int(ev[0])
    assert c != 0 and np.all(ev[:h] == c) and np.all(ev[h:] == P - c)
    p = rng.integers(0, P, (4, h), dtype=np.uint32)
    p[0] = 0  # the reference's case: only the out-of-space coefficient
    kappa = [int(x) for x in rng.integers(1, P, 4)]
    f = np.empty((4, n), np.uint32)
    fft_part = np.empty((4, n), np.uint32)
    for k in range(4):
        coeffs = np.zeros(n, np.uint32)
        coeffs[:h] = p[k]
        fft_part[k] = O.evaluate(tw, coeffs, log)
        coeffs[h] = kappa[k]
        f[k] = O.evaluate(tw, coeffs, log)
    g, lam = O.fri_decompose(f)
    assert lam == [kappa[k] * c % P for k in range(4)]
    np.testing.assert_array_equal(g, fft_part)
    for k in range(4):
        assert not O.interpolate(itw, g[k], log)[h:].any()
