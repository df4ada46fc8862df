"""GPU parity: twiddle tree, interpolate, evaluate (LDE) against the CPU oracle, bit-exact.

Mirrors the reference's differential tests (crates/stwo/src/prover/backend/simd/fft/
rfft.rs:911-1011, ifft.rs:571-725, simd/circle.rs:606-736) with the oracle standing in for
CpuBackend.
"""
import numpy as np
import pytest

import oracle as O
from proving_amd.backend import BaseColumn, canonic_half_initial

pytestmark = pytest.mark.gpu
P = O.P


def rand_col(rng, log):
    return rng.integers(0, P, 1 << log, dtype=np.uint32)


@pytest.mark.parametrize("root_log", [0, 1, 2, 3, 5, 10, 14, 20])
def test_twiddles_match_cpu(backend, root_log):
    init = canonic_half_initial(root_log + 1)
    t = backend.precompute_twiddles(init, root_log)
    tw, itw = t.to_cpu()
    etw, eitw = O.precompute_twiddles(init, root_log)
    np.testing.assert_array_equal(tw, etw)
    np.testing.assert_array_equal(itw, eitw)
    t.free()


def test_twiddles_non_canonic_root(backend):
    t = backend.precompute_twiddles(12345, 9)
    tw, itw = t.to_cpu()
    etw, eitw = O.precompute_twiddles(12345, 9)
    np.testing.assert_array_equal(tw, etw)
    np.testing.assert_array_equal(itw, eitw)
    t.free()


@pytest.mark.parametrize("log", [1, 2, 3, 4, 5, 8, 12, 13, 14, 16, 18])
def test_interpolate_matches_cpu(ctx, backend, log):
    rng = np.random.default_rng(log)
    tree = backend.precompute_canonic_twiddles(log + 1)
    _, itw = O.canonic_twiddles(log + 1)
    ev = rand_col(rng, log)
    col = BaseColumn.from_cpu(ctx, ev)
    backend.interpolate(col, tree)
    np.testing.assert_array_equal(col.to_cpu(), O.interpolate(itw, ev, log))
    col.free()
    tree.free()


@pytest.mark.parametrize("n,m", [(1, 1), (2, 2), (1, 2), (0, 1), (0, 2), (3, 3), (3, 4), (0, 4), (5, 7), (12, 13),
                                 (13, 14), (14, 15), (16, 17), (17, 18), (10, 14), (18, 19)])
def test_evaluate_matches_cpu(ctx, backend, n, m):
    rng = np.random.default_rng(100 * n + m)
    tree = backend.precompute_canonic_twiddles(max(m, 1))
    tw, _ = O.canonic_twiddles(max(m, 1))
    coeffs = rand_col(rng, n)
    c = BaseColumn.from_cpu(ctx, coeffs)
    out = backend.evaluate(c, m, tree)
    np.testing.assert_array_equal(out.to_cpu(), O.evaluate(tw, coeffs, m))
    for b in (c, out):
        b.free()
    tree.free()


# pass shapes whose strided layers leave a remainder of one or two layers: the planner folds it into
# the block pass (block + remainder <= 13) or splits the last two strided passes
@pytest.mark.parametrize("kind,n,m,shape", [
    ("interpolate", 15, 15, (11, 3, 3, 3)),   # 4 strided layers: remainder 1 -> block 12
    ("interpolate", 17, 17, (11, 4, 4, 3)),   # remainder 2 -> block 13
    ("interpolate", 18, 18, (13, 4, 4, 4)),   # remainder 1, block full -> passes of 3 and 2
    ("interpolate", 16, 16, (13, 2, 3, 3)),   # remainder 1 with 2-layer passes
    ("evaluate", 16, 17, (12, 3, 3, 3)),      # the LDE plans its 16 computed layers
    ("evaluate", 18, 19, (13, 4, 2, 4)),
])
def test_pass_plans_with_remainders(ctx, backend, kind, n, m, shape):
    rng = np.random.default_rng(7 * n + m)
    saved = backend.fft_tuning(m)
    backend.set_fft_tuning(m, *shape)
    try:
        tree = backend.precompute_canonic_twiddles(m)
        tw, itw = O.canonic_twiddles(m)
        vals = rand_col(rng, n)
        col = BaseColumn.from_cpu(ctx, vals)
        if kind == "interpolate":
            backend.interpolate(col, tree)
            np.testing.assert_array_equal(col.to_cpu(), O.interpolate(itw, vals, n))
        else:
            out = backend.evaluate(col, m, tree)
            np.testing.assert_array_equal(out.to_cpu(), O.evaluate(tw, vals, m))
            out.free()
        col.free()
        tree.free()
    finally:
        backend.set_fft_tuning(m, *saved)


# every family of the quad-layout pass kernels: block passes of 2^9..2^13 words and strided passes of
# 1..8 layers, each with 2..16 rows per lane (shape = block_log, stride_layers, r_block, r_stride)
QUAD_CASES = [(kind, n, m, (blk, st, r, r))
              for r in (1, 2, 3, 4)
              for kind, n, m, blk, st in [("interpolate", 9, 9, 13, 8), ("evaluate", 10, 11, 13, 8),
                                          ("interpolate", 11, 11, 13, 8), ("evaluate", 12, 13, 12, 8),
                                          ("interpolate", 13, 13, 13, 8), ("interpolate", 14, 14, 13, 1),
                                          ("evaluate", 15, 16, 13, 2), ("interpolate", 16, 16, 13, 3),
                                          ("evaluate", 17, 18, 13, 4), ("interpolate", 18, 18, 13, 5),
                                          ("evaluate", 19, 20, 13, 6), ("interpolate", 20, 20, 13, 7),
                                          ("evaluate", 19, 20, 11, 8)]]


@pytest.mark.parametrize("kind,n,m,shape", QUAD_CASES)
def test_quad_pass_kernels(ctx, backend, kind, n, m, shape):
    test_pass_plans_with_remainders(ctx, backend, kind, n, m, shape)


def test_round_trip_closed_form(ctx, backend):
    """Interpolating values 0..2^n (natural order) gives a polynomial with f(domain.at(i)) = i:
    evaluate(interpolate(x)) == x, and the coefficients match the oracle."""
    log = 15
    tree = backend.precompute_canonic_twiddles(log + 1)
    nat = np.arange(1 << log, dtype=np.uint32)
    br = np.array([int(bin(j)[2:].zfill(log)[::-1], 2) for j in range(1 << log)], np.uint32)
    col = BaseColumn.from_cpu(ctx, nat[br])
    backend.interpolate(col, tree)
    out = backend.evaluate(col, log, tree)
    np.testing.assert_array_equal(out.to_cpu(), nat[br])


def test_batched_uneven_sizes(ctx, backend):
    rng = np.random.default_rng(7)
    logs = [3, 12, 5, 12, 16, 3, 9, 14]
    tree = backend.precompute_canonic_twiddles(max(logs) + 2)
    tw, itw = O.canonic_twiddles(max(logs) + 2)
    evs = [rand_col(rng, lg) for lg in logs]
    cols = [BaseColumn.from_cpu(ctx, e) for e in evs]
    backend.interpolate_columns(cols, tree)
    coeffs = [O.interpolate(itw, e, lg) for e, lg in zip(evs, logs)]
    for c, e in zip(cols, coeffs):
        np.testing.assert_array_equal(c.to_cpu(), e)
    for blowup in (1, 2):
        outs = backend.evaluate_polynomials(cols, blowup, tree)
        for o, e, lg in zip(outs, coeffs, logs):
            np.testing.assert_array_equal(o.to_cpu(), O.evaluate(tw, e, lg + blowup))


def test_bit_reverse(ctx, backend):
    rng = np.random.default_rng(3)
    v = rand_col(rng, 11)
    col = BaseColumn.from_cpu(ctx, v)
    col.bit_reverse()
    br = np.array([int(bin(j)[2:].zfill(11)[::-1], 2) for j in range(1 << 11)])
    np.testing.assert_array_equal(col.to_cpu(), v[br])


def test_domain_not_in_tower_rejected(ctx, backend):
    from proving_amd import HipError
    tree = backend.precompute_canonic_twiddles(8)
    col = BaseColumn(ctx, 10)
    with pytest.raises(HipError, match="not a doubling"):
        backend.interpolate(col, tree)


@pytest.mark.parametrize("log", [1, 2, 5, 12, 17])
def test_split_at_mid(ctx, backend, log):
    """PolyOps::split_at_mid: halves of the coefficient vector, left in place."""
    rng = np.random.default_rng(40 + log)
    coeffs = rand_col(rng, log)
    poly = BaseColumn.from_cpu(ctx, coeffs)
    left, right = backend.split_at_mid(poly)
    assert left.log_size == right.log_size == log - 1
    np.testing.assert_array_equal(left.to_cpu(), coeffs[: 1 << (log - 1)])
    np.testing.assert_array_equal(right.to_cpu(), coeffs[1 << (log - 1):])


def test_split_at_mid_rejects_overlap(ctx, backend):
    from proving_amd import HipError
    col = BaseColumn(ctx, 6)
    with pytest.raises(HipError, match="overlap"):
        ctx.check(backend.lib.sh_split_at_mid(backend.h, col.ptr, 6, col.ptr, col.ptr + 4 * 40), "split_at_mid")
