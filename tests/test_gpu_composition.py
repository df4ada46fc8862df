"""GPU parity of the wide-Fibonacci composition kernel (ComponentProver::
evaluate_constraint_quotients_on_domain for WideFibonacciEval) against the oracle, bit-exact."""
import numpy as np
import pytest

import oracle as O
from oracle import protocol as PR
from proving_amd.backend import BaseColumn, SecureColumnByCoords
from proving_amd.workloads import wide_fib_trace

pytestmark = pytest.mark.gpu
P = O.P


def powers_for(m, seed):
    rng = np.random.default_rng(seed)
    alpha = [int(v) for v in rng.integers(0, P, 4)]
    pw = [[1, 0, 0, 0]]
    for _ in range(m - 1):
        pw.append(PR.qmul(pw[-1], alpha))
    return pw[::-1]


@pytest.mark.parametrize("n_cols,log_n,extra,expansion", [
    (3, 3, 1, 0), (5, 4, 1, 0), (16, 6, 1, 0), (64, 14, 1, 0), (10, 5, 2, 0), (12, 6, 1, 1), (256, 10, 1, 0)])
def test_composition_matches_oracle(ctx, backend, n_cols, log_n, extra, expansion):
    trace = wide_fib_trace(n_cols, log_n, 30 + n_cols)
    E = log_n + extra
    Lc = E + expansion  # committed evaluation size
    tw, itw = O.canonic_twiddles(Lc)
    ldes = [O.evaluate(tw, O.interpolate(itw, c, log_n), Lc) for c in trace]
    powers = powers_for(n_cols - 2, n_cols)
    exp = PR.wide_fib_composition(ldes, log_n, E, expansion, powers)
    cols = [BaseColumn.from_cpu(ctx, l) for l in ldes]
    acc = SecureColumnByCoords.from_cpu(ctx, np.zeros((4, 1 << E), np.uint32))
    backend.wide_fibonacci_composition(cols, log_n, E, powers, acc, log_expansion=expansion)
    np.testing.assert_array_equal(acc.to_cpu(), exp)
    # accumulates: a second call adds the same again
    backend.wide_fibonacci_composition(cols, log_n, E, powers, acc, log_expansion=expansion)
    np.testing.assert_array_equal(acc.to_cpu(), (exp.astype(np.uint64) * 2 % P).astype(np.uint32))


def test_composition_rejects_bad_shapes(ctx, backend):
    from proving_amd import HipError
    cols = [BaseColumn(ctx, 4) for _ in range(2)]
    acc = SecureColumnByCoords(ctx, 4)
    with pytest.raises(HipError, match="3..4096"):
        backend.wide_fibonacci_composition(cols, 3, 4, [[1, 0, 0, 0]], acc)


@pytest.mark.parametrize("n_cols,log_n", [(2, 3), (3, 0), (17, 9), (256, 12)])
def test_trace_generation_matches_host(ctx, backend, n_cols, log_n):
    exp = wide_fib_trace(n_cols, log_n, 9)
    c0, c1 = BaseColumn.from_cpu(ctx, exp[0]), BaseColumn.from_cpu(ctx, exp[1])
    cols = backend.wide_fibonacci_trace(c0, c1, n_cols)
    for k in (0, 1, n_cols // 2, n_cols - 1):
        np.testing.assert_array_equal(cols[k].to_cpu(), exp[k])
