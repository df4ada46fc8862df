"""Graph-captured commitment: every replay equals the eager result, with eager work and new
input data between replays (graph replays after the first are where captured memsets have
gone wrong on this hardware; the captured commit uses none)."""
import numpy as np
import pytest

import oracle as O
from proving_amd.backend import BaseColumn, HashLayer
from proving_amd.workloads import splitmix64_columns

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager(ctx, backend):
    logs = [10] * 24 + [13] * 6 + [15] * 2
    b = 1
    L = max(logs) + b
    tree = backend.precompute_canonic_twiddles(L)
    host = splitmix64_columns(11, logs)
    evals = [BaseColumn.from_cpu(ctx, c) for c in host]
    coeffs = [BaseColumn(ctx, lg) for lg in logs]
    ldes = [BaseColumn(ctx, lg + b) for lg in logs]
    layers = [HashLayer(ctx, k) for k in range(L + 1)]
    _, _, _, eager_root = backend.commit_trace(evals, b, tree, coeffs, ldes, layers)
    g = backend.commit_trace_graph(evals, b, tree, coeffs, ldes, layers)
    # unrelated eager work on other buffers between replays
    other = [BaseColumn.from_cpu(ctx, c) for c in splitmix64_columns(12, [12] * 4)]
    for rep in range(4):
        layers[0].upload(np.zeros(32, np.uint8))
        g.launch()
        assert g.root() == eager_root, rep
        backend.commit_trace(other, b, tree)
        ctx.sync()
    # new input data: the graph reads the same buffers, so it commits the new columns
    host2 = splitmix64_columns(13, logs)
    for e, c in zip(evals, host2):
        e.upload(c)
    g.launch()
    r2 = g.root()
    tw, itw = O.canonic_twiddles(L)
    exp_ldes = [O.evaluate(tw, O.interpolate(itw, c, lg), lg + b) for c, lg in zip(host2, logs)]
    assert r2 == O.merkle_commit_lifted(exp_ldes, L)[0][0].tobytes()
    for lde, exp in zip(ldes[::7], exp_ldes[::7]):
        np.testing.assert_array_equal(lde.to_cpu(), exp)
    g.free()


def test_mixed_graph_replays_match_eager(ctx, backend):
    """The captured batched commit with the mixed-degree tree: replays equal the eager commit,
    and with new inputs the oracle's mixed-degree root."""
    logs = [10] * 24 + [13] * 6 + [15] * 2 + [9] * 3
    b = 1
    L = max(logs) + b
    tree = backend.precompute_canonic_twiddles(L)
    evals = [BaseColumn.from_cpu(ctx, c) for c in splitmix64_columns(21, logs)]
    coeffs = [BaseColumn(ctx, lg) for lg in logs]
    ldes = [BaseColumn(ctx, lg + b) for lg in logs]
    layers = [HashLayer(ctx, k) for k in range(L + 1)]
    _, _, _, eager_root = backend.commit_trace(evals, b, tree, coeffs, ldes, layers, mixed=True)
    g = backend.commit_trace_graph(evals, b, tree, coeffs, ldes, layers, mixed=True)
    other = [BaseColumn.from_cpu(ctx, c) for c in splitmix64_columns(22, [12] * 4)]
    for rep in range(4):
        layers[0].upload(np.zeros(32, np.uint8))
        g.launch()
        assert g.root() == eager_root, rep
        backend.commit_trace(other, b, tree, mixed=True)
        ctx.sync()
    host2 = splitmix64_columns(23, logs)
    for e, c in zip(evals, host2):
        e.upload(c)
    g.launch()
    tw, itw = O.canonic_twiddles(L)
    exp_ldes = [O.evaluate(tw, O.interpolate(itw, c, lg), lg + b) for c, lg in zip(host2, logs)]
    assert g.root() == O.merkle_commit_mixed(exp_ldes)[0][0].tobytes()
    g.free()


def test_captured_entry_point_sequence_replays(ctx, backend):
    """sh_graph_warmup/capture around a sequence of entry points that allocate (interpolate into
    pooled copies, evaluate into a new set): the capture hands out the warm-up's blocks in the
    same order, and every replay, with eager work and new inputs between replays, equals the
    oracle."""
    from proving_amd.backend import Graph
    logs = [9] * 5 + [12] * 3
    L = max(logs) + 1
    tree = backend.precompute_canonic_twiddles(L)
    tw, itw = O.canonic_twiddles(L)
    evals = [BaseColumn.from_cpu(ctx, c) for c in splitmix64_columns(31, logs)]

    def work():
        coeffs = []
        for e in evals:
            c = BaseColumn(ctx, e.log_size)
            c.copy_from(e)
            coeffs.append(c)
        backend.interpolate_columns(coeffs, tree)
        tmp = BaseColumn(ctx, 11)   # freed inside the sequence: its block is reused below
        tmp.free()
        return coeffs, backend.evaluate_polynomials(coeffs, 1, tree)

    g = Graph(ctx)
    with g.warmup():
        coeffs, ldes = work()
    for c in coeffs:
        c.free()
    ldes.free()
    with g.capture():
        coeffs, ldes = work()
    other = [BaseColumn.from_cpu(ctx, c) for c in splitmix64_columns(32, [12] * 4)]
    for rep, seed in enumerate((33, 33, 34)):
        host = splitmix64_columns(seed, logs)
        for e, h in zip(evals, host):
            e.upload(h)
        g.launch()
        for k in (0, 4, 5, 7):
            ec = O.interpolate(itw, host[k], logs[k])
            np.testing.assert_array_equal(coeffs[k].to_cpu(), ec, err_msg=f"replay {rep} column {k}")
            np.testing.assert_array_equal(ldes[k].to_cpu(), O.evaluate(tw, ec, logs[k] + 1))
        backend.interpolate_columns(other, tree)
        ctx.sync()
    for c in coeffs:
        c.free()
    ldes.free()
    g.free()
