"""GPU parity: lifted Merkle leaves and layers, mixed-degree commit_on_layer, against the
oracle and the golden (hashlib, reference-test) vectors.

Mirrors crates/stwo/src/prover/backend/simd/blake2s_lifted.rs:367-428 and
crates/stwo/src/prover/vcs_lifted/prover.rs:184-312.
"""
import numpy as np
import pytest

import oracle as O
from proving_amd.backend import BaseColumn

pytestmark = pytest.mark.gpu
P = O.P


def lcg_cols(seed, logs):
    x = seed
    cols = []
    for lg in logs:
        c = []
        for _ in range(1 << lg):
            x = (x * 6364136223846793005 + 1442695040888963407) & ((1 << 64) - 1)
            c.append((x >> 33) % P)
        cols.append(np.array(c, np.uint32))
    return cols


def test_lifted_leaves_reference_kat(ctx, backend, golden):
    k = golden["reference_kats"]["lifted_leaves_small"]
    cols = [BaseColumn.from_cpu(ctx, np.arange(1 << lg, dtype=np.uint32)) for lg in k["logs"]]
    leaves = backend.build_leaves(cols, 4).to_cpu()
    assert leaves[0].tobytes().hex() == k["leaf0"]
    assert leaves[4].tobytes().hex() == k["leaf4"]
    assert leaves[15].tobytes().hex() == k["leaf15"]
    assert [l.tobytes().hex() for l in leaves] == k["all_leaves"]
    layers, root = backend.commit_lifted(cols, 4)
    assert root.hex() == k["root"]


def test_trees_match_golden(ctx, backend, golden):
    for case in golden["hashlib"]["trees"]:
        host = lcg_cols(case["seed"], case["logs"])
        cols = [BaseColumn.from_cpu(ctx, c) for c in host]
        layers, root = backend.commit_lifted(cols, case["lifting_log"])
        assert root.hex() == case["lifted_root"], case["logs"]
        leaves = layers[-1].to_cpu()
        for i, h in case["leaves_sample"].items():
            assert leaves[int(i)].tobytes().hex() == h
        _, mroot = backend.commit_mixed(cols)
        assert mroot.hex() == case["mixed_root"], case["logs"]


@pytest.mark.parametrize("logs,L", [([4] * 16, 4), ([1] * 3, 1), ([5] * 40 + [7] * 17 + [9] * 5, 11),
                                    ([2, 3, 3, 6, 6, 6, 10], 10), ([12] * 64, 12), ([8] * 1, 13),
                                    # group ends on a block boundary with more groups to come; a carried
                                    # partial block completed exactly at a group's end; long groups
                                    ([3] * 16 + [5] * 16 + [6] * 20, 6), ([3] * 10 + [4] * 6 + [5] * 3, 7),
                                    ([6] * 100 + [8] * 33 + [9] * 15, 9)])
def test_lifted_tree_matches_oracle(ctx, backend, logs, L):
    rng = np.random.default_rng(len(logs) * 31 + L)
    host = [rng.integers(0, P, 1 << lg, dtype=np.uint32) for lg in logs]
    order = sorted(range(len(logs)), key=lambda i: logs[i])
    cols = [BaseColumn.from_cpu(ctx, c) for c in host]
    layers, root = backend.commit_lifted(cols, L)
    exp = O.merkle_commit_lifted(host, L)
    assert root == exp[0][0].tobytes()
    for k in range(L + 1):
        np.testing.assert_array_equal(layers[k].to_cpu(), exp[k])
    leaves = backend.build_leaves([cols[i] for i in order], L).to_cpu()
    np.testing.assert_array_equal(leaves, exp[L])


def test_next_layer(ctx, backend):
    from proving_amd.backend import HashLayer
    rng = np.random.default_rng(11)
    for log in (1, 2, 9, 12):
        prev = rng.integers(0, 256, (1 << log, 32), dtype=np.uint8)
        hl = HashLayer(ctx, log)
        hl.upload(prev)
        out = backend.build_next_layer(hl).to_cpu()
        np.testing.assert_array_equal(out, O.merkle_next_layer(prev))


def test_empty_tree(ctx, backend):
    import hashlib
    layers, root = backend.commit_lifted([], 0)
    assert root == hashlib.blake2s(b"").digest()


def test_commit_on_layer_cases(ctx, backend):
    rng = np.random.default_rng(5)
    from proving_amd.backend import HashLayer
    for log, ncols, with_prev in [(3, 0, False), (3, 0, True), (6, 5, True), (6, 16, True),
                                  (6, 17, False), (9, 33, True), (2, 1, True)]:
        host = [rng.integers(0, P, 1 << log, dtype=np.uint32) for _ in range(ncols)]
        cols = [BaseColumn.from_cpu(ctx, c) for c in host]
        prev = None
        prev_host = None
        if with_prev:
            prev_host = rng.integers(0, 256, (2 << log, 32), dtype=np.uint8)
            prev = HashLayer(ctx, log + 1)
            prev.upload(prev_host)
        out = backend.commit_on_layer(log, prev, cols).to_cpu()
        np.testing.assert_array_equal(out, O.commit_on_layer(log, prev_host, host))


def test_mixed_tree_matches_oracle(ctx, backend):
    rng = np.random.default_rng(8)
    logs = [12, 12, 16, 16, 16, 18, 20, 13, 12]
    host = [rng.integers(0, P, 1 << lg, dtype=np.uint32) for lg in logs]
    cols = [BaseColumn.from_cpu(ctx, c) for c in host]
    layers, root = backend.commit_mixed(cols)
    exp = O.merkle_commit_mixed(host)
    assert root == exp[0][0].tobytes()
    for k in (0, 5, 12, 13, 17, 20):
        np.testing.assert_array_equal(layers[k].to_cpu(), exp[k])


def test_mixed_tree_long_gaps_every_layer(ctx, backend):
    """Column levels separated by 4-5 column-free levels (more than one launch reduces), up to 40
    columns on one level, a single column at the top level: every layer equals the oracle's."""
    rng = np.random.default_rng(9)
    logs = [20] * 3 + [14] * 40 + [9] * 17 + [3] * 2 + [1]
    host = [rng.integers(0, P, 1 << lg, dtype=np.uint32) for lg in logs]
    order = rng.permutation(len(logs))  # the tree sorts the columns itself
    cols = [BaseColumn.from_cpu(ctx, host[i]) for i in order]
    layers, root = backend.commit_mixed(cols)
    exp = O.merkle_commit_mixed([host[i] for i in order])
    assert root == exp[0][0].tobytes()
    for k in range(21):
        np.testing.assert_array_equal(layers[k].to_cpu(), exp[k], err_msg=f"layer {k}")
