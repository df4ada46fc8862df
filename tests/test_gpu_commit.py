"""GPU parity of the batched trace commitment (interpolate + LDE + lifted Merkle) at the
benchmark configurations, against the oracle, bit-exact.

* 64 columns x 2^14, seed 0, blowup 1: every LDE word and the root.
* 512 columns of mixed sizes (256 @ 2^12, 128 @ 2^16, 96 @ 2^18, 32 @ 2^20), seed 1: root of the
  lifted tree and of the mixed-degree tree (commit_on_layer), plus spot columns.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from proving_amd.backend import BaseColumn

pytestmark = pytest.mark.gpu


def run_commit(ctx, backend, logs, seed, blowup=1):
    host = O.splitmix64_columns(seed, logs)
    L = max(logs) + blowup
    tree = backend.precompute_canonic_twiddles(L)
    evals = [BaseColumn.from_cpu(ctx, c) for c in host]
    coeffs, ldes, layers, root = backend.commit_trace(evals, blowup, tree)
    return host, tree, evals, coeffs, ldes, layers, root


def test_commit_64x2p14_bit_exact(ctx, backend):
    logs = [14] * 64
    host, tree, evals, coeffs, ldes, layers, root = run_commit(ctx, backend, logs, 0)
    tw, itw = O.canonic_twiddles(15)
    exp_ldes = []
    for i, c in enumerate(host):
        co = O.interpolate(itw, c, 14)
        np.testing.assert_array_equal(coeffs[i].to_cpu(), co)
        lde = O.evaluate(tw, co, 15)
        np.testing.assert_array_equal(ldes[i].to_cpu(), lde)
        exp_ldes.append(lde)
    exp = O.merkle_commit_lifted(exp_ldes, 15)
    assert root == exp[0][0].tobytes()
    np.testing.assert_array_equal(layers[15].to_cpu(), exp[15])
    # one leaf hash per row across all 64 columns == the mixed-degree root here
    assert root == O.merkle_commit_mixed(exp_ldes)[0][0].tobytes()


def test_commit_mixed_512_columns_root(ctx, backend):
    logs = [12] * 256 + [16] * 128 + [18] * 96 + [20] * 32
    host, tree, evals, coeffs, ldes, layers, root = run_commit(ctx, backend, logs, 1)
    tw, itw = O.canonic_twiddles(21)
    exp_ldes = [O.evaluate(tw, O.interpolate(itw, c, lg), lg + 1) for c, lg in zip(host, logs)]
    for i in (0, 255, 256, 383, 384, 479, 480, 511):
        np.testing.assert_array_equal(ldes[i].to_cpu(), exp_ldes[i])
    exp = O.merkle_commit_lifted(exp_ldes, 21)
    assert root == exp[0][0].tobytes()
    # the mixed-degree tree that injects smaller columns at their matching layer
    exp_mixed = O.merkle_commit_mixed(exp_ldes)
    _, mroot = backend.commit_mixed(ldes)
    assert mroot == exp_mixed[0][0].tobytes()
    # the batched commit with the mixed-degree tree (FFT classes largest first, each column level
    # started when its class is ready): same LDEs, the oracle's mixed root and layers
    co2, ld2, lay2, mroot2 = backend.commit_trace(evals, 1, tree, mixed=True)
    assert mroot2 == exp_mixed[0][0].tobytes()
    for i in (0, 255, 256, 383, 384, 479, 480, 511):
        np.testing.assert_array_equal(ld2[i].to_cpu(), exp_ldes[i])
    for k in (21, 20, 19, 17, 13, 12, 6, 1):
        np.testing.assert_array_equal(lay2[k].to_cpu(), exp_mixed[k], err_msg=f"layer {k}")


def test_commit_repeatable(ctx, backend):
    """Repeated commits reuse scratch buffers; results must not change."""
    logs = [10] * 20 + [13] * 3
    host, tree, evals, coeffs, ldes, layers, root = run_commit(ctx, backend, logs, 5)
    for _ in range(3):
        _, _, _, r2 = backend.commit_trace(evals, 1, tree, coeffs=coeffs, ldes=ldes, layers=layers)
        assert r2 == root


def test_autotune_keeps_results(ctx, backend):
    """The pass shape changes only speed: commit before and after autotuning gives the same tree."""
    logs = [9] * 5 + [14] * 4 + [15] * 2
    host, tree, evals, coeffs, ldes, layers, root = run_commit(ctx, backend, logs, 9)
    backend.autotune_fft(logs, 1, tree)
    for m in (9, 10, 14, 15, 16):
        backend.set_fft_tuning(m, 11, 3, 3, 3)
    _, _, _, r1 = backend.commit_trace(evals, 1, tree, coeffs=coeffs, ldes=ldes, layers=layers)
    backend.autotune_fft(logs, 1, tree)
    _, _, _, r2 = backend.commit_trace(evals, 1, tree, coeffs=coeffs, ldes=ldes, layers=layers)
    assert r1 == root and r2 == root


@pytest.mark.parametrize("m,b,tune", [
    (13, 1, (9, 4, 2, 1)),    # last strided pass: 4 layers, 2 rows per lane
    (14, 1, (9, 5, 2, 2)),    # 5 layers, 4 rows
    (16, 2, (10, 6, 2, 3)),   # 6 layers, 8 rows, extension onto 4 subdomains
    (17, 1, (10, 7, 2, 4)),   # 7 layers, 16 rows
    (20, 1, (12, 8, 2, 3)),   # the 256 x 2^20 trace's shape
    (20, 1, (12, 8, 2, 2)),
    (19, 1, (11, 4, 2, 4)),   # two strided passes; the last one a single register stage
    (14, 3, (9, 5, 2, 2)),    # extension onto 8 subdomains
])
def test_fused_interpolate_extend(ctx, backend, m, b, tune):
    """The interpolation's last strided pass fused with the extension's first (sh_set_fft_fusion):
    every coefficient and extension word, and the root, equal the separate transforms' and the
    oracle's, for each fused pass shape (layers 4..8, 2..16 rows per lane, one register stage)."""
    saved = {k: backend.fft_tuning(k) for k in (m, m + b)}
    n = 3
    host = O.splitmix64_columns(40 + m, [m] * n)
    tree = backend.precompute_canonic_twiddles(m + b)
    evals = [BaseColumn.from_cpu(ctx, c) for c in host]
    try:
        for k in (m, m + b):
            backend.set_fft_tuning(k, *tune)
        out = {}
        for mode in (0, 1, 2):
            backend.set_fft_fusion(mode)
            co, ld, _, root = backend.commit_trace(evals, b, tree)
            out[mode] = ([c.to_cpu() for c in co], [x.to_cpu() for x in ld], root)
        tw, itw = O.canonic_twiddles(m + b)
        co_exp = O.interpolate(itw, host[1], m)
        np.testing.assert_array_equal(out[1][0][1], co_exp)
        np.testing.assert_array_equal(out[1][1][1], O.evaluate(tw, co_exp, m + b))
        for mode in (0, 2):
            for i in range(n):
                np.testing.assert_array_equal(out[mode][0][i], out[1][0][i], err_msg=f"coeffs mode {mode} col {i}")
                np.testing.assert_array_equal(out[mode][1][i], out[1][1][i], err_msg=f"lde mode {mode} col {i}")
            assert out[mode][2] == out[1][2]
    finally:
        backend.set_fft_fusion(1)
        for k, t in saved.items():
            backend.set_fft_tuning(k, *t)


def test_pool_reuses_freed_buffers(ctx, backend):
    """sh_free returns a block to the context pool; the next allocation of that size gets it back
    (no device sync), and work through reused blocks is unchanged."""
    from proving_amd.backend import BaseColumn as BC
    a = BC(ctx, 16)
    pa = a.ptr
    a.free()
    b = BC(ctx, 16)
    assert b.ptr == pa
    logs = [9] * 4 + [12] * 2
    host, tree, evals, coeffs, ldes, layers, root = run_commit(ctx, backend, logs, 4)
    coeffs.free()  # the commit's outputs: one pooled allocation each
    ldes.free()
    for _ in range(2):
        _, _, _, r2 = backend.commit_trace(evals, 1, tree)
        assert r2 == root
    ctx.check(ctx.lib.sh_pool_trim(ctx.h), "pool_trim")
    _, _, _, r3 = backend.commit_trace(evals, 1, tree)
    assert r3 == root


@pytest.mark.parametrize("off,n,val", [(0, 1 << 14, 0), (1, 1, 7), (3, 29, 0xAB), (5, 4096 + 11, 0xFF),
                                       (16, 16, 1), (12, 1 << 20, 0x5A)])
def test_memset_bytes(ctx, off, n, val):
    """sh_memset (the zero-fill of accumulators and counters): exactly bytes [off, off + n) take the
    value, at any alignment and length (16-byte stores over the aligned middle, bytes at the ends)."""
    rng = np.random.default_rng(off + n)
    host = rng.integers(0, 1 << 31, 1 << 19, dtype=np.uint32)
    col = BaseColumn.from_cpu(ctx, host)
    ctx.check(ctx.lib.sh_memset(ctx.h, ctypes.c_void_p(col.ptr + off), val, n), "sh_memset")
    exp = host.view(np.uint8).copy()
    exp[off:off + n] = val
    np.testing.assert_array_equal(col.to_cpu().view(np.uint8), exp)
    col.free()
