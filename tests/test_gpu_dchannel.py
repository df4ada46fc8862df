"""The device-resident Blake2sChannel (dchannel.hip) and the entry points that read their
arguments from device memory, against the oracle, bit-exact: every mix, draw, proof of work,
power table, sample and DEEP quotient a proof takes from the device channel equals what the
oracle's channel and drivers give for the same state (crates/stwo/src/core/channel/
blake2s.rs:131-240 are the reference's channel tests this mirrors)."""
import ctypes

import numpy as np
import pytest

import oracle as O
from oracle import protocol as PR
from proving_amd.backend import (BaseColumn, Channel, ColumnRefs, DeviceBuffer, DeviceChannel, FriParams,
                                 SecureColumnByCoords, _ptr_array, fri_prove, fri_prove_dch, fri_quotients_dpoint,
                                 tree_decommit)

pytestmark = pytest.mark.gpu
P = O.P


def dev_words(ctx, arr):
    a = np.ascontiguousarray(np.asarray(arr, dtype=np.uint32).reshape(-1))
    buf = DeviceBuffer(ctx, 4 * max(1, len(a)))
    if len(a):
        buf.upload(a)
    return buf


def started_oracle_channel(seed):
    """An oracle channel in a state with a nonzero digest and n_draws."""
    rng = np.random.default_rng(seed)
    oc = O.Channel()
    oc.mix_u32s(rng.integers(0, 1 << 32, 7, dtype=np.uint64).astype(np.uint32).tolist())
    for _ in range(seed % 3 + 1):
        oc.draw_u32s()
    return oc


def device_from(backend, oc):
    hc = Channel(backend.lib)
    ctypes.memmove(hc.s.digest, oc.digest, 32)
    hc.s.n_draws = oc.n_draws
    return DeviceChannel(backend, hc)


def random_qm31_point(seed):
    oc = O.Channel()
    oc.mix_u64(seed)
    return [int(x) for x in oc.random_point()]


@pytest.mark.parametrize("n", [0, 1, 7, 8, 9, 15, 16, 17, 24, 100, 4 * 264])
def test_mix_words(ctx, backend, n):
    """mix_u32s over n words: one block (n <= 8), exact multiples of the block, and the 264
    sampled values of the 256-column proof."""
    oc = started_oracle_channel(n)
    dch = device_from(backend, oc)
    w = np.random.default_rng(n).integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    dch.mix_words(dev_words(ctx, w).ptr, n)
    oc.mix_u32s(w.tolist())
    assert dch.state() == (oc.digest, oc.n_draws)


def test_draws_and_random_point(ctx, backend):
    oc = started_oracle_channel(5)
    dch = device_from(backend, oc)
    felts = DeviceBuffer(ctx, 16 * 5)
    dch.draw_secure_felts(5, felts.ptr)
    pt = DeviceBuffer(ctx, 32)
    dch.draw_random_point(pt.ptr)
    exp = np.array([oc.draw_secure_felt() for _ in range(5)])
    exp_pt = oc.random_point()
    np.testing.assert_array_equal(felts.download(np.uint32, 20).reshape(5, 4), exp)
    np.testing.assert_array_equal(pt.download(np.uint32, 8), exp_pt)
    assert dch.state() == (oc.digest, oc.n_draws)
    # the point is on the circle: x^2 + y^2 = 1 over QM31
    x, y = [int(v) for v in exp_pt[:4]], [int(v) for v in exp_pt[4:]]
    assert PR.qadd(PR.qmul(x, x), PR.qmul(y, y)) == [1, 0, 0, 0]


@pytest.mark.parametrize("n,rev", [(1, 0), (5, 0), (64, 1), (65, 0), (254, 1), (1000, 0)])
def test_secure_powers_dev(ctx, backend, n, rev):
    felt = [int(x) for x in np.random.default_rng(n).integers(0, P, 4)]
    src = dev_words(ctx, felt)
    out = DeviceBuffer(ctx, 16 * n)
    backend.secure_powers_dev(src.ptr, n, rev, out.ptr)
    exp, p = [], [1, 0, 0, 0]
    for _ in range(n):
        exp.append(p)
        p = PR.qmul(p, felt)
    if rev:
        exp = exp[::-1]
    np.testing.assert_array_equal(out.download(np.uint32, 4 * n).reshape(n, 4), np.array(exp, np.uint32))


@pytest.mark.parametrize("bits", [1, 8, 16, 20])
def test_grind_dchannel(ctx, backend, bits):
    """GrindOps::grind then mix_u64(nonce): the smallest valid nonce, as the oracle finds it."""
    oc = started_oracle_channel(bits)
    dch = device_from(backend, oc)
    ctl = DeviceBuffer(ctx, 16)
    dch.grind(bits, ctl.ptr)
    nonce = O.grind(oc.digest, bits)
    assert O.verify_pow_nonce(oc.digest, bits, nonce)
    oc.mix_u64(nonce)
    c = ctl.download(np.uint32, 2)
    assert int(c[0]) | (int(c[1]) << 32) == nonce
    assert dch.state() == (oc.digest, oc.n_draws)


def test_eval_at_dpoint(ctx, backend):
    """compute_samples at a device point: columns of several sizes, each at the point doubled
    its own number of times."""
    rng = np.random.default_rng(3)
    logs = [3, 6, 6, 9, 12, 12, 12]
    dbls = [0, 2, 0, 5, 1, 0, 3]
    polys = [rng.integers(0, P, 1 << lg, dtype=np.uint32) for lg in logs]
    cols = [BaseColumn.from_cpu(ctx, p) for p in polys]
    pt = random_qm31_point(11)
    d_pt = dev_words(ctx, pt)
    vals = DeviceBuffer(ctx, 16 * len(cols))
    backend.eval_at_dpoint(cols, dbls, d_pt.ptr, vals.ptr)
    got = vals.download(np.uint32, 4 * len(cols)).reshape(-1, 4)
    for i, (p, d) in enumerate(zip(polys, dbls)):
        np.testing.assert_array_equal(got[i], O.eval_at_point(p, PR._repeated_double(pt, d)))


@pytest.mark.parametrize("logs,L", [([9] * 6, 9), ([5, 9, 7, 9, 5, 8, 6], 9), ([12] * 40 + [11] * 9, 12)])
def test_fri_quotients_dpoint(ctx, backend, logs, L):
    """The device-point DEEP quotient equals the oracle's compute_fri_quotients with one sample
    per column, all at one point: one size class, several (lifted and accumulated), and many
    columns of two sizes."""
    b = 1
    rng = np.random.default_rng(len(logs) + L)
    evals = [rng.integers(0, P, 1 << lg, dtype=np.uint32) for lg in logs]
    pt = random_qm31_point(L)
    values = [[int(x) for x in rng.integers(0, P, 4)] for _ in logs]
    rc = [int(x) for x in rng.integers(0, P, 4)]
    tw, itw = O.canonic_twiddles(L + 1)
    exp = PR.fri_quotients(evals, [[(pt, v)] for v in values], rc, L, b, tw, itw)
    tree = backend.precompute_canonic_twiddles(L + 1)
    cols = [BaseColumn.from_cpu(ctx, e) for e in evals]
    d_pt, d_vals, d_rc = dev_words(ctx, pt), dev_words(ctx, values), dev_words(ctx, rc)
    out = fri_quotients_dpoint(backend, tree, cols, d_pt.ptr, d_vals.ptr, d_rc.ptr, L, b)
    np.testing.assert_array_equal(out.to_cpu(), np.asarray(exp, np.uint32).reshape(4, -1))


def test_fri_prove_dch_matches_host_channel(ctx, backend):
    """FRI from a device channel: the same proof, queries and final channel as the host-channel
    driver, the extra words come back with the commit phase, and the trees opened in the same
    gather equal sh_tree_decommit's openings (a preprocessed tree of another height included)."""
    log = 10
    fri = FriParams(6, 1, 1, 20, 1)
    rng = np.random.default_rng(8)
    # a low-degree QM31 evaluation: 4 coordinate polynomials of degree < 2^(log - 1)
    tw, _ = O.canonic_twiddles(log + 1)
    col = np.stack([O.evaluate(tw, np.concatenate([rng.integers(0, P, 1 << (log - 1), dtype=np.uint32),
                                                   np.zeros(1 << (log - 1), np.uint32)]), log)
                    for _ in range(4)])
    tree = backend.precompute_canonic_twiddles(log + 1)
    q = SecureColumnByCoords.from_cpu(ctx, col)
    oc = started_oracle_channel(9)
    host = Channel(backend.lib)
    ctypes.memmove(host.s.digest, oc.digest, 32)
    host.s.n_draws = oc.n_draws
    dch = DeviceChannel(backend, host)
    felts_h, queries_h = fri_prove(backend, tree, host, fri, q, as_array=True)
    extra = dev_words(ctx, np.arange(37, dtype=np.uint32) * 3 + 1)
    # trees: a lifted tree at the FRI column's height with mixed column sizes, a preprocessed one
    # of a lower height, and an empty one
    trees = []
    for logs, pre in (([log, log - 2, log, log - 4], False), ([log - 3, log - 3], True), ([], True)):
        cols = [BaseColumn.from_cpu(ctx, rng.integers(0, P, 1 << lg, dtype=np.uint32)) for lg in logs]
        L = max(logs) if logs else 0
        layers, _ = backend.commit_lifted(sorted(cols, key=lambda c: c.log_size) if cols else [], L)
        trees.append((cols, layers, pre))
    out = Channel(backend.lib)
    felts_d, queries_d, got_extra, openings = fri_prove_dch(backend, tree, dch, out, fri, q, extra=(extra.ptr, 37),
                                                            trees=trees)
    assert felts_d.tolist() == felts_h.tolist()
    assert queries_d == queries_h
    assert (bytes(out.s.digest), out.s.n_draws) == (bytes(host.s.digest), host.s.n_draws)
    np.testing.assert_array_equal(got_extra, np.arange(37, dtype=np.uint32) * 3 + 1)
    for (cols, layers, pre), (vals, wit) in zip(trees, openings):
        L = len(layers) - 1
        pos = PR.pp_query_positions(queries_d, log, L) if pre else queries_d
        if not cols:
            assert vals.size == 0 and len(wit) == 0
            continue
        ev, ew = tree_decommit(backend, cols, layers, pos)
        np.testing.assert_array_equal(vals, ev)
        np.testing.assert_array_equal(wit, ew)


def test_composition_dev_matches_host_powers(ctx, backend):
    rng = np.random.default_rng(4)
    n_cols, log_n = 10, 7
    eval_log = log_n + 1
    cols = [BaseColumn.from_cpu(ctx, rng.integers(0, P, 1 << eval_log, dtype=np.uint32)) for _ in range(n_cols)]
    powers = rng.integers(0, P, (n_cols - 2, 4), dtype=np.uint32)
    exp = PR.wide_fib_composition([c.to_cpu() for c in cols], log_n, eval_log, 0, powers.tolist())
    acc = SecureColumnByCoords(ctx, eval_log)
    for c in acc.columns:
        ctx.check(ctx.lib.sh_memset(ctx.h, c.ptr, 0, c.nbytes), "memset")
    d_pw = dev_words(ctx, powers)
    backend.wide_fibonacci_composition_dev(cols, log_n, eval_log, d_pw.ptr, acc, 0)
    np.testing.assert_array_equal(acc.to_cpu(), exp)


@pytest.mark.parametrize("n_cols,log_n", [(3, 4), (12, 8), (300, 20)])
def test_wide_fib_oods_check_values(ctx, backend, n_cols, log_n):
    """The device sanity check's two values equal the oracle's composition rebuilt from its
    samples and the oracle's constraint evaluation at the point, on random samples (which
    disagree, so the flag is set); made to agree, the flag clears."""
    rng = np.random.default_rng(n_cols)
    vals = rng.integers(0, P, (n_cols + 8, 4), dtype=np.uint32)
    alpha = [int(x) for x in rng.integers(0, P, 4)]
    z = random_qm31_point(n_cols)
    m = n_cols - 2
    pw, p = [], [1, 0, 0, 0]
    for _ in range(m):
        pw.append(p)
        p = PR.qmul(p, alpha)
    rhs = PR.wide_fib_composition_at_point([list(map(int, v)) for v in vals[:n_cols]], alpha, z, log_n)
    comp = [list(map(int, v)) for v in vals[n_cols:]]
    x = PR._repeated_double(z, log_n - 1)[:4]
    lhs = PR.qadd(PR.from_partial_evals(comp[:4]), PR.qmul(x, PR.from_partial_evals(comp[4:])))
    d_pw = dev_words(ctx, np.array(pw[::-1], np.uint32))
    d_z, out = dev_words(ctx, z), DeviceBuffer(ctx, 64)
    d_vals = dev_words(ctx, vals)
    backend.wide_fibonacci_oods_check(d_vals.ptr, n_cols, d_pw.ptr, d_z.ptr, log_n, out.ptr)
    got = out.download(np.uint32, 9)
    assert got[1:5].tolist() == [int(v) for v in lhs] and got[5:9].tolist() == [int(v) for v in rhs]
    assert got[0] == (1 if list(lhs) != list(rhs) else 0)
    # shift the first composition coordinate so that lhs == rhs: the flag clears
    fix = PR.qsub(rhs, lhs)
    vals[n_cols] = np.array(PR.qadd([int(v) for v in vals[n_cols]], fix), np.uint32)
    d_vals = dev_words(ctx, vals)
    backend.wide_fibonacci_oods_check(d_vals.ptr, n_cols, d_pw.ptr, d_z.ptr, log_n, out.ptr)
    got = out.download(np.uint32, 9)
    assert got[0] == 0 and got[1:5].tolist() == got[5:9].tolist()
