"""Parity of the C++ protocol drivers (stwo_prover.h) with the oracle's restatement: the Blake2s
channel, lifted-Merkle decommitment, the DEEP-quotient driver and the FRI prover, byte-exact;
proofs from the HIP path are accepted by the oracle's independent FRI verifier.
"""
import numpy as np
import pytest

import oracle as O
from oracle import protocol as PR
from proving_amd.backend import (BaseColumn, Channel, FriParams, SecureColumnByCoords, fri_prove, fri_quotients,
                                 gather, merkle_decommit)

pytestmark = pytest.mark.gpu
P = O.P


def test_channel_matches_oracle():
    a, b = Channel(), O.Channel()
    ops = [("u32s", [1, 2, 3]), ("u32s", []), ("u64", 0x1234567890abcdef), ("felts", [[1, 2, 3, 4], [5, 6, 7, 8]]),
           ("root", bytes(range(32)))]
    for op, arg in ops:
        if op == "u32s":
            a.mix_u32s(arg)
            b.mix_u32s(arg)
        elif op == "u64":
            a.mix_u64(arg)
            b.mix_u64(arg)
        elif op == "felts":
            a.mix_felts(arg)
            b.mix_felts(np.array(arg, np.uint32))
        else:
            a.mix_root(arg)
            b.mix_hash(arg)
        assert a.digest == b.digest
        assert list(a.draw_secure_felt()) == list(b.draw_secure_felt())
        assert list(a.draw_random_point()) == list(b.random_point())
    qa = a.draw_queries(12, 19)
    raw = []
    while len(raw) < 19:
        for w in b.draw_u32s():
            if len(raw) < 19:
                raw.append(int(w) & 0xfff)
    assert qa == raw


def test_merkle_decommit_matches_oracle(ctx, backend):
    rng = np.random.default_rng(11)
    host = [rng.integers(0, P, 1 << 10, dtype=np.uint32) for _ in range(4)]
    cols = [BaseColumn.from_cpu(ctx, c) for c in host]
    layers, root = backend.commit_lifted(cols, 10)
    exp = O.merkle_commit_lifted(host, 10)
    assert root == exp[0][0].tobytes()
    for positions in ([0], [1023], [0, 1], [3, 4, 5, 100, 101, 511, 512, 1000], list(range(0, 1024, 7)), []):
        assert merkle_decommit(backend, layers, positions) == PR.merkle_decommit(exp, positions)


def test_gather(ctx, backend):
    rng = np.random.default_rng(12)
    host = [rng.integers(0, P, 1 << lg, dtype=np.uint32) for lg in (3, 9, 14)]
    cols = [BaseColumn.from_cpu(ctx, c) for c in host]
    picks = [(0, 7), (1, 0), (2, 16383), (1, 300), (0, 0), (2, 5)]
    got = gather(backend, [cols[c] for c, _ in picks], [r for _, r in picks])
    assert got == [int(host[c][r]) for c, r in picks]


def lifted_case(seed, logs, b, n_points=2):
    rng = np.random.default_rng(seed)
    L = max(logs) + b
    tw, itw = O.canonic_twiddles(L)
    coeffs = [rng.integers(0, P, 1 << lg, dtype=np.uint32) for lg in logs]
    evals = [O.evaluate(tw, c, lg + b) for c, lg in zip(coeffs, logs)]
    ch = O.Channel()
    ch.mix_u64(seed)
    pts = [[int(v) for v in ch.random_point()] for _ in range(n_points)]
    samples = []
    for i, (c, lg) in enumerate(zip(coeffs, logs)):
        m = (0, 1, 2)[i % 3] if n_points >= 2 else 1
        s = []
        for k in range(m):
            p = pts[k]
            for _ in range(L - (lg + b)):
                x, y = p[:4], p[4:]
                xx = PR.qmul(x, x)
                p = PR.qsub(PR.qadd(xx, xx), [1, 0, 0, 0]) + PR.qmulm(PR.qmul(x, y), 2)
            s.append((pts[k], [int(v) for v in O.eval_at_point(c, p)]))
        samples.append(s)
    return L, tw, itw, evals, samples


@pytest.mark.parametrize("logs,b", [([5, 8, 6, 8, 4, 7], 1), ([3, 3, 9], 2), ([10] * 5, 1), ([6, 12, 9, 11, 12], 1)])
def test_fri_quotients_match_oracle(ctx, backend, logs, b):
    L, tw, itw, evals, samples = lifted_case(len(logs) + b, logs, b)
    rc = [77, 88, 99, 1234567]
    exp = PR.fri_quotients(evals, samples, rc, L, b, tw, itw)
    tree = backend.precompute_canonic_twiddles(L)
    cols = [BaseColumn.from_cpu(ctx, e) for e in evals]
    out = fri_quotients(backend, tree, cols, samples, rc, L, b)
    np.testing.assert_array_equal(out.to_cpu(), exp)


def test_fri_quotients_large_rows_match_verifier_formula(ctx, backend):
    """At a size the oracle driver would take long on, check rows against fri_answers."""
    logs, b = [12] * 8 + [16] * 4 + [18] * 2, 1
    L, tw, itw, evals, samples = lifted_case(3, logs, b)
    tree = backend.precompute_canonic_twiddles(L)
    cols = [BaseColumn.from_cpu(ctx, e) for e in evals]
    rc = [5, 4, 3, 2]
    out = fri_quotients(backend, tree, cols, samples, rc, L, b).to_cpu()
    positions = [0, 1, 2, 3, 1000, 77777, (1 << L) - 2, (1 << L) - 1]
    qv = [[PR.lifted_value(e, L, p) for p in positions] for e in evals]
    rows = PR.fri_answers([lg + b for lg in logs], samples, rc, positions, qv, L)
    for k, p in enumerate(positions):
        assert rows[k] == [int(out[c][p]) for c in range(4)]


def params_of(d):
    return FriParams(d["pow_bits"], d["log_blowup_factor"], d["log_last_layer_degree_bound"], d["n_queries"],
                     d["fold_step"])


@pytest.mark.parametrize("pdict,logs", [
    ({"pow_bits": 10, "log_blowup_factor": 1, "log_last_layer_degree_bound": 0, "n_queries": 3, "fold_step": 1},
     [5, 8, 6, 8, 4, 7]),
    ({"pow_bits": 6, "log_blowup_factor": 2, "log_last_layer_degree_bound": 2, "n_queries": 30, "fold_step": 1},
     [3, 3, 9]),
    ({"pow_bits": 12, "log_blowup_factor": 1, "log_last_layer_degree_bound": 3, "n_queries": 70, "fold_step": 1},
     [6, 12, 9, 11, 12]),
    # fold_step > 1: packed leaves, multi-step folds, a shorter last fold
    ({"pow_bits": 0, "log_blowup_factor": 2, "log_last_layer_degree_bound": 2, "n_queries": 3, "fold_step": 2},
     [8, 6, 8]),
    ({"pow_bits": 6, "log_blowup_factor": 1, "log_last_layer_degree_bound": 0, "n_queries": 40, "fold_step": 3},
     [6, 12, 9, 11, 12]),
    ({"pow_bits": 4, "log_blowup_factor": 2, "log_last_layer_degree_bound": 1, "n_queries": 20, "fold_step": 4},
     [10, 7, 10]),
])
def test_fri_prove_matches_oracle_and_verifies(ctx, backend, pdict, logs):
    b = pdict["log_blowup_factor"]
    L, tw, itw, evals, samples = lifted_case(len(logs), logs, b)
    rc = [31, 41, 59, 26]
    quot = PR.fri_quotients(evals, samples, rc, L, b, tw, itw)
    ch_o = O.Channel()
    ch_o.mix_u64(7)
    proof, queries = PR.fri_prove(quot, L, pdict, ch_o)
    felts_o = PR.serialize_fri_proof(proof)

    tree = backend.precompute_canonic_twiddles(L)
    col = SecureColumnByCoords(ctx, L)
    for c in range(4):
        col.columns[c].upload(np.ascontiguousarray(quot[c]))
    ch = Channel()
    ch.mix_u64(7)
    felts, qpos = fri_prove(backend, tree, ch, params_of(pdict), col)
    assert qpos == queries
    assert ch.digest == ch_o.digest
    assert felts == felts_o

    def fle(qs):
        qv = [[PR.lifted_value(e, L, p) for p in qs] for e in evals]
        return PR.fri_answers([lg + b for lg in logs], samples, rc, qs, qv, L)

    chv = O.Channel()
    chv.mix_u64(7)
    assert PR.fri_verify(PR.deserialize_fri_proof(felts), pdict, chv, L - b, fle) == qpos


def test_fri_prove_rejects_high_degree(ctx, backend):
    from proving_amd import HipError
    rng = np.random.default_rng(1)
    col = SecureColumnByCoords(ctx, 8)
    for c in range(4):
        col.columns[c].upload(rng.integers(0, P, 1 << 8, dtype=np.uint32))
    tree = backend.precompute_canonic_twiddles(8)
    p = FriParams(4, 1, 0, 3, 1)
    with pytest.raises(HipError, match="invalid degree"):
        fri_prove(backend, tree, Channel(), p, col)
    with pytest.raises(HipError, match="fold_step"):
        fri_prove(backend, tree, Channel(), FriParams(4, 1, 0, 3, 0), col)
    # the first layer's folds must not pass the last layer (fri.rs:165-168)
    with pytest.raises(HipError, match="smaller line domain"):
        fri_prove(backend, tree, Channel(), FriParams(4, 1, 4, 3, 4), col)


def test_deep_fri_chain_on_device(ctx, backend):
    """Trace commit -> DEEP quotients -> FRI on the device; the oracle verifier accepts the proof
    with first-layer answers rebuilt from values gathered on the device at the query positions."""
    logs, b = [10] * 6 + [13] * 3 + [14] * 2, 1
    L = max(logs) + b
    rng = np.random.default_rng(21)
    host = [rng.integers(0, P, 1 << lg, dtype=np.uint32) for lg in logs]
    tree = backend.precompute_canonic_twiddles(L)
    evs = [BaseColumn.from_cpu(ctx, h) for h in host]
    coeffs, ldes, layers, root = backend.commit_trace(evs, b, tree)
    ch = Channel()
    ch.mix_root(root)
    z = ch.draw_random_point()
    samples = []
    for i, c in enumerate(coeffs):
        p = [int(v) for v in z]
        for _ in range(L - (logs[i] + b)):
            x, y = p[:4], p[4:]
            xx = PR.qmul(x, x)
            p = PR.qsub(PR.qadd(xx, xx), [1, 0, 0, 0]) + PR.qmulm(PR.qmul(x, y), 2)
        samples.append([(list(z), list(backend.eval_at_point(c, p)))])
    rc = list(ch.draw_secure_felt())
    quot = fri_quotients(backend, tree, ldes, samples, rc, L, b)
    pd = {"pow_bits": 8, "log_blowup_factor": b, "log_last_layer_degree_bound": 1, "n_queries": 16, "fold_step": 1}
    ch_before = ch.digest
    felts, qpos = fri_prove(backend, tree, ch, params_of(pd), quot)

    def fle(qs):
        rows = []
        for e, lg in zip(ldes, logs):
            n = lg + b
            rows.append(gather(backend, [e] * len(qs), [((p >> (L - n + 1)) << 1) + (p & 1) for p in qs]))
        return PR.fri_answers([lg + b for lg in logs], samples, rc, qs, rows, L)

    chv = O.Channel()
    chv.s.digest[:] = list(ch_before)
    chv.s.n_draws = 0
    assert PR.fri_verify(PR.deserialize_fri_proof(felts), pd, chv, L - b, fle) == qpos


def test_fri_deep_phase_full_size_verifies(ctx, backend):
    """The FRI + DEEP phase at 200 columns x 2^21 (LDE 2^22), 3 OOD points, 32 queries: the
    device proof is accepted by the oracle FRI verifier with first-layer answers rebuilt from
    values gathered on the device."""
    from proving_amd.backend import HashLayer
    from proving_amd.workloads import FRI_DEEP, splitmix64_columns
    w = FRI_DEEP
    n, N, b = w["log_size"], w["n_cols"], w["fri"]["log_blowup_factor"]
    L = n + b
    tree = backend.precompute_canonic_twiddles(L)
    host = splitmix64_columns(w["seed"], [n] * N)
    evals = [BaseColumn.from_cpu(ctx, c) for c in host]
    del host
    coeffs = [BaseColumn(ctx, n) for _ in range(N)]
    ldes = [BaseColumn(ctx, L) for _ in range(N)]
    backend.commit_trace(evals, b, tree, coeffs, ldes, [HashLayer(ctx, k) for k in range(L + 1)])
    ch = Channel()
    ch.mix_u64(w["seed"])
    points = [[int(x) for x in ch.draw_random_point()] for _ in range(w["n_points"])]
    samples = [[(p, [int(v) for v in backend.eval_at_point(c, p)]) for p in points] for c in coeffs]
    ch.mix_felts([v for col in samples for _, v in col])
    before = ch.digest
    rc = [int(x) for x in ch.draw_secure_felt()]
    quot = fri_quotients(backend, tree, ldes, samples, rc, L, b)
    felts, qpos = fri_prove(backend, tree, ch, params_of(w["fri"]), quot)
    assert len(qpos) <= 32

    def fle(qs):
        rows = [gather(backend, [e] * len(qs), qs) for e in ldes]
        return PR.fri_answers([L] * N, samples, rc, qs, rows, L)

    chv = O.Channel()
    chv.s.digest[:] = list(before)
    chv.s.n_draws = 0
    assert [int(v) for v in chv.draw_secure_felt()] == rc
    assert PR.fri_verify(PR.deserialize_fri_proof(felts), w["fri"], chv, L - b, fle) == qpos
