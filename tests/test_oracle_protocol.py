"""The oracle's protocol restatement checked against itself from the other side: the Merkle
decommitment is accepted by the independent verifier (and tampering is not), FRI proofs made by
the prover restatement pass the verifier restatement, the DEEP-quotient driver agrees with the
verifier's per-row formula and gives low-degree output, and the felt encoding round-trips.

Mirrors crates/stwo/src/prover/pcs/quotient_ops.rs tests (test_quotients_are_low_degree,
prove_and_verify_pcs) and crates/stwo/src/prover/fri.rs tests (valid proof accepted, tampered
proof rejected), at the sizes the reference tests use.
"""
import numpy as np
import pytest

import oracle as O
from oracle import protocol as PR

P = O.P


def fri_test_params():
    return {"pow_bits": 10, "log_blowup_factor": 1, "log_last_layer_degree_bound": 0, "n_queries": 3, "fold_step": 1}


def double_q(pt):
    """CirclePoint::double over QM31: (2x^2 - 1, 2xy)."""
    x, y = list(pt[:4]), list(pt[4:])
    xx = PR.qmul(x, x)
    return PR.qsub(PR.qadd(xx, xx), [1, 0, 0, 0]) + PR.qmulm(PR.qmul(x, y), 2)


def random_points(seed, n):
    ch = O.Channel()
    ch.mix_u64(seed)
    return [[int(v) for v in ch.random_point()] for _ in range(n)]


def sample(coeffs, point, n_doublings):
    for _ in range(n_doublings):
        point = double_q(point)
    return [int(v) for v in O.eval_at_point(coeffs, point)]


# ---------------------------------------------------------------- Merkle
def test_merkle_decommit_verifies_and_rejects_tampering():
    rng = np.random.default_rng(4)
    cols = [rng.integers(0, P, 1 << 6, dtype=np.uint32) for _ in range(4)]
    layers = O.merkle_commit_lifted(cols, 6)
    root = layers[0][0].tobytes()
    positions = [1, 2, 3, 17, 40, 63]
    rows = [[int(c[p]) for c in cols] for p in positions]
    wit = PR.merkle_decommit(layers, positions)
    assert PR.merkle_verify(root, 6, positions, rows, wit)
    # a sibling pair (2, 3) needs no witness at the leaves; 6 positions, 5 leaf-level siblings
    assert len(wit) >= 4
    bad = [list(r) for r in rows]
    bad[2][1] = (bad[2][1] + 1) % P
    assert not PR.merkle_verify(root, 6, positions, bad, wit)
    assert not PR.merkle_verify(root, 6, positions, rows, wit[:-1])
    assert not PR.merkle_verify(root, 6, positions, rows, wit + [wit[0]])
    wit2 = list(wit)
    wit2[0] = bytes(32)
    assert not PR.merkle_verify(root, 6, positions, rows, wit2)


def test_merkle_decommit_all_positions_needs_no_witness():
    rng = np.random.default_rng(5)
    cols = [rng.integers(0, P, 1 << 4, dtype=np.uint32) for _ in range(4)]
    layers = O.merkle_commit_lifted(cols, 4)
    assert PR.merkle_decommit(layers, list(range(16))) == []
    # one position: one sibling per level
    assert len(PR.merkle_decommit(layers, [9])) == 4


# ---------------------------------------------------------------- quotients
def test_quotients_are_low_degree():
    """test_quotients_are_low_degree (crates/stwo/src/prover/pcs/quotient_ops.rs:186-221):
    coefficients 0..8, blowup 4, two samples; the quotient has degree < 2^3 - 1."""
    log_size, b = 3, 4
    coeffs = np.arange(1 << log_size, dtype=np.uint32)
    tw, itw = O.canonic_twiddles(log_size + b)
    ev = O.evaluate(tw, coeffs, log_size + b)
    pts = random_points(1, 2)
    smp = [(p, sample(coeffs, p, 0)) for p in pts]
    rc = [3, 5, 7, 11]
    q = PR.fri_quotients([ev], [smp], rc, log_size + b, b, tw, itw)
    for c in range(4):
        co = O.interpolate(itw, q[c], log_size + b)
        assert not co[(1 << log_size) - 1:].any()


def lifted_workload(seed, n_cols=10, trace_log=8, b=1):
    """prepare_polys (crates/stwo/src/prover/pcs/quotient_ops.rs:223-237): columns of random sizes
    below the lifting size plus one at full size, coefficients 0..2^k; one or two samples each."""
    rng = np.random.default_rng(seed)
    logs = [int(rng.integers(4, trace_log - 1)) for _ in range(n_cols - 1)] + [trace_log]
    L = trace_log + b
    tw, itw = O.canonic_twiddles(L)
    coeffs = [np.arange(1 << lg, dtype=np.uint32) for lg in logs]
    evals = [O.evaluate(tw, c, lg + b) for c, lg in zip(coeffs, logs)]
    pts = random_points(seed, 2)
    mask = [int(rng.integers(1, 3)) for _ in range(n_cols)]
    samples = [[(pts[k], sample(c, pts[k], L - (lg + b))) for k in range(m)]
               for c, lg, m in zip(coeffs, logs, mask)]
    return logs, L, tw, itw, evals, samples


def test_lifted_quotients_low_degree_and_match_row_formula():
    logs, L, tw, itw, evals, samples = lifted_workload(0)
    b = 1
    rc = [123, 456, 789, 1011]
    q = PR.fri_quotients(evals, samples, rc, L, b, tw, itw)
    # the combined quotient is computed on the 2^(L-b) subdomain and extended, so only the top
    # coefficient of that subdomain (the one degree the division by a line removes) carries a check
    for c in range(4):
        co = O.interpolate(itw, q[c], L)
        assert not co[(1 << (L - b)) - 1:].any()
    positions = [0, 1, 5, 100, 257, (1 << L) - 1]
    qv = [[PR.lifted_value(e, L, p) for p in positions] for e in evals]
    rows = PR.fri_answers([lg + b for lg in logs], samples, rc, positions, qv, L)
    for k, p in enumerate(positions):
        assert rows[k] == [int(q[c][p]) for c in range(4)]


def test_wrong_sample_value_breaks_degree():
    """With a wrong sample value the quotient keeps its top coefficient (the degree check has teeth)."""
    logs, L, tw, itw, evals, samples = lifted_workload(1)
    i = next(k for k, s in enumerate(samples) if len(s) == 2)
    pt, v = samples[i][1]
    samples[i][1] = (pt, [(v[0] + 1) % P] + v[1:])
    q = PR.fri_quotients(evals, samples, [5, 6, 7, 8], L, 1, tw, itw)
    assert any(O.interpolate(itw, q[c], L)[(1 << (L - 1)) - 1:].any() for c in range(4))


# ---------------------------------------------------------------- FRI
def fri_round_trip(params, seed=0):
    logs, L, tw, itw, evals, samples = lifted_workload(seed, trace_log=8, b=params["log_blowup_factor"])
    b = params["log_blowup_factor"]
    rc = [9, 8, 7, 6]
    q = PR.fri_quotients(evals, samples, rc, L, b, tw, itw)
    ch = O.Channel()
    ch.mix_u64(42)
    proof, queries = PR.fri_prove(q, L, params, ch)

    def first_layer_evals(qs):
        qv = [[PR.lifted_value(e, L, p) for p in qs] for e in evals]
        return PR.fri_answers([lg + b for lg in logs], samples, rc, qs, qv, L)

    return proof, queries, ch, L, first_layer_evals


@pytest.mark.parametrize("params", [
    fri_test_params(),
    {"pow_bits": 4, "log_blowup_factor": 2, "log_last_layer_degree_bound": 2, "n_queries": 20, "fold_step": 1},
    # fold_step > 1 with packed leaves (crates/stwo/src/prover/fri.rs:512-532 test parameters)
    {"pow_bits": 0, "log_blowup_factor": 2, "log_last_layer_degree_bound": 2, "n_queries": 3, "fold_step": 2},
    {"pow_bits": 0, "log_blowup_factor": 2, "log_last_layer_degree_bound": 2, "n_queries": 3, "fold_step": 3},
    {"pow_bits": 0, "log_blowup_factor": 2, "log_last_layer_degree_bound": 2, "n_queries": 3, "fold_step": 4},
    {"pow_bits": 6, "log_blowup_factor": 1, "log_last_layer_degree_bound": 0, "n_queries": 30, "fold_step": 3},
])
def test_fri_prove_verify_round_trip(params):
    proof, queries, ch, L, fle = fri_round_trip(params)
    felts = PR.serialize_fri_proof(proof)
    assert PR.deserialize_fri_proof(felts) == proof
    chv = O.Channel()
    chv.mix_u64(42)
    vq = PR.fri_verify(PR.deserialize_fri_proof(felts), params, chv, L - params["log_blowup_factor"], fle)
    assert vq == queries
    assert chv.digest == ch.digest
    fs, last_log = params["fold_step"], params["log_last_layer_degree_bound"] + params["log_blowup_factor"]
    assert len(proof["inner_layers"]) == -(-(L - fs - last_log) // fs)
    assert len(proof["last_layer_poly"]) == 1 << params["log_last_layer_degree_bound"]
    assert O.verify_pow_nonce(_digest_before_pow(proof, params, L), params["pow_bits"], proof["proof_of_work"])


def _digest_before_pow(proof, params, L):
    ch = O.Channel()
    ch.mix_u64(42)
    ch.mix_hash(proof["first_layer"]["commitment"])
    ch.draw_secure_felt()
    for lp in proof["inner_layers"]:
        ch.mix_hash(lp["commitment"])
        ch.draw_secure_felt()
    ch.mix_felts(np.array(proof["last_layer_poly"], np.uint32).reshape(-1))
    return ch.digest


@pytest.mark.parametrize("fold_step", [1, 3])
@pytest.mark.parametrize("tamper", ["fri_witness", "hash_witness", "last_layer", "pow", "inner_commitment",
                                    "first_eval"])
def test_fri_verify_rejects_tampering(tamper, fold_step):
    params = {"pow_bits": 4, "log_blowup_factor": 1, "log_last_layer_degree_bound": 1, "n_queries": 8,
              "fold_step": fold_step}
    proof, queries, ch, L, fle = fri_round_trip(params, seed=2)
    evals_fn = fle
    if tamper == "fri_witness":
        lp = proof["inner_layers"][1 if fold_step == 1 else 0]
        lp["fri_witness"][0] = [(lp["fri_witness"][0][0] + 1) % P] + list(lp["fri_witness"][0][1:])
    elif tamper == "hash_witness":
        proof["first_layer"]["hash_witness"][0] = bytes(32)
    elif tamper == "last_layer":
        proof["last_layer_poly"][0] = [(proof["last_layer_poly"][0][0] + 1) % P] + list(proof["last_layer_poly"][0][1:])
    elif tamper == "pow":
        proof["proof_of_work"] += 1
    elif tamper == "inner_commitment":
        proof["inner_layers"][0]["commitment"] = bytes(32)
    elif tamper == "first_eval":
        def evals_fn(qs):
            v = fle(qs)
            v[0] = [(v[0][0] + 1) % P] + v[0][1:]
            return v
    chv = O.Channel()
    chv.mix_u64(42)
    with pytest.raises(AssertionError):
        PR.fri_verify(proof, params, chv, L - 1, evals_fn)


def test_fri_prove_rejects_high_degree():
    """commit_last_layer panics with "invalid degree" (crates/stwo/src/prover/fri.rs:215-221)."""
    rng = np.random.default_rng(3)
    col = rng.integers(0, P, (4, 1 << 6), dtype=np.uint32)
    with pytest.raises(AssertionError, match="invalid degree"):
        PR.fri_prove(col, 6, fri_test_params(), O.Channel())


# ---------------------------------------------------------------- wide-Fibonacci composition
def _composition_case(n_cols, log_n, seed, tamper=False):
    from proving_amd.workloads import wide_fib_trace
    trace = wide_fib_trace(n_cols, log_n, seed)
    if tamper:
        trace[n_cols // 2][3] = (int(trace[n_cols // 2][3]) + 1) % P
    L = log_n + 1
    tw, itw = O.canonic_twiddles(L)
    coeffs = [O.interpolate(itw, c, log_n) for c in trace]
    ldes = [O.evaluate(tw, co, L) for co in coeffs]
    ch = O.Channel()
    ch.mix_u64(seed)
    alpha = [int(v) for v in ch.draw_secure_felt()]
    z = [int(v) for v in ch.random_point()]
    m = n_cols - 2
    pw = [[1, 0, 0, 0]]
    for _ in range(m - 1):
        pw.append(PR.qmul(pw[-1], alpha))
    powers = pw[::-1]  # constraint j gets alpha^(m-1-j)
    comp = PR.wide_fib_composition(ldes, log_n, L, 0, powers)
    comp_coeffs = [O.interpolate(itw, comp[t], L) for t in range(4)]
    lhs = PR.from_partial_evals([[int(v) for v in O.eval_at_point(comp_coeffs[t], z)] for t in range(4)])
    mask = [[int(v) for v in O.eval_at_point(co, z)] for co in coeffs]
    rhs = PR.wide_fib_composition_at_point(mask, alpha, z, log_n)
    return lhs, rhs, comp_coeffs


@pytest.mark.parametrize("n_cols,log_n", [(3, 3), (5, 4), (16, 6), (40, 5)])
def test_wide_fib_composition_matches_point_evaluation(n_cols, log_n):
    """The domain-side composition (interpolated) agrees with the verifier's point-side
    evaluation at a random out-of-domain point (crates/stwo/src/core/verifier.rs:102-117), which
    fixes the weights' order, the denominators and the constraint."""
    lhs, rhs, _ = _composition_case(n_cols, log_n, 7)
    assert lhs == rhs


def test_wide_fib_composition_detects_a_bad_row():
    lhs, rhs, _ = _composition_case(16, 5, 8, tamper=True)
    assert lhs != rhs


# ---------------------------------------------------------------- full prove -> verify
def _wf_params(pow_bits=4, last=0, n_queries=3):
    return {"pow_bits": pow_bits, "log_blowup_factor": 1, "log_last_layer_degree_bound": last,
            "n_queries": n_queries, "fold_step": 1}


@pytest.mark.parametrize("n_cols,log_n", [(4, 4), (12, 6)])
def test_wide_fib_prove_verify_round_trip(n_cols, log_n):
    """test_wide_fib_prove_with_blake (crates/examples/src/wide_fibonacci/mod.rs:183-237) on the
    oracle: the proof verifies, and its felts decode back to the same proof."""
    from proving_amd.workloads import wide_fib_trace
    trace = wide_fib_trace(n_cols, log_n, 11)
    fri = _wf_params()
    tw, itw = O.canonic_twiddles(log_n + 1)
    felts, proof = PR.wide_fib_prove(trace, log_n, fri, tw, itw)
    assert PR.wide_fib_verify(felts, n_cols, log_n, fri)
    d = PR.deserialize_pcs_proof(felts)
    assert d["commitments"] == proof["commitments"] and d["queried_values"] == proof["queried_values"]


@pytest.mark.parametrize("tamper", ["trace", "queried", "sampled", "pow", "commitment"])
def test_wide_fib_verify_rejects(tamper):
    from proving_amd.workloads import wide_fib_trace
    n_cols, log_n = 6, 5
    trace = wide_fib_trace(n_cols, log_n, 12)
    if tamper == "trace":
        trace[3][5] = (int(trace[3][5]) + 1) % P
    fri = _wf_params()
    tw, itw = O.canonic_twiddles(log_n + 1)
    felts, proof = PR.wide_fib_prove(trace, log_n, fri, tw, itw)
    if tamper == "queried":
        proof["queried_values"][1][2][0] = (proof["queried_values"][1][2][0] + 1) % P
    elif tamper == "sampled":
        proof["sampled_values"][1][0][0] = PR.qadd(proof["sampled_values"][1][0][0], [1, 0, 0, 0])
    elif tamper == "pow":
        proof["oods_proof_of_work"] += 1
    elif tamper == "commitment":
        proof["commitments"][1] = bytes(32)
    felts = PR.serialize_pcs_proof(proof)
    with pytest.raises(AssertionError):
        PR.wide_fib_verify(felts, n_cols, log_n, fri)
