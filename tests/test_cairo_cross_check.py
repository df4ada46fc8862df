"""Cross-check against a proof the reference prover made: the Cairo proof of
test_prove_verify_all_opcode_components (crates/prover/tests/slow_tests.rs:420-463, fixture
tests/golden/cairo_proof_all_opcodes.npz written by tests/golden/make_cairo_fixture.py).

Checked with this build's code:
  * the CairoSerde layout of the whole felt list (claim, interaction claim, commitment-scheme
    proof, FRI proof, salt) parses with nothing left over;
  * the Fiat-Shamir transcript replayed with the library's Blake2s channel (and the oracle's)
    reproduces the reference's: the three proof-of-work nonces the reference prover ground on its
    own transcript (interaction 24 bits, OODS 16 bits, FRI 26 bits) verify, and a changed input
    before a nonce makes it fail;
  * the query positions drawn from that transcript open all four committed trees (preprocessed,
    trace, interaction, composition: lifted Merkle trees over columns of mixed sizes, the
    preprocessed one at its own height) against the proof's roots; on the GPU the leaf hashes are
    recomputed by the library's leaf kernel.
  * the FRI layer schedule and opening sizes (fold_step 3, packed leaves) equal those this build's
    FRI driver produces for the same queries;
  * the whole FRI decommitment verifies against the proof's roots: the first layer's values at the
    queries are the DEEP quotients of the queried values (fri_answers with the periodicity samples
    of two-sample columns and the random-coefficient order across the four trees), then every
    circle and line fold of fold_step 3 and the last-layer polynomial.  The oracle does this on the
    CPU; on the GPU the quotient numerators come from the library's numerator kernel and every line
    fold from its fold kernel.  The sample points come from the proof's own layout: a column
    sampled once is sampled at the OODS point, one sampled twice is the last cumulative sum of a
    logup (mask offsets [-1, 0], crates/constraint-framework/src/lib.rs:208-209), and a wrong
    offset, or a wrong periodicity, fails the first layer's Merkle root.  The column log sizes
    that order the queried values come from the claim (tests/cairo_transcript.py TRACE_COLUMNS);
    the Cairo AIR's constraints are not restated.
"""
import ctypes
import os

import numpy as np
import pytest

import cairo_transcript as CT
import oracle as O
from oracle import protocol as PR

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cairo_proof_all_opcodes.npz")


@pytest.fixture(scope="module")
def proof():
    with np.load(FIXTURE) as z:
        return CT.parse_cairo_proof(z["felts"])


class _OracleChannel(O.Channel):
    def mix_root(self, root):
        self.mix_hash(root)


def _lib_channel():
    from proving_amd.backend import Channel
    return Channel()


def _clone(ch):
    c = type(ch)()
    ctypes.memmove(ctypes.byref(c.s), ctypes.byref(ch.s), ctypes.sizeof(ch.s))
    return c


def _queries(ch, log_size, n):
    """draw_queries + Queries::new (crates/stwo/src/core/queries.rs): sorted, deduplicated."""
    return sorted(set(_clone(ch).draw_queries(log_size, n)))


def _verifies(root, height, positions, rows, witness):
    try:
        return PR.merkle_verify(root, height, positions, rows, witness)
    except AssertionError:  # two rows for one position disagree
        return False


def test_layout(proof):
    assert proof["config"] == {"pow_bits": 26, "log_blowup_factor": 1, "log_last_layer_degree_bound": 0,
                               "n_queries": 70, "fold_step": 3}
    assert len(proof["commitments"]) == 4 and proof["channel_salt"] == 0
    # a row of every tree holds one value per sampled column
    n = len(proof["queried_values"][3]) // 8
    for t in range(4):
        assert len(proof["queried_values"][t]) == n * len(proof["sampled_values"][t])
    assert [len(c) for c in proof["sampled_values"][3]] == [1] * 8
    assert len(proof["fri"]["last_layer_poly"]) == 1 << proof["fri"]["last_layer_log"]


@pytest.mark.parametrize("make_channel", [_lib_channel, _OracleChannel], ids=["library", "oracle"])
def test_transcript_reproduces_reference_grinds(proof, make_channel):
    checks, _ = CT.replay(proof, make_channel(), O.verify_pow_nonce)
    assert checks == {"interaction_pow": True, "oods_pow": True, "fri_pow": True}


def test_changed_transcript_breaks_grinds(proof):
    import copy
    p = copy.deepcopy(proof)
    p["claims"][0] = (p["claims"][0][0], "log", p["claims"][0][2] + 1)   # a claimed log size
    checks, _ = CT.replay(p, _lib_channel(), O.verify_pow_nonce)
    assert not checks["interaction_pow"]
    p = copy.deepcopy(proof)
    v = p["sampled_values"][1][0][0]
    p["sampled_values"][1][0][0] = ((v[0] + 1) % CT.P,) + tuple(v[1:])   # one sampled value
    checks, _ = CT.replay(p, _lib_channel(), O.verify_pow_nonce)
    assert checks["interaction_pow"] and checks["oods_pow"] and not checks["fri_pow"]


def _open(proof):
    """Lifting log size (the one height at which the composition tree opens), the queries, and the
    preprocessed tree's height."""
    _, ch = CT.replay(proof, _lib_channel(), O.verify_pow_nonce)
    nq = proof["config"]["n_queries"]
    hits = []
    for L in range(4, 31):
        q = _queries(ch, L, nq)
        rows = CT.tree_rows(proof, 3, len(q)) if len(proof["queried_values"][3]) % len(q) == 0 else None
        if rows and _verifies(proof["commitments"][3], L, q, rows, proof["decommitments"][3]):
            hits.append(L)
    assert len(hits) == 1, hits
    L = hits[0]
    q = _queries(ch, L, nq)
    pp = [h for h in range(1, 31)
          if _verifies(proof["commitments"][0], h, CT.preprocessed_positions(q, L, h),
                       CT.tree_rows(proof, 0, len(q)), proof["decommitments"][0])]
    assert len(pp) == 1, pp
    return L, q, pp[0]


def test_decommitments_open_against_reference_roots(proof):
    L, q, pp_height = _open(proof)
    assert L == 21 and len(q) == 70 and pp_height == 26
    for t in (1, 2):
        assert PR.merkle_verify(proof["commitments"][t], L, q, CT.tree_rows(proof, t, len(q)),
                                proof["decommitments"][t])


def test_fri_opening_shape_matches_reference(proof):
    """The FRI layer schedule, coset openings and packed-leaf decommitments this build's FRI driver
    makes for the drawn queries have exactly the sizes of the reference proof's (fold_step 3: six
    inner layers, the last folding by 2)."""
    L, q, _ = _open(proof)
    fri = proof["fri"]
    got = [(len(fri["first_layer"]["fri_witness"]), len(fri["first_layer"]["decommitment"]))]
    got += [(len(lp["fri_witness"]), len(lp["decommitment"])) for lp in fri["inner_layers"]]
    assert CT.fri_opening_shape(q, L, proof["config"]) == got
    # and with the queries of a wrong lifting size the shape differs
    assert CT.fri_opening_shape([x >> 1 for x in q], L - 1, proof["config"]) != got


def test_tampered_openings_rejected(proof):
    L, q, _ = _open(proof)
    rows = CT.tree_rows(proof, 1, len(q))
    bad = [list(r) for r in rows]
    bad[5][100] = (bad[5][100] + 1) % CT.P
    assert not _verifies(proof["commitments"][1], L, q, bad, proof["decommitments"][1])
    wit = list(proof["decommitments"][2])
    wit[3] = bytes([wit[3][0] ^ 1]) + wit[3][1:]
    assert not _verifies(proof["commitments"][2], L, q, CT.tree_rows(proof, 2, len(q)), wit)
    # a position off by one opens nothing
    q2 = list(q)
    q2[0] ^= 2
    assert not _verifies(proof["commitments"][3], L, sorted(q2), CT.tree_rows(proof, 3, len(q)),
                         proof["decommitments"][3])


@pytest.mark.gpu
def test_gpu_leaf_hashes_open_reference_trees(proof, backend, ctx):
    """Leaf hashes of every queried row of the four trees from the library's leaf kernel
    (each row's values as one row of equal-size columns), then the Merkle paths to the roots."""
    from proving_amd.backend import BaseColumn
    L, q, pp_height = _open(proof)
    k = max(1, (len(q) - 1).bit_length())
    for t in range(4):
        rows = np.array(CT.tree_rows(proof, t, len(q)), np.uint32)
        cols = [BaseColumn.from_cpu(ctx, np.pad(rows[:, c], (0, (1 << k) - len(q)))) for c in range(rows.shape[1])]
        leaves = backend.build_leaves(cols, k).to_cpu()
        for c in cols:
            c.free()
        got = [bytes(leaves[i]) for i in range(len(q))]
        want = [CT._leaf_hash(r) for r in rows.tolist()]
        assert got == want, f"tree {t}"
        height = pp_height if t == 0 else L
        pos = CT.preprocessed_positions(q, L, pp_height) if t == 0 else q
        assert PR.merkle_verify_leaves(proof["commitments"][t], height, pos, got, proof["decommitments"][t])


# ---------------------------------------------------------------- the FRI decommitment
def _shift(z, g_index, k):
    """z + k * P(g_index) for k in {-1, 0, 1}: CirclePoint::mul_signed + add
    (crates/stwo/src/core/circle.rs:101-103, 106-118), P an M31 point, z a QM31 point [x4, y4]."""
    if k == 0:
        return list(z)
    g = O.index_to_point(g_index)
    gx, gy = [int(g[0]), 0, 0, 0], [int(g[1]) if k > 0 else (CT.P - int(g[1])) % CT.P, 0, 0, 0]
    x, y = z[:4], z[4:]
    return PR.qsub(PR.qmul(x, gx), PR.qmul(y, gy)) + PR.qadd(PR.qmul(x, gy), PR.qmul(y, gx))


def _pcs_inputs(proof, two_offsets=(-1, 0), interaction_log_shift=0):
    """Arguments of the oracle's pcs_verify_values for the reference proof: tree roots and heights,
    extended column log sizes, sample points (mask_points, crates/constraint-framework/src/
    component.rs:202-213: z + offset * CanonicCoset(L - b).step(); the composition columns at z,
    crates/stwo/src/core/verifier.rs:98-100), the proof in the oracle's layout, and the channel
    just before verify_values mixes the sampled values."""
    L, q, pp_height = _open(proof)
    b = proof["config"]["log_blowup_factor"]
    _, ch = CT.replay(proof, _OracleChannel(), O.verify_pow_nonce, until="oods")
    z = [int(v) for v in ch.random_point()]   # CirclePoint::get_random_point, verifier.rs:95
    trace, inter = CT.column_log_sizes(proof)
    npp = len(proof["sampled_values"][0])
    logs = [[pp_height] * npp, [lg + b for lg in trace], [lg + b + interaction_log_shift for lg in inter], [L] * 8]
    step = 1 << (31 - (L - b))
    points = [[[_shift(z, step, k) for k in ((0,) if len(c) == 1 else two_offsets if len(c) == 2 else ())]
               for c in tree] for tree in proof["sampled_values"]]
    queried = [CT.columns_in_order(proof, t, len(q), logs[t]) for t in range(4)]
    fri = proof["fri"]

    def layer(lp):
        return {"commitment": lp["commitment"], "fri_witness": [list(w) for w in lp["fri_witness"]],
                "hash_witness": lp["decommitment"]}

    oproof = {"sampled_values": [[[list(v) for v in c] for c in t] for t in proof["sampled_values"]],
              "queried_values": queried, "decommitments": proof["decommitments"],
              "fri_proof": {"first_layer": layer(fri["first_layer"]),
                            "inner_layers": [layer(lp) for lp in fri["inner_layers"]],
                            "last_layer_poly": [list(v) for v in fri["last_layer_poly"]],
                            "proof_of_work": fri["proof_of_work"]}}
    return (proof["commitments"], [pp_height, L, L, L], logs, points, oproof, ch, dict(proof["config"]))


def test_fri_decommitment_verifies(proof):
    """CommitmentSchemeVerifier::verify_values (crates/stwo/src/core/pcs/verifier.rs:55-130) on the
    reference proof with the oracle's fri_answers and folds: the four trees open, and the FRI
    first layer, its six inner layers and the last layer accept the values computed here."""
    assert PR.pcs_verify_values(*_pcs_inputs(proof))


@pytest.mark.parametrize("case", ["next_row_mask", "no_prev_row", "periodicity_log"])
def test_fri_decommitment_rejects_wrong_samples(proof, case):
    """Each of these changes only the DEEP quotients, and the first FRI layer's root refuses them:
    the cumulative-sum columns sampled at offsets [0, 1] or [0, 0] instead of [-1, 0], or the
    periodicity sample taken with the interaction columns' log sizes one too large."""
    kw = {"next_row_mask": {"two_offsets": (0, 1)}, "no_prev_row": {"two_offsets": (0, 0)},
          "periodicity_log": {"interaction_log_shift": 1}}[case]
    with pytest.raises(AssertionError, match="CommitmentInvalid"):
        PR.pcs_verify_values(*_pcs_inputs(proof, **kw))


@pytest.mark.gpu
def test_gpu_numerators_and_folds_verify_reference_fri(proof, backend, ctx):
    """The same decommitment with the library doing the arithmetic: the DEEP quotient numerators
    of every sample batch at the queried rows from sh_accumulate_numerators (the rows as columns,
    blowup 0), and every line fold of every FRI subset from sh_fold_line on a twiddle tree rooted
    at that subset's coset.  The row sums over the batches (numerator minus first-linear term times
    p.y, times the CM31 denominator inverse: accumulate_row_quotients,
    crates/stwo/src/core/pcs/quotients.rs:155-188) are taken on the host and must equal the
    oracle's fri_answers; the reference's roots then accept the result."""
    from proving_amd.backend import BaseColumn, SecureColumnByCoords
    calls = {"answers": 0, "folds": 0}

    def answers(flat_logs, flat_samples, rc, queries, qv, L):
        batches = list(PR._batches(PR._build_samples(flat_logs, flat_samples, rc, L)).items())
        k = max(1, (len(queries) - 1).bit_length())
        cols = [BaseColumn.from_cpu(ctx, np.pad(np.array(c, np.uint32), (0, (1 << k) - len(queries)))) for c in qv]
        parts = backend.accumulate_numerators(cols, [(list(pt), terms) for pt, terms in batches], 0)
        nums = [(pt, part.to_cpu(), [int(x) for x in fl]) for pt, part, fl in parts]
        for c in cols:
            c.free()
        for _, part, _ in parts:
            part.free()
        half = O.canonic_half_initial(L)
        got = []
        for r, pos in enumerate(queries):
            px, py = (int(c) for c in PR.circle_domain_at(half, L - 1, PR.bitrev(pos, L)))
            acc = [0, 0, 0, 0]
            for pt, num, fl in nums:
                pt = [int(x) for x in pt]
                n = PR.qsub([int(num[c][r]) for c in range(4)], PR.qmulm(fl, py))
                t0 = PR._cmul([(pt[0] - px) % CT.P, pt[1]], [pt[6], pt[7]])
                t1 = PR._cmul([(pt[4] - py) % CT.P, pt[5]], [pt[2], pt[3]])
                den = PR._cinv([(t0[0] - t1[0]) % CT.P, (t0[1] - t1[1]) % CT.P])
                acc = PR.qadd(acc, PR._cmul(n[:2], den) + PR._cmul(n[2:], den))
            got.append(acc)
        assert got == PR.fri_answers(flat_logs, flat_samples, rc, queries, qv, L)
        calls["answers"] += 1
        return got

    def fold_coset(evals, init_idx, log, alpha):
        alphas = [list(alpha)]
        for _ in range(log - 1):
            alphas.append(PR.qmul(alphas[-1], alphas[-1]))
        tw = backend.precompute_twiddles(init_idx, log)
        src = SecureColumnByCoords.from_cpu(ctx, np.array(evals, np.uint32).T.copy())
        out = backend.fold_line(src, init_idx, alphas, tw)
        v = [int(x) for x in out.to_cpu()[:, 0]]
        for c in (src, out, tw):
            c.free()
        calls["folds"] += 1
        return v

    def fold_circle(evals, init_idx, fold_step, alpha):
        return PR.fold_circle(evals, init_idx, fold_step, alpha, fold_coset_fn=fold_coset)

    assert PR.pcs_verify_values(*_pcs_inputs(proof), answers_fn=answers, fold_circle_fn=fold_circle,
                                fold_coset_fn=fold_coset)
    assert calls["answers"] == 1 and calls["folds"] > 6 * 32
