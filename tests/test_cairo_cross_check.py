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
    FRI driver produces for the same queries.
Not checked: the values inside the FRI layer openings, whose first-layer values are the DEEP
quotients of the Cairo AIR's masks (the AIR is not part of this build).
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
