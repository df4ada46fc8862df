"""Reader and Fiat-Shamir transcript replay for a reference-produced Cairo proof.

The proof is the CairoSerde felt list of CairoProof (crates/cairo-air/src/serde_utils.rs:12-58):
the claim, the interaction proof-of-work nonce, the flattened interaction claim, then the
commitment-scheme proof (config, commitments, OODS nonce, sampled values, Merkle decommitments,
queried values sorted by column size and transposed to rows, FRI proof) and the channel salt.

The replay follows verify_cairo_ex (crates/cairo-air/src/verifier.rs:256-329), verify_ex
(crates/stwo/src/core/verifier.rs:29-118), CommitmentSchemeVerifier::verify_values
(crates/stwo/src/core/pcs/verifier.rs:55-132) and FriVerifier::commit / verify_proof_of_work /
sample_query_positions (crates/stwo/src/core/fri.rs:158-359), with the channel under test.  Its
three proof-of-work nonces (interaction 24 bits, OODS 16 bits, FRI 26 bits) were ground by the
reference prover on its own transcript, so each one that verifies shows that every channel
operation before it matched the reference's.  The Cairo AIR itself (constraints, masks, lookup
sums) is not restated: only the claim's layout, which the transcript mixes, is.
"""
import hashlib
import struct

P = (1 << 31) - 1
INTERACTION_POW_BITS = 24   # crates/cairo-air/src/verifier.rs:248
OODS_POW_BITS = 16          # crates/stwo/src/core/proof_of_work.rs:8
MEMORY_ADDRESS_TO_ID_SPLIT = 16  # 1 << (LOG_MEMORY_ADDRESS_BOUND 29 - MAX_SEQUENCE_LOG_SIZE 25)
FELT252_N_WORDS, FELT252_BITS_PER_WORD = 28, 9  # crates/common/src/prover_types/cpu.rs:385-386

# CairoClaim fields in declaration order (crates/cairo-air/src/claims.rs:17-86) and the payload
# of Some(claim): "log" = Claim { log_size }, an integer = an empty Claim of a component with that
# fixed LOG_SIZE (crates/cairo-air/src/components/<name>.rs:6), "big" = memory_id_to_big's
# Claim { big_log_sizes: Vec<u32> }.
COMPONENTS = (
    [(n, "log") for n in (
        "add_opcode", "add_opcode_small", "add_ap_opcode", "assert_eq_opcode", "assert_eq_opcode_imm",
        "assert_eq_opcode_double_deref", "blake_compress_opcode", "call_opcode_abs", "call_opcode_rel_imm",
        "jnz_opcode_non_taken", "jnz_opcode_taken", "jump_opcode_abs", "jump_opcode_double_deref",
        "jump_opcode_rel", "jump_opcode_rel_imm", "mul_opcode", "mul_opcode_small", "qm_31_add_mul_opcode",
        "ret_opcode", "verify_instruction", "blake_round", "blake_g")]
    + [("blake_round_sigma", 4), ("triple_xor_32", "log"), ("verify_bitwise_xor_12", 20)]
    + [(n, "log") for n in (
        "add_mod_builtin", "bitwise_builtin", "mul_mod_builtin", "pedersen_builtin",
        "pedersen_builtin_narrow_windows", "poseidon_builtin", "range_check96_builtin", "range_check_builtin",
        "ec_op_builtin", "partial_ec_mul_generic", "pedersen_aggregator_window_bits_18",
        "partial_ec_mul_window_bits_18")]
    + [("pedersen_points_table_window_bits_18", 23), ("pedersen_aggregator_window_bits_9", "log"),
       ("partial_ec_mul_window_bits_9", "log"), ("pedersen_points_table_window_bits_9", 15)]
    + [(n, "log") for n in ("poseidon_aggregator", "poseidon_3_partial_rounds_chain", "poseidon_full_round_chain",
                            "cube_252")]
    + [("poseidon_round_keys", 6), ("range_check_252_width_27", "log"), ("memory_address_to_id", "log"),
       ("memory_id_to_big", "big"), ("memory_id_to_small", "log"),
       ("range_check_6", 6), ("range_check_8", 8), ("range_check_11", 11), ("range_check_12", 12),
       ("range_check_18", 18), ("range_check_20", 20), ("range_check_4_3", 7), ("range_check_4_4", 8),
       ("range_check_9_9", 18), ("range_check_7_2_5", 14), ("range_check_3_6_6_3", 18),
       ("range_check_4_4_4_4", 16), ("range_check_3_3_3_3_3", 15), ("verify_bitwise_xor_4", 8),
       ("verify_bitwise_xor_7", 14), ("verify_bitwise_xor_8", 16), ("verify_bitwise_xor_9", 18)]
)


class _Reader:
    def __init__(self, felts):
        self.f = [int(x) for x in felts]
        self.i = 0

    def u(self):
        v = self.f[self.i]
        self.i += 1
        return v

    def take(self, n):
        v = self.f[self.i:self.i + n]
        assert len(v) == n, "proof ends early"
        self.i += n
        return v

    def vec(self, item):
        return [item() for _ in range(self.u())]

    def qm31(self):
        return tuple(self.take(4))

    def hash(self):
        """Blake2sHash: 8 little-endian u32 words (crates/cairo-serialize/src/serialize.rs:213-221)."""
        return struct.pack("<8I", *self.take(8))


def _fri_layer(r):
    """FriLayerProof: fri_witness (QM31s), decommitment (hash witness), commitment."""
    return {"fri_witness": r.vec(r.qm31), "decommitment": r.vec(r.hash), "commitment": r.hash()}


def parse_cairo_proof(felts):
    """CairoProof felts -> dict; asserts the layout consumes every felt."""
    r = _Reader(felts)
    pd = {}
    # PublicData { public_memory: PublicMemory { program, public_segments, output, safe_call_ids },
    # initial_state, final_state } (crates/cairo-air/src/air.rs:91-95, 443-450); the segments go
    # out as FullSegmentRanges, 11 x (start id, start value, stop id, stop value) (air.rs:298-330)
    pd["program"] = r.vec(lambda: (r.u(), r.take(8)))
    pd["segments"] = r.take(44)
    pd["output"] = r.vec(lambda: (r.u(), r.take(8)))
    pd["safe_call_ids"] = r.take(2)
    pd["initial_state"] = r.take(3)
    pd["final_state"] = r.take(3)
    claims = []
    for name, kind in COMPONENTS:
        tag = r.u()  # Option: 0 = Some, 1 = None (serialize.rs:170-180)
        assert tag in (0, 1), f"bad option tag for {name}"
        if tag == 1:
            claims.append((name, kind, None))
        elif kind == "log":
            claims.append((name, kind, r.u()))
        elif kind == "big":
            claims.append((name, kind, r.vec(r.u)))
        else:
            claims.append((name, kind, ()))
    proof = {"public_data": pd, "claims": claims, "interaction_pow": r.u(),
             "interaction_claim": r.vec(r.qm31)}
    proof["config"] = dict(zip(("pow_bits", "log_blowup_factor", "log_last_layer_degree_bound", "n_queries",
                                "fold_step"), r.take(5)))
    proof["commitments"] = r.vec(r.hash)
    proof["oods_pow"] = r.u()
    proof["sampled_values"] = r.vec(lambda: r.vec(lambda: r.vec(r.qm31)))
    proof["decommitments"] = r.vec(lambda: r.vec(r.hash))
    proof["queried_values"] = r.vec(lambda: r.vec(r.u))
    fri = {"first_layer": _fri_layer(r), "inner_layers": r.vec(lambda: _fri_layer(r))}
    fri["last_layer_poly"] = r.vec(r.qm31)
    fri["last_layer_log"] = r.u()
    fri["proof_of_work"] = r.u()
    proof["fri"] = fri
    proof["channel_salt"] = r.u()
    assert r.i == len(r.f), f"{len(r.f) - r.i} felts left after the proof"
    return proof


def flat_claim(claims):
    """CairoClaim::flatten_claim (crates/cairo-air/src/claims.rs:467-880): enable bits and log
    sizes in field order; memory_id_to_big contributes one bit per big log size, padded with
    zeros to MEMORY_ADDRESS_TO_ID_SPLIT."""
    bits, logs = [], []
    for name, kind, payload in claims:
        if kind == "big":
            assert payload is not None, "memory_id_to_big is required"
            assert len(payload) <= MEMORY_ADDRESS_TO_ID_SPLIT
            bits += [1] * len(payload) + [0] * (MEMORY_ADDRESS_TO_ID_SPLIT - len(payload))
            logs += payload
        elif payload is None:
            bits.append(0)
        else:
            bits.append(1)
            logs.append(payload if kind == "log" else kind)
    return bits, logs


def pack_into_secure_felts(values):
    """crates/cairo-air/src/utils.rs:52-64: groups of four M31s, the last one zero-padded."""
    v = [int(x) % P for x in values]
    return [tuple(v[i:i + 4] + [0] * (4 - len(v[i:i + 4]))) for i in range(0, len(v), 4)]


def felt252_limbs(words):
    """split(value, 2^9 - 1) (crates/common/src/prover_types/felt.rs:19-53): the 9-bit limbs of the
    little-endian 256-bit value, lowest first, 28 of them."""
    x = sum(int(w) << (32 * i) for i, w in enumerate(words))
    return [(x >> (FELT252_BITS_PER_WORD * k)) & ((1 << FELT252_BITS_PER_WORD) - 1) for k in range(FELT252_N_WORDS)]


def _leaf_hash(m31s):
    """Blake2sMerkleHasher::update_leaf + finalize (crates/stwo/src/core/vcs_lifted/blake2_merkle.rs:38-44)."""
    return hashlib.blake2s(b"".join(struct.pack("<I", v) for v in m31s)).digest()


def public_data_parts(pd):
    """PublicData::pack_into_u32s (crates/cairo-air/src/air.rs:162-222)."""
    public = list(pd["initial_state"]) + list(pd["final_state"]) + list(pd["segments"]) + list(pd["safe_call_ids"])
    public += [i for i, _ in pd["output"]] + [i for i, _ in pd["program"]]
    output = [l for _, v in pd["output"] for l in felt252_limbs(v)]
    program = [l for _, v in pd["program"] for l in felt252_limbs(v)]
    return public, output, program


def replay(proof, channel, verify_pow, until=None):
    """Runs the verifier's channel operations over the proof; returns the proof-of-work checks and
    the channel just before the queries are drawn.  `channel` is a fresh Blake2s channel
    (mix_u32s, mix_u64, mix_felts, mix_root, digest); `verify_pow(digest, bits, nonce)`.
    until="oods": stop right after the OODS nonce is mixed, where verify_ex draws the OODS point."""
    ch = channel
    cfg = proof["config"]
    checks = {}
    ch.mix_felts([(proof["channel_salt"] % P, 0, 0, 0)])
    # FriParams::mix_into (crates/stwo/src/core/fri.rs:75-86): note n_queries before the last-layer bound
    ch.mix_u32s([cfg["pow_bits"], cfg["log_blowup_factor"], cfg["n_queries"], cfg["log_last_layer_degree_bound"],
                 cfg["fold_step"]])
    ch.mix_root(proof["commitments"][0])
    # CairoClaim::mix_into -> FlatClaim::mix_into (crates/cairo-air/src/flat_claims.rs:12-25)
    bits, logs = flat_claim(proof["claims"])
    pd = proof["public_data"]
    ch.mix_felts(pack_into_secure_felts([len(bits)]))
    ch.mix_felts(pack_into_secure_felts(bits))
    ch.mix_felts(pack_into_secure_felts(logs))
    ch.mix_felts(pack_into_secure_felts([len(pd["program"])]))
    # PublicData::mix_into (crates/cairo-air/src/air.rs:142-160)
    public, output, program = public_data_parts(pd)
    ch.mix_felts(pack_into_secure_felts(public))
    ch.mix_root(_leaf_hash(output))
    ch.mix_root(_leaf_hash(program))
    ch.mix_root(proof["commitments"][1])
    checks["interaction_pow"] = verify_pow(ch.digest, INTERACTION_POW_BITS, proof["interaction_pow"])
    ch.mix_u64(proof["interaction_pow"])
    # the lookup elements are drawn here; draws leave the digest as it is
    ch.mix_felts(proof["interaction_claim"])
    ch.mix_root(proof["commitments"][2])
    # verify_ex: random coefficient drawn, composition tree committed, OODS grind checked
    ch.mix_root(proof["commitments"][3])
    checks["oods_pow"] = verify_pow(ch.digest, OODS_POW_BITS, proof["oods_pow"])
    ch.mix_u64(proof["oods_pow"])
    if until == "oods":
        return checks, ch
    # OODS point drawn; verify_values mixes every sampled value, draws the FRI coefficient
    ch.mix_felts([v for tree in proof["sampled_values"] for col in tree for v in col])
    fri = proof["fri"]
    ch.mix_root(fri["first_layer"]["commitment"])
    for layer in fri["inner_layers"]:
        ch.mix_root(layer["commitment"])
    ch.mix_felts(fri["last_layer_poly"])
    checks["fri_pow"] = verify_pow(ch.digest, cfg["pow_bits"], fri["proof_of_work"])
    ch.mix_u64(fri["proof_of_work"])
    return checks, ch


def tree_rows(proof, tree, n_queries):
    """Rows of one tree's queried values: the serialized layout is query-major, columns sorted by
    log size within a row (crates/cairo-air/src/utils.rs:221-266), i.e. the order the lifted leaf
    hash absorbs them."""
    vals = proof["queried_values"][tree]
    assert len(vals) % n_queries == 0
    w = len(vals) // n_queries
    return [vals[k * w:(k + 1) * w] for k in range(n_queries)]


def preprocessed_positions(queries, lifting_log, pp_log):
    """prepare_preprocessed_query_positions (crates/stwo/src/core/pcs/utils.rs:185-203)."""
    if lifting_log < pp_log:
        return [(p >> 1 << (pp_log - lifting_log + 1)) + (p & 1) for p in queries]
    return [(p >> (lifting_log - pp_log + 1) << 1) + (p & 1) for p in queries]


def merkle_witness_count(positions, height):
    """Number of hashes MerkleProverLifted::decommit sends for these leaf positions
    (crates/stwo/src/prover/vcs_lifted/prover.rs:118-152): one per node whose sibling is not known."""
    q = sorted(set(positions))
    n = 0
    for _ in range(height):
        nxt, i = [], 0
        while i < len(q):
            first = q[i]
            if i + 1 < len(q) and q[i + 1] == first ^ 1:
                i += 2
            else:
                n += 1
                i += 1
            nxt.append(first >> 1)
        q = nxt
    return n


def fri_opening_shape(queries, col_log, params):
    """(values in fri_witness, hashes in the Merkle decommitment) of the first and every inner FRI
    layer that FriProver::decommit_on_queries (crates/stwo/src/prover/fri.rs:253-284, 408-438) sends
    for these queries: the layer schedule of commit_inner_layers (fri.rs:145-196), cosets of
    2^fold_step positions, packed leaves when a layer folds more than once."""
    fs = params["fold_step"]
    last_log = params["log_last_layer_degree_bound"] + params["log_blowup_factor"]
    layers = [(col_log, fs, col_log >= 2 and fs > 1)]
    line_log = col_log - fs
    while line_log > last_log:
        k = fs if line_log > last_log + fs else line_log - last_log
        layers.append((line_log, k, line_log >= 2 and k > 1))
        line_log -= k
    shape, qs = [], sorted(set(queries))
    for log, k, pack in layers:
        cosets = sorted(set(q >> k for q in qs))
        positions = [p for c in cosets for p in range(c << k, (c + 1) << k)]
        n_values = len(positions) - len(qs)
        mpos = sorted(set(p >> 2 for p in positions)) if pack else positions
        shape.append((n_values, merkle_witness_count(mpos, log - 2 if pack else log)))
        qs = cosets
    return shape


# N_TRACE_COLUMNS of every component (crates/cairo-air/src/components/<name>.rs, the constant of
# that name, lines 5-11): the width of its block of the trace tree.  memory_address_to_id is
# MEMORY_ADDRESS_TO_ID_SPLIT * 2 (memory_address_to_id.rs:31-32), verify_bitwise_xor_12
# 1 << (2 * EXPAND_BITS) (verify_bitwise_xor_12.rs:4-8), memory_id_to_big BIG_N_COLUMNS =
# FELT252_N_WORDS + 1 per big log size (memory_id_to_big.rs:21-23, 217-223).
TRACE_COLUMNS = {
    "add_opcode": 103, "add_opcode_small": 39, "add_ap_opcode": 17, "assert_eq_opcode": 12,
    "assert_eq_opcode_imm": 9, "assert_eq_opcode_double_deref": 19, "blake_compress_opcode": 174,
    "call_opcode_abs": 25, "call_opcode_rel_imm": 24, "jnz_opcode_non_taken": 9, "jnz_opcode_taken": 47,
    "jump_opcode_abs": 14, "jump_opcode_double_deref": 21, "jump_opcode_rel": 16, "jump_opcode_rel_imm": 13,
    "mul_opcode": 130, "mul_opcode_small": 37, "qm_31_add_mul_opcode": 73, "ret_opcode": 16,
    "verify_instruction": 17, "blake_round": 212, "blake_g": 53, "blake_round_sigma": 1, "triple_xor_32": 21,
    "verify_bitwise_xor_12": 16, "add_mod_builtin": 267, "bitwise_builtin": 89, "mul_mod_builtin": 426,
    "pedersen_builtin": 3, "pedersen_builtin_narrow_windows": 3, "poseidon_builtin": 6,
    "range_check96_builtin": 12, "range_check_builtin": 17, "ec_op_builtin": 273, "partial_ec_mul_generic": 624,
    "pedersen_aggregator_window_bits_18": 206, "partial_ec_mul_window_bits_18": 297,
    "pedersen_points_table_window_bits_18": 1, "pedersen_aggregator_window_bits_9": 234,
    "partial_ec_mul_window_bits_9": 311, "pedersen_points_table_window_bits_9": 1, "poseidon_aggregator": 342,
    "poseidon_3_partial_rounds_chain": 169, "poseidon_full_round_chain": 126, "cube_252": 141,
    "poseidon_round_keys": 1, "range_check_252_width_27": 20, "memory_address_to_id": 32,
    "memory_id_to_big": FELT252_N_WORDS + 1, "memory_id_to_small": 9, "range_check_6": 1, "range_check_8": 1,
    "range_check_11": 1, "range_check_12": 1, "range_check_18": 2, "range_check_20": 8, "range_check_4_3": 1,
    "range_check_4_4": 1, "range_check_9_9": 8, "range_check_7_2_5": 1, "range_check_3_6_6_3": 1,
    "range_check_4_4_4_4": 1, "range_check_3_3_3_3_3": 1, "verify_bitwise_xor_4": 1, "verify_bitwise_xor_7": 1,
    "verify_bitwise_xor_8": 2, "verify_bitwise_xor_9": 1,
}


def component_instances(claims):
    """(name, log size) of every component the claim enables, in CairoClaim::log_sizes order
    (crates/cairo-air/src/claims.rs:383-465, the field order); memory_id_to_big once per big log
    size (components/memory_id_to_big.rs:217-240)."""
    out = []
    for name, kind, payload in claims:
        if payload is None:
            continue
        if kind == "log":
            out.append((name, payload))
        elif kind == "big":
            out += [(name, lg) for lg in payload]
        else:
            out.append((name, kind))
    return out


def column_log_sizes(proof):
    """Log sizes (before blowup) of the trace tree's and the interaction tree's columns, in column
    order (CairoClaim::log_sizes, claims.rs:383-465).  Trace tree: N_TRACE_COLUMNS per component.
    Interaction tree: a component's interaction columns end with the 4 coordinate columns of its
    last cumulative sum, which logup samples at offsets [-1, 0] (crates/constraint-framework/src/
    lib.rs:208-209) and every other interaction column at [0] (:201-202), so the blocks are read off
    the sample counts: each ends at its 4th consecutive column with two samples."""
    inst = component_instances(proof["claims"])
    trace = [lg for name, lg in inst for _ in range(TRACE_COLUMNS[name])]
    blocks, cur, twos = [], 0, 0
    for col in proof["sampled_values"][2]:
        cur += 1
        twos = twos + 1 if len(col) == 2 else 0
        if twos == 4:
            blocks.append(cur)
            cur = twos = 0
    assert cur == 0 and len(blocks) == len(inst), "interaction blocks do not match the components"
    interaction = [lg for (_, lg), k in zip(inst, blocks) for _ in range(k)]
    assert len(trace) == len(proof["sampled_values"][1])
    return trace, interaction


def columns_in_order(proof, tree, n_queries, col_logs):
    """Undoes sort_and_transpose_queried_values (crates/cairo-air/src/utils.rs:221-266): the
    per-column queried values, column order, from the rows (columns stably sorted by log size)."""
    rows = tree_rows(proof, tree, n_queries)
    order = sorted(range(len(col_logs)), key=lambda i: col_logs[i])
    cols = [None] * len(col_logs)
    for j, i in enumerate(order):
        cols[i] = [int(r[j]) for r in rows]
    return cols
