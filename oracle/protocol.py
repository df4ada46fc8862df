"""Restatement of the reference's host protocol over the C oracle: lifted-Merkle decommitment and
verification, FRI prover and verifier, DEEP-quotient driver, FriProof felt encoding.

TEST INFRASTRUCTURE ONLY (the checker for the C++ drivers); never imported by the product path.
Each function cites the reference function it restates.
"""
import numpy as np

import oracle as O

P = O.P


# ---------------------------------------------------------------- small QM31 helpers
def qadd(a, b):
    return [(int(x) + int(y)) % P for x, y in zip(a, b)]


def qsub(a, b):
    return [(int(x) - int(y)) % P for x, y in zip(a, b)]


def qmul(a, b):
    return [int(v) for v in O.qm31_mul([int(x) for x in a], [int(x) for x in b])]


def qinv(a):
    return [int(v) for v in O.qm31_inv([int(x) for x in a])]


def qmulm(a, m):
    return [int(x) * int(m) % P for x in a]


def bitrev(i, n):
    return int(bin(i)[2:].zfill(n)[::-1], 2) if n else i


def coset_at(init, log, i):
    """Coset(init, log).at(i) (crates/stwo/src/core/circle.rs:357-359)."""
    return O.index_to_point((init + (1 << (31 - log)) * i) % (1 << 31))


def circle_domain_at(half_init, half_log, i):
    return O.circle_domain_at(half_init, half_log, i)


# ---------------------------------------------------------------- Merkle (lifted)
def merkle_decommit(layers, positions):
    """Hash witness of MerkleProverLifted::decommit (crates/stwo/src/prover/vcs_lifted/prover.rs:
    118-152); layers root-first as returned by oracle.merkle_commit_lifted."""
    L = len(layers) - 1
    q = sorted(set(positions))
    witness = []
    for layer in range(L - 1, -1, -1):
        prev = layers[layer + 1]
        nxt = []
        i = 0
        while i < len(q):
            first = q[i]
            if i + 1 < len(q) and q[i + 1] == first ^ 1:
                i += 2
            else:
                witness.append(prev[first ^ 1].tobytes())
                i += 1
            nxt.append(first >> 1)
        q = nxt
    return witness


def merkle_verify(root, height, positions, rows, witness):
    """MerkleVerifierLifted::verify for columns of one size (crates/stwo/src/core/vcs_lifted/
    verifier.rs:115-212): rows[k] holds every column's value at positions[k]."""
    hashes = [O.blake2s(b"".join(int(v).to_bytes(4, "little") for v in row)) for row in rows]
    return merkle_verify_leaves(root, height, positions, hashes, witness)


def merkle_verify_leaves(root, height, positions, leaf_hashes, witness):
    """The same check from the leaf hashes of the queried rows (leaf_hashes[k] at positions[k])."""
    leaves = {}
    for pos, h in zip(positions, leaf_hashes):
        assert leaves.setdefault(pos, h) == h
    prev = sorted(leaves.items())
    wit = iter(witness)
    for _ in range(height):
        cur = []
        i = 0
        while i < len(prev):
            idx0, h0 = prev[i]
            if i + 1 < len(prev) and prev[i + 1][0] == idx0 ^ 1:
                children = (h0, prev[i + 1][1])
                i += 2
            else:
                w = next(wit, None)
                if w is None:
                    return False
                children = (h0, w) if idx0 & 1 == 0 else (w, h0)
                i += 1
            cur.append((idx0 >> 1, O.blake2s(children[0] + children[1])))
        prev = cur
    if next(wit, None) is not None:
        return False
    return len(prev) == 1 and prev[0][1] == root


# ---------------------------------------------------------------- FRI prover
def line_interpolate(values_bitrev, init, log):
    """LineEvaluation::interpolate (crates/stwo/src/prover/line.rs:54-64) + line_ifft
    (crates/stwo/src/prover/line.rs:88-99): coefficients in bit-reversed order."""
    v = [list(values_bitrev[bitrev(i, log)]) for i in range(1 << log)]
    d_init, d_log = init, log
    while d_log > 0:
        size = 1 << d_log
        for c in range(0, len(v), size):
            for i in range(size // 2):
                xi = O.m31_inv(int(coset_at(d_init, d_log, i)[0]))
                a, b = v[c + i], v[c + size // 2 + i]
                v[c + i], v[c + size // 2 + i] = qadd(a, b), qmulm(qsub(a, b), xi)
        d_init, d_log = (d_init * 2) % (1 << 31), d_log - 1
    inv = O.m31_inv(len(v))
    return [qmulm(x, inv) for x in v]


def squared_alpha_powers(alpha, n):
    """squared_alpha_powers (crates/stwo/src/prover/fri.rs:79-87): [a, a^2, a^4, ...], n terms."""
    out = []
    for _ in range(n):
        out.append(list(alpha))
        alpha = qmul(alpha, alpha)
    return out


def _pack_leaves_input(cols4):
    """PackLeavesOps::pack_leaves_input (crates/stwo/src/prover/backend/cpu/merkle_lifted.rs:101-124):
    column coord + 4*offset holds coordinate coord of the values at rows 4i + offset."""
    return [np.ascontiguousarray(np.asarray(cols4[c])[off::4]) for off in range(4) for c in range(4)]


def _commit_fri_layer(cols4, log, pack):
    """FriFirstLayerProver::new / FriInnerLayerProver::new (crates/stwo/src/prover/fri.rs:294-306,
    357-368): one leaf per value, or per four consecutive values when the leaves are packed."""
    if pack:
        return O.merkle_commit_lifted(_pack_leaves_input(cols4), log - 2)
    return O.merkle_commit_lifted([np.ascontiguousarray(c) for c in cols4], log)


def _fold_queries(queries, k):
    """Queries::fold (crates/stwo/src/core/queries.rs:48-55)."""
    return sorted(set(q >> k for q in queries))


def _decommit_layer(col, layers, queries, fold_step=1, pack=False):
    """compute_decommitment_positions_and_witness_evals (crates/stwo/src/prover/fri.rs:408-438):
    every coset of 2^fold_step positions holding a query is opened, the values the verifier cannot
    rebuild go to the witness; then the layer's Merkle decommitment (fri.rs:308-341), at positions
    divided by 4 when the leaves are packed."""
    positions, witness_evals = [], []
    i = 0
    while i < len(queries):
        j = i
        while j < len(queries) and queries[j] >> fold_step == queries[i] >> fold_step:
            j += 1
        start = (queries[i] >> fold_step) << fold_step
        for p in range(start, start + (1 << fold_step)):
            positions.append(p)
            if p not in queries[i:j]:
                witness_evals.append([int(col[c][p]) for c in range(4)])
        i = j
    mpos = sorted(set(p >> 2 for p in positions)) if pack else positions
    return {"fri_witness": witness_evals, "hash_witness": merkle_decommit(layers, mpos),
            "commitment": layers[0][0].tobytes()}


def fri_prove(column, log_size, params, ch):
    """FriProver::commit + decommit (crates/stwo/src/prover/fri.rs:113-284).
    column: (4, 2^log_size) QM31 evaluation on CanonicCoset(log_size).  Returns (proof, queries)."""
    fs = params["fold_step"]
    assert fs >= 1
    last_log = params["log_last_layer_degree_bound"] + params["log_blowup_factor"]
    pack0 = log_size >= 2 and fs > 1
    first_layers = _commit_fri_layer(column, log_size, pack0)
    ch.mix_hash(first_layers[0][0].tobytes())
    alpha = ch.draw_secure_felt()
    cur = O.fold_circle_into_line(column, log_size, alpha)
    line_log, line_init = log_size - 1, O.canonic_half_initial(log_size)

    def fold(cur, line_init, line_log, alphas):
        for a in alphas:
            cur = O.fold_line(cur, line_init, line_log, a)
            line_log, line_init = line_log - 1, (line_init * 2) % (1 << 31)
        return cur, line_init, line_log

    # the circle-to-line fold is followed by fold_step - 1 line folds with alpha^2, alpha^4, ...
    cur, line_init, line_log = fold(cur, line_init, line_log, squared_alpha_powers(qmul(alpha, alpha), fs - 1))
    assert line_log >= last_log, "The circle-to-line fold results in a smaller line domain than the last layer."
    inner = []
    while line_log > last_log:
        k = fs if line_log > last_log + fs else line_log - last_log
        pack = line_log >= 2 and k > 1
        layers = _commit_fri_layer(cur, line_log, pack)
        ch.mix_hash(layers[0][0].tobytes())
        alpha = ch.draw_secure_felt()
        inner.append((cur, layers, k, pack))
        cur, line_init, line_log = fold(cur, line_init, line_log, squared_alpha_powers(alpha, k))
    coeffs = line_interpolate([cur[:, i] for i in range(cur.shape[1])], line_init, line_log)
    ordered = [coeffs[bitrev(i, line_log)] for i in range(len(coeffs))]
    bound = 1 << params["log_last_layer_degree_bound"]
    assert all(not any(c) for c in ordered[bound:]), "invalid degree"
    kept = ordered[:bound]
    lb = bound.bit_length() - 1
    last_poly = [kept[bitrev(i, lb)] for i in range(bound)]
    ch.mix_felts(np.array(last_poly, np.uint32).reshape(-1))
    nonce = O.grind(ch.digest, params["pow_bits"])
    ch.mix_u64(nonce)
    raw = []
    while len(raw) < params["n_queries"]:
        for w in ch.draw_u32s():
            if len(raw) < params["n_queries"]:
                raw.append(int(w) & ((1 << log_size) - 1))
    queries = sorted(set(raw))
    first = _decommit_layer(column, first_layers, queries, fs, pack0)
    inner_proofs = []
    lq = _fold_queries(queries, fs)
    for col, layers, k, pack in inner:
        inner_proofs.append(_decommit_layer(col, layers, lq, k, pack))
        lq = _fold_queries(lq, k)
    proof = {"first_layer": first, "inner_layers": inner_proofs, "last_layer_poly": last_poly,
             "log_last_layer_degree_bound": params["log_last_layer_degree_bound"], "proof_of_work": nonce}
    return proof, queries


def serialize_fri_proof(proof):
    """CairoSerde felt encoding (crates/cairo-serialize/src/serialize.rs:21-90, 167-211)."""
    out = []

    def layer(lp):
        out.append(len(lp["fri_witness"]))
        for q in lp["fri_witness"]:
            out.extend(int(x) for x in q)
        out.append(len(lp["hash_witness"]))
        for h in lp["hash_witness"]:
            out.extend(int.from_bytes(h[4 * i:4 * i + 4], "little") for i in range(8))
        out.extend(int.from_bytes(lp["commitment"][4 * i:4 * i + 4], "little") for i in range(8))

    layer(proof["first_layer"])
    out.append(len(proof["inner_layers"]))
    for lp in proof["inner_layers"]:
        layer(lp)
    out.append(len(proof["last_layer_poly"]))
    for q in proof["last_layer_poly"]:
        out.extend(int(x) for x in q)
    out.append(proof["log_last_layer_degree_bound"])
    out.append(proof["proof_of_work"])
    return out


def deserialize_fri_proof(felts):
    it = iter(felts)

    def hash_():
        return b"".join(int(next(it)).to_bytes(4, "little") for _ in range(8))

    def layer():
        fw = [[next(it) for _ in range(4)] for _ in range(next(it))]
        hw = [hash_() for _ in range(next(it))]
        return {"fri_witness": fw, "hash_witness": hw, "commitment": hash_()}

    first = layer()
    inner = [layer() for _ in range(next(it))]
    lp = [[next(it) for _ in range(4)] for _ in range(next(it))]
    lb = next(it)
    pow_ = next(it)
    assert next(it, None) is None, "trailing felts"
    return {"first_layer": first, "inner_layers": inner, "last_layer_poly": lp,
            "log_last_layer_degree_bound": lb, "proof_of_work": pow_}


# ---------------------------------------------------------------- FRI verifier
def _line_eval_at_point(coeffs_bitrev, x):
    """LinePoly::eval_at_point (crates/stwo/src/core/poly/line.rs:131-139)."""
    doublings = []
    for _ in range(len(coeffs_bitrev).bit_length() - 1):
        doublings.append(x)
        s = qmul(x, x)
        x = qsub(qadd(s, s), [1, 0, 0, 0])
    vals = [list(c) for c in coeffs_bitrev]
    for f in reversed(doublings):
        vals = [qadd(vals[2 * t], qmul(vals[2 * t + 1], f)) for t in range(len(vals) // 2)]
    return vals[0]


def _circle_domain_index(half_init, half_log, i):
    """CircleDomain::index_at (crates/stwo/src/core/poly/circle/domain.rs:50-60): the half coset,
    then its negation."""
    if i < 1 << half_log:
        return (half_init + (1 << (31 - half_log)) * i) % (1 << 31)
    return ((1 << 31) - (half_init + (1 << (31 - half_log)) * (i - (1 << half_log)))) % (1 << 31)


def fold_coset(evals, init_idx, log, alpha):
    """fold_coset (crates/stwo/src/core/fri.rs:766-786): log line folds of the values on
    Coset(init_idx, log) (bit-reversed), alpha squared after each fold."""
    ev = [list(e) for e in evals]
    a = list(alpha)
    for i in range(log):
        dl, di = log - i, (init_idx << i) % (1 << 31)
        for j in range(0, 1 << dl, 2):
            xi = O.m31_inv(int(coset_at(di, dl, bitrev(j, dl))[0]))
            f0, f1 = qadd(ev[j], ev[j + 1]), qmulm(qsub(ev[j], ev[j + 1]), xi)
            ev[j >> 1] = qadd(f0, qmul(a, f1))
        a = qmul(a, a)
    return ev[0]


def fold_circle(evals, init_idx, fold_step, alpha, fold_coset_fn=None):
    """SparseEvaluation::fold_circle for one subset (crates/stwo/src/core/fri.rs:725-756): the
    circle-to-line fold on CircleDomain(Coset(init_idx, fold_step - 1)), then fold_step - 1 line
    folds with alpha^2, alpha^4, ... (by fold_coset_fn when given)."""
    half = fold_step - 1
    buf = []
    for i in range(1 << half):
        k = bitrev(2 * i, fold_step)
        y = int(O.index_to_point(_circle_domain_index(init_idx, half, k))[1])
        yi = O.m31_inv(y)
        f0, f1 = qadd(evals[2 * i], evals[2 * i + 1]), qmulm(qsub(evals[2 * i], evals[2 * i + 1]), yi)
        buf.append(qadd(qmul(alpha, f1), f0))
    return buf[0] if fold_step == 1 else (fold_coset_fn or fold_coset)(buf, init_idx, half, qmul(alpha, alpha))


def fri_verify(proof, params, ch, log_degree_bound, first_layer_query_evals_fn, fold_circle_fn=None,
               fold_coset_fn=None):
    """FriVerifier::commit / verify_proof_of_work / sample_query_positions / decommit
    (crates/stwo/src/core/fri.rs:156-378, 480-700).  first_layer_query_evals_fn(queries) returns the
    column's QM31 values at the sorted queries.  fold_circle_fn / fold_coset_fn (same arguments as
    fold_circle / fold_coset) replace the folds of a subset, for checking another implementation
    of them.  Returns the queries, raises on failure."""
    fold_circle_fn = fold_circle_fn or fold_circle
    fold_coset_fn = fold_coset_fn or fold_coset
    b, fs = params["log_blowup_factor"], params["fold_step"]
    ch.mix_hash(proof["first_layer"]["commitment"])
    col_log = log_degree_bound + b
    alpha0 = ch.draw_secure_felt()
    pack0 = col_log >= 2 and fs > 1
    assert log_degree_bound >= fs, "InvalidNumFriLayers"
    layer_bound = log_degree_bound - fs
    line_init, line_log = O.canonic_half_initial(layer_bound + b + 1), layer_bound + b
    inner = []
    n_inner = len(proof["inner_layers"])
    for k, lp in enumerate(proof["inner_layers"]):
        ch.mix_hash(lp["commitment"])
        step = fs
        if k == n_inner - 1:
            step = layer_bound - params["log_last_layer_degree_bound"]
            assert 1 <= step <= fs, "InvalidNumFriLayers"
        assert layer_bound >= step, "InvalidNumFriLayers"
        inner.append((line_init, line_log, ch.draw_secure_felt(), lp, step, line_log >= 2 and step > 1))
        layer_bound -= step
        line_init, line_log = (line_init << step) % (1 << 31), line_log - step
    assert layer_bound == params["log_last_layer_degree_bound"], "InvalidNumFriLayers"
    assert len(proof["last_layer_poly"]) <= 1 << params["log_last_layer_degree_bound"]
    ch.mix_felts(np.array(proof["last_layer_poly"], np.uint32).reshape(-1))
    assert O.verify_pow_nonce(ch.digest, params["pow_bits"], proof["proof_of_work"]), "ProofOfWork"
    ch.mix_u64(proof["proof_of_work"])
    raw = []
    while len(raw) < params["n_queries"]:
        for w in ch.draw_u32s():
            if len(raw) < params["n_queries"]:
                raw.append(int(w) & ((1 << col_log) - 1))
    queries = sorted(set(raw))

    def rebuild(qs, evals, lp, log, step, pack):
        """compute_decommitment_positions_and_rebuild_evals + build_merkle_verification_inputs
        (fri.rs:629-700), then the layer's Merkle check."""
        wit = iter(lp["fri_witness"])
        ev = iter(evals)
        positions, subsets = [], []
        i = 0
        while i < len(qs):
            j = i
            while j < len(qs) and qs[j] >> step == qs[i] >> step:
                j += 1
            start = (qs[i] >> step) << step
            sub = []
            for p in range(start, start + (1 << step)):
                positions.append(p)
                if p in qs[i:j]:
                    sub.append(list(next(ev)))
                else:
                    w = next(wit, None)
                    assert w is not None, "EvaluationsInvalid"
                    sub.append(list(w))
            subsets.append((start, sub))
            i = j
        assert next(wit, None) is None, "EvaluationsInvalid"
        values = [v for _, sub in subsets for v in sub]
        if pack:
            mpos = sorted(set(p >> 2 for p in positions))
            rows = [[x for v in values[4 * k:4 * k + 4] for x in v] for k in range(len(mpos))]
            ok = merkle_verify(lp["commitment"], log - 2, mpos, rows, lp["hash_witness"])
        else:
            ok = merkle_verify(lp["commitment"], log, positions, values, lp["hash_witness"])
        assert ok, "CommitmentInvalid"
        return subsets

    subsets = rebuild(queries, first_layer_query_evals_fn(queries), proof["first_layer"], col_log, fs, pack0)
    half_init = O.canonic_half_initial(col_log)
    evals = [fold_circle_fn(sub, _circle_domain_index(half_init, col_log - 1, bitrev(start, col_log)), fs, alpha0)
             for start, sub in subsets]
    qs = _fold_queries(queries, fs)
    for li, llog, alpha, lp, step, pack in inner:
        subsets = rebuild(qs, evals, lp, llog, step, pack)
        evals = [fold_coset_fn(sub, (li + (1 << (31 - llog)) * bitrev(start, llog)) % (1 << 31), step, alpha)
                 for start, sub in subsets]
        qs = _fold_queries(qs, step)
    # last layer (fri.rs:320-334)
    for q, e in zip(qs, evals):
        x = coset_at(line_init, line_log, bitrev(q, line_log))[0]
        assert e == _line_eval_at_point(proof["last_layer_poly"], [int(x), 0, 0, 0]), "LastLayerEvaluationsInvalid"
    return queries


# ---------------------------------------------------------------- DEEP quotients
def _build_samples(column_logs, samples, random_coeff, L):
    """build_samples_with_randomness_and_periodicity (crates/stwo/src/core/pcs/quotients.rs:
    274-319): per column, [(point[8], value[4], random power)]; a column with two samples gets a
    periodicity sample first, at the second point shifted by CanonicCoset(L).step() doubled
    log_size times."""
    pw = [1, 0, 0, 0]
    swr = []
    for log, smp in zip(column_logs, samples):
        out = []
        if len(smp) == 2:
            g = O.index_to_point(((1 << (31 - L)) << log) % (1 << 31))
            pt = [int(x) for x in smp[1][0]]
            x, y = pt[:4], pt[4:]
            gx, gy = [int(g[0]), 0, 0, 0], [int(g[1]), 0, 0, 0]
            sx = qsub(qmul(x, gx), qmul(y, gy))
            sy = qadd(qmul(x, gy), qmul(y, gx))
            out.append((sx + sy, [int(v) for v in smp[1][1]], pw))
            pw = qmul(pw, random_coeff)
        for pt, v in smp:
            out.append(([int(x) for x in pt], [int(x) for x in v], pw))
            pw = qmul(pw, random_coeff)
        swr.append(out)
    return swr


def _batches(swr):
    """ColumnSampleBatch::new_vec (crates/stwo/src/core/pcs/quotients.rs:78-95): group by point in
    order of first occurrence; terms (column index, value, random power)."""
    batches = {}
    for ci, smp in enumerate(swr):
        for pt, v, r in smp:
            batches.setdefault(tuple(pt), []).append((ci, v, r))
    return batches


def _cmul(a, b):
    return [(a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P]


def _cinv(a):
    n = O.m31_inv((a[0] * a[0] + a[1] * a[1]) % P)
    return [a[0] * n % P, (P - a[1]) * n % P]


def lifted_value(col, L, pos):
    """A column of 2^n values read at a position of the 2^L lifted domain
    (crates/stwo/src/prover/vcs_lifted/prover.rs:108-116)."""
    n = len(col).bit_length() - 1
    return int(col[((pos >> (L - n + 1)) << 1) + (pos & 1)])


def fri_answers(column_logs, samples, random_coeff, positions, queried_values, L):
    """fri_answers + accumulate_row_quotients + denominator_inverses (crates/stwo/src/core/pcs/
    quotients.rs:115-258): the DEEP quotient at each query position of the lifted domain, from
    the columns' lifted values there (queried_values[i][k]: column i at positions[k])."""
    batches = _batches(_build_samples(column_logs, samples, random_coeff, L))
    consts = {pt: [O.line_coeffs(list(pt), v, r) for _, v, r in terms] for pt, terms in batches.items()}
    half = O.canonic_half_initial(L)
    res = []
    for k, pos in enumerate(positions):
        px, py = (int(c) for c in circle_domain_at(half, L - 1, bitrev(pos, L)))
        acc = [0, 0, 0, 0]
        for pt, terms in batches.items():
            num = [0, 0, 0, 0]
            for (ci, _, _), co in zip(terms, consts[pt]):
                a, bb, c = [int(x) for x in co[:4]], [int(x) for x in co[4:8]], [int(x) for x in co[8:]]
                val = qmulm(c, queried_values[ci][k])
                num = qadd(num, qsub(val, qadd(qmulm(a, py), bb)))
            # (Re(p.x) - px) * Im(p.y) - (Re(p.y) - py) * Im(p.x), an element of CM31
            t0 = _cmul([(pt[0] - px) % P, pt[1]], [pt[6], pt[7]])
            t1 = _cmul([(pt[4] - py) % P, pt[5]], [pt[2], pt[3]])
            den = _cinv([(t0[0] - t1[0]) % P, (t0[1] - t1[1]) % P])
            acc = qadd(acc, _cmul(num[:2], den) + _cmul(num[2:], den))
        res.append(acc)
    return res


def fri_quotients(columns, samples, random_coeff, L, b, tw, itw):
    """compute_fri_quotients (crates/stwo/src/prover/pcs/quotient_ops.rs:86-160) with
    build_samples_with_randomness_and_periodicity (crates/stwo/src/core/pcs/quotients.rs:274-319).
    columns: evaluations (uint32 arrays); samples[i]: list of (point[8], value[4])."""
    swr = _build_samples([len(c).bit_length() - 1 for c in columns], samples, random_coeff, L)
    accs = []
    for log in sorted(set(len(c).bit_length() - 1 for c in columns)):
        idx = [i for i, c in enumerate(columns) if len(c).bit_length() - 1 == log]
        for pt, terms in _batches([swr[i] for i in idx]).items():
            coeffs = np.stack([O.line_coeffs(list(pt), v, r) for _, v, r in terms])
            part, fl = O.accumulate_numerators([columns[i] for i in idx], [ci for ci, _, _ in terms], coeffs, b)
            accs.append((list(pt), part, fl))
    accs.sort(key=lambda a: tuple(a[0]))
    per_point = []
    i = 0
    while i < len(accs):
        j = i
        while j < len(accs) and accs[j][0] == accs[i][0]:
            j += 1
        fl = [0, 0, 0, 0]
        for a in accs[i:j]:
            fl = qadd(fl, a[2])
        part = O.lift_and_accumulate([a[1] for a in accs[i:j]])
        per_point.append((accs[i][0], part, fl))
        i = j
    return O.compute_quotients_and_combine([p[1] for p in per_point], [p[0] for p in per_point],
                                           [p[2] for p in per_point], L, b, tw, itw)


# ---------------------------------------------------------------- wide-Fibonacci composition
def _m31_point_add(p, q):
    return ((p[0] * q[0] - p[1] * q[1]) % P, (p[0] * q[1] + p[1] * q[0]) % P)


def coset_vanishing_m31(coset_initial_idx, coset_log, p):
    """coset_vanishing (crates/stwo/src/core/constraints.rs:12-35) at an M31 point, with point
    arithmetic: p - initial + step/2, then x doubled coset_log - 1 times."""
    init = O.index_to_point(coset_initial_idx)
    half_step = O.index_to_point((1 << (31 - coset_log)) >> 1)
    q = _m31_point_add((int(p[0]), int(p[1])), (int(init[0]), (P - int(init[1])) % P))
    q = _m31_point_add(q, (int(half_step[0]), int(half_step[1])))
    x = q[0]
    for _ in range(1, coset_log):
        x = (2 * x * x - 1) % P
    return x


def wide_fib_composition(cols, trace_log, eval_log, log_expansion, powers):
    """evaluate_constraint_quotients_on_domain for WideFibonacciEval
    (crates/constraint-framework/src/prover/component_prover.rs:77-193, SubDomain mode;
    crates/examples/src/wide_fibonacci/mod.rs:87-98): the 4 coordinate columns of
    sum_j powers[j] * (c_{j+2} - c_j^2 - c_{j+1}^2) * denom_inv[row >> trace_log]."""
    n = 1 << eval_log
    c = [np.asarray(col[:n], np.uint64) for col in cols]
    acc = np.zeros((4, n), np.uint64)
    for j in range(len(cols) - 2):
        k = (c[j + 2] + 2 * P - (c[j] * c[j] % P) - (c[j + 1] * c[j + 1] % P)) % P
        for t in range(4):
            acc[t] = (acc[t] + k * np.uint64(int(powers[j][t])) % P) % P
    # denominators on the evaluation domain CanonicCoset(eval_log + e).circle_domain().split(e).0
    half_init = 1 << (31 - (eval_log + log_expansion + 1))
    n_den = 1 << (eval_log - trace_log)
    den = [0] * n_den
    for i in range(n_den):
        p = O.circle_domain_at(half_init, eval_log - 1, i)
        v = coset_vanishing_m31(1 << (31 - (trace_log + 1)), trace_log, p)
        den[bitrev(i, eval_log - trace_log)] = O.m31_inv(v)
    row_d = np.array(den, np.uint64)[np.arange(n) >> trace_log]
    return (acc * row_d % P).astype(np.uint32)


def from_partial_evals(e):
    """SecureField::from_partial_evals: e0 + e1*i + e2*u + e3*iu (crates/stwo/src/core/fields/qm31.rs)."""
    out = [0, 0, 0, 0]
    for t in range(4):
        basis = [0, 0, 0, 0]
        basis[t] = 1
        out = qadd(out, qmul(e[t], basis))
    return out


def wide_fib_composition_at_point(mask_values, alpha, point, trace_log):
    """FrameworkComponent::evaluate_constraint_quotients_at_point (crates/constraint-framework/src/
    component.rs:220-243) with PointEvaluationAccumulator (crates/stwo/src/core/air/accumulation.rs:
    18-33): Horner over the constraints in order, each times 1/coset_vanishing(trace coset, point).
    mask_values[k] is trace column k's value at the point (QM31)."""
    x, y = [int(v) for v in point[:4]], [int(v) for v in point[4:]]
    init = O.index_to_point(1 << (31 - (trace_log + 1)))
    half_step = O.index_to_point((1 << (31 - trace_log)) >> 1)
    # point - init + step/2 over QM31 x M31 points
    def add_m31(px, py, q):
        return (qsub(qmulm(px, q[0]), qmulm(py, q[1])), qadd(qmulm(px, q[1]), qmulm(py, q[0])))
    px, py = add_m31(x, y, (int(init[0]), (P - int(init[1])) % P))
    px, py = add_m31(px, py, (int(half_step[0]), int(half_step[1])))
    v = px
    for _ in range(1, trace_log):
        s = qmul(v, v)
        v = qsub(qadd(s, s), [1, 0, 0, 0])
    dinv = qinv(v)
    acc = [0, 0, 0, 0]
    for j in range(len(mask_values) - 2):
        a, b, c = mask_values[j], mask_values[j + 1], mask_values[j + 2]
        cons = qsub(c, qadd(qmul(a, a), qmul(b, b)))
        acc = qadd(qmul(acc, alpha), qmul(dinv, cons))
    return acc


# ---------------------------------------------------------------- PCS and prove / verify
def _hash_words(h):
    return [int.from_bytes(h[4 * i:4 * i + 4], "little") for i in range(8)]


def _words_hash(w):
    return b"".join(int(x).to_bytes(4, "little") for x in w)


def _double_point(pt):
    x, y = list(pt[:4]), list(pt[4:])
    xx = qmul(x, x)
    return qsub(qadd(xx, xx), [1, 0, 0, 0]) + qmulm(qmul(x, y), 2)


def _repeated_double(pt, n):
    for _ in range(n):
        pt = _double_point(pt)
    return pt


def pp_query_positions(query_positions, max_log_size, pp_max_log_size):
    """prepare_preprocessed_query_positions (crates/stwo/src/core/pcs/utils.rs:185-203)."""
    if pp_max_log_size == 0:
        return []
    if max_log_size < pp_max_log_size:
        return [((p >> 1) << (pp_max_log_size - max_log_size + 1)) + (p & 1) for p in query_positions]
    return [((p >> (max_log_size - pp_max_log_size + 1)) << 1) + (p & 1) for p in query_positions]


class OracleTree:
    """CommitmentTreeProver on the oracle: coefficients, evaluations (blown up), lifted tree."""

    def __init__(self, coeffs, evals, lifting_log_size):
        self.coeffs, self.evals, self.L = coeffs, evals, lifting_log_size
        if evals:
            self.layers = O.merkle_commit_lifted(evals, lifting_log_size)
        else:
            assert lifting_log_size == 0
            self.layers = [np.frombuffer(O.blake2s(b""), np.uint8).reshape(1, 32)]
        self.root = self.layers[0][0].tobytes()

    def decommit(self, positions):
        queried = [[lifted_value(e, self.L, p) for p in positions] for e in self.evals]
        witness = merkle_decommit(self.layers, positions) if self.L > 0 else []
        return queried, witness


def pcs_prove_values(trees, sampled_points, ch, fri, tw, itw, oods_pow):
    """CommitmentSchemeProver::prove_values (crates/stwo/src/prover/pcs/mod.rs:224-315) on the
    oracle; returns the proof dict (same fields as the device driver's)."""
    L = trees[-1].L
    b = fri["log_blowup_factor"]
    sampled_values, flat_cols, flat_samples = [], [], []
    for tree, pts_tree in zip(trees, sampled_points):
        vt = []
        for c, e, pts in zip(tree.coeffs, tree.evals, pts_tree):
            n_e = len(e).bit_length() - 1
            vals = [[int(v) for v in O.eval_at_point(c, _repeated_double(p, L - n_e))] for p in pts]
            vt.append(vals)
            flat_cols.append(e)
            flat_samples.append(list(zip([list(p) for p in pts], vals)))
        sampled_values.append(vt)
    ch.mix_felts(np.array([v for t in sampled_values for col in t for v in col], np.uint32).reshape(-1))
    rc = [int(v) for v in ch.draw_secure_felt()]
    quot = fri_quotients(flat_cols, flat_samples, rc, L, b, tw, itw)
    fproof, queries = fri_prove(quot, L, fri, ch)
    pp = pp_query_positions(queries, L, trees[0].L)
    qv, dec = [], []
    for i, t in enumerate(trees):
        q, w = t.decommit(pp if i == 0 else queries)
        qv.append(q)
        dec.append(w)
    return {"fri_params": dict(fri), "commitments": [t.root for t in trees], "oods_proof_of_work": oods_pow,
            "sampled_values": sampled_values, "decommitments": dec, "queried_values": qv,
            "fri_felts": serialize_fri_proof(fproof), "query_positions": queries}


def serialize_pcs_proof(proof):
    """CairoSerialize for CommitmentSchemeProof (crates/cairo-serialize/src/serialize.rs:99-143)."""
    f = proof["fri_params"]
    out = [f["pow_bits"], f["log_blowup_factor"], f["log_last_layer_degree_bound"], f["n_queries"], f["fold_step"]]
    out.append(len(proof["commitments"]))
    for h in proof["commitments"]:
        out.extend(_hash_words(h))
    out.append(proof["oods_proof_of_work"])
    out.append(len(proof["sampled_values"]))
    for t in proof["sampled_values"]:
        out.append(len(t))
        for col in t:
            out.append(len(col))
            for v in col:
                out.extend(int(x) for x in v)
    out.append(len(proof["decommitments"]))
    for w in proof["decommitments"]:
        out.append(len(w))
        for h in w:
            out.extend(_hash_words(h))
    out.append(len(proof["queried_values"]))
    for t in proof["queried_values"]:
        out.append(len(t))
        for col in t:
            out.append(len(col))
            out.extend(int(x) for x in col)
    out.extend(proof["fri_felts"])
    return out


def deserialize_pcs_proof(felts):
    it = iter(felts)
    nx = lambda: int(next(it))
    fri = {k: nx() for k in ("pow_bits", "log_blowup_factor", "log_last_layer_degree_bound", "n_queries",
                             "fold_step")}
    commitments = [_words_hash([nx() for _ in range(8)]) for _ in range(nx())]
    oods_pow = nx()
    sampled = [[[[nx() for _ in range(4)] for _ in range(nx())] for _ in range(nx())] for _ in range(nx())]
    dec = [[_words_hash([nx() for _ in range(8)]) for _ in range(nx())] for _ in range(nx())]
    qv = [[[nx() for _ in range(nx())] for _ in range(nx())] for _ in range(nx())]
    fri_felts = list(it)
    return {"fri_params": fri, "commitments": commitments, "oods_proof_of_work": oods_pow, "sampled_values": sampled,
            "decommitments": dec, "queried_values": qv, "fri_proof": deserialize_fri_proof(fri_felts)}


def _verify_tree(root, height, col_logs, positions, queried, witness):
    """MerkleVerifierLifted::verify (crates/stwo/src/core/vcs_lifted/verifier.rs:115-212)."""
    if height == 0:
        return True
    order = sorted(range(len(col_logs)), key=lambda i: col_logs[i])
    rows = [[queried[i][k] for i in order] for k in range(len(positions))]
    return merkle_verify(root, height, positions, rows, witness)


def pcs_verify_values(tree_roots, tree_heights, tree_col_logs, sampled_points, proof, ch, fri, answers_fn=None,
                      **fold_fns):
    """CommitmentSchemeVerifier::verify_values (crates/stwo/src/core/pcs/verifier.rs:57-130).
    tree_col_logs: extended (blown-up) log sizes per tree. Raises AssertionError on failure.
    answers_fn (same arguments as fri_answers) and fold_fns (fri_verify's fold_circle_fn,
    fold_coset_fn) replace the DEEP quotients and the folds, for checking other implementations."""
    answers_fn = answers_fn or fri_answers
    ch.mix_felts(np.array([v for t in proof["sampled_values"] for col in t for v in col],
                          np.uint32).reshape(-1))
    rc = [int(v) for v in ch.draw_secure_felt()]
    L = tree_heights[-1]
    b = fri["log_blowup_factor"]
    flat_logs, flat_samples, flat_cols_q = [], [], []
    for logs_t, pts_t, vals_t in zip(tree_col_logs, sampled_points, proof["sampled_values"]):
        for lg, pts, vals in zip(logs_t, pts_t, vals_t):
            flat_logs.append(lg)
            flat_samples.append(list(zip(pts, vals)))
    state = {}

    def first_layer_evals(queries):
        pp = pp_query_positions(queries, L, tree_heights[0])
        for i, (root, h, logs) in enumerate(zip(tree_roots, tree_heights, tree_col_logs)):
            pos = pp if i == 0 else queries
            assert _verify_tree(root, h, logs, pos, proof["queried_values"][i], proof["decommitments"][i]), \
                f"tree {i} decommitment invalid"
        qv = [col for t in proof["queried_values"] for col in t]
        return answers_fn(flat_logs, flat_samples, rc, queries, qv, L)

    fri_verify(proof["fri_proof"], fri, ch, L - b, first_layer_evals, **fold_fns)
    return True


def wide_fib_prove(trace, log_n, fri, tw, itw):
    """The prover side of test_wide_fib_prove_with_blake (crates/examples/src/wide_fibonacci/
    mod.rs:183-220) through prove_ex (crates/stwo/src/prover/mod.rs:38-135), on the oracle.
    Requires log_blowup_factor 1. Returns (felts, proof dict)."""
    b = fri["log_blowup_factor"]
    assert b == 1
    L = log_n + b
    ch = O.Channel()
    trees = [OracleTree([], [], 0)]
    ch.mix_hash(trees[0].root)
    coeffs = [O.interpolate(itw, c, log_n) for c in trace]
    trees.append(OracleTree(coeffs, [O.evaluate(tw, c, L) for c in coeffs], L))
    ch.mix_hash(trees[1].root)
    alpha = [int(v) for v in ch.draw_secure_felt()]
    m = len(trace) - 2
    pw = [[1, 0, 0, 0]]
    for _ in range(m - 1):
        pw.append(qmul(pw[-1], alpha))
    comp = wide_fib_composition(trees[1].evals, log_n, log_n + 1, 0, pw[::-1])
    comp_coeffs = [O.interpolate(itw, comp[t], log_n + 1) for t in range(4)]
    half = 1 << log_n
    polys = [c[:half] for c in comp_coeffs] + [c[half:] for c in comp_coeffs]   # split_at_mid
    trees.append(OracleTree(polys, [O.evaluate(tw, p, L) for p in polys], L))
    ch.mix_hash(trees[2].root)
    oods_pow = O.grind(ch.digest, 16)
    ch.mix_u64(oods_pow)
    z = [int(v) for v in ch.random_point()]
    pts = [[], [[z] for _ in trace], [[z] for _ in range(8)]]
    proof = pcs_prove_values(trees, pts, ch, fri, tw, itw, oods_pow)
    return serialize_pcs_proof(proof), proof


def wide_fib_verify(felts, n_cols, log_n, fri):
    """verify (crates/stwo/src/core/verifier.rs:19-119) for one wide-Fibonacci component, after
    the commitment-scheme setup of test_wide_fib_prove_with_blake (preprocessed tree empty and
    not lifted; trace lifting log_n + blowup). Raises AssertionError on a bad proof."""
    proof = deserialize_pcs_proof(felts)
    assert proof["fri_params"] == dict(fri), "config mismatch"
    b = fri["log_blowup_factor"]
    L = log_n + b
    roots = proof["commitments"]
    assert len(roots) == 3, "expected preprocessed, trace and composition trees"
    ch = O.Channel()
    ch.mix_hash(roots[0])
    ch.mix_hash(roots[1])
    alpha = [int(v) for v in ch.draw_secure_felt()]
    split_bound = (log_n + 1) - 1                      # composition_log_degree_bound - COMPOSITION_LOG_SPLIT
    L = max(L, split_bound + b)
    max_log_degree_bound = L - b
    ch.mix_hash(roots[2])
    assert O.verify_pow_nonce(ch.digest, 16, proof["oods_proof_of_work"]), "OodsProofOfWork"
    ch.mix_u64(proof["oods_proof_of_work"])
    z = [int(v) for v in ch.random_point()]
    pts = [[], [[z] for _ in range(n_cols)], [[z] for _ in range(8)]]
    sv = proof["sampled_values"]
    assert [len(t) for t in sv] == [0, n_cols, 8] and all(len(c) == 1 for t in sv[1:] for c in t), "structure"
    comp = [c[0] for c in sv[2]]
    left = from_partial_evals(comp[:4])
    right = from_partial_evals(comp[4:])
    zx = _repeated_double(z, max_log_degree_bound - 1)[:4]
    lhs = qadd(left, qmul(zx, right))                   # extract_composition_oods_eval
    rhs = wide_fib_composition_at_point([c[0] for c in sv[1]], alpha, z, max_log_degree_bound)
    assert lhs == rhs, "OodsNotMatching"
    heights = [0, L, L]
    col_logs = [[], [log_n + b] * n_cols, [max_log_degree_bound + b] * 8]
    return pcs_verify_values(roots, heights, col_logs, pts, proof, ch, fri)
