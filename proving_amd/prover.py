"""Host protocol above the HIP backend: the commitment scheme prover and `prove()` for the
wide-Fibonacci AIR, with the proof written as CairoSerde felts.

Restates, over device-resident columns:
- CommitmentSchemeProver::commit / prove_values (crates/stwo/src/prover/pcs/mod.rs:81-315);
- CommitmentTreeProver::new / decommit (crates/stwo/src/prover/pcs/mod.rs:360-399);
- prove_ex (crates/stwo/src/prover/mod.rs:38-135) for one FrameworkComponent of
  WideFibonacciEval (crates/examples/src/wide_fibonacci/mod.rs:77-98);
- CommitmentSchemeProof / StarkProof felt encoding (crates/cairo-serialize/src/serialize.rs:99-153).

Every kernel runs on the device; the host handles the channel, a few QM31 scalars, and the
small proof vectors. There is no CPU fallback: without libstwo_hip.so nothing here imports.
"""
import ctypes
import threading
import weakref

import numpy as np

from . import field as F
from .backend import (BaseColumn, Channel, ColumnRefs, ColumnSet, DeviceBuffer, DeviceChannel, FriParams, Graph,
                      HashLayerSet, SecureColumnByCoords, _logs, _ptr_array, fri_commit_dch, fri_prove,
                      fri_quotients_arrays, fri_quotients_dpoint, tree_decommit)

P = F.P
OODS_POW_BITS = 16  # crates/stwo/src/core/proof_of_work.rs:8
COMPOSITION_LOG_SPLIT = 1  # crates/stwo/src/core/verifier.rs:17
# Proofs keep their channel in device memory by default (DeviceChannel): from the trace root to
# the FRI commit phase the host never waits for the device.  False: the host channel throughout.
DEVICE_CHANNEL = True
# With the device channel, a trace proved again from the same buffers replays a captured hipGraph
# of the proof's whole device work (trace commit, composition and its commit, OODS proof of work
# and samples, DEEP quotients, FRI commit phase and proof of work: _ProofGraph); a trace committed
# through start_wide_fibonacci replays a captured graph of the batched commit
# (CommitmentSchemeProver._commit_captured).
GRAPH_COMMITS = True


def _mix_root(channel, layers, root):
    """MerkleChannel::mix_root: on the device from the tree's root layer when the channel is a
    DeviceChannel (the root need not be known on the host), else on the host."""
    if isinstance(channel, DeviceChannel):
        channel.mix_words(layers[0].ptr, 8)
    else:
        channel.mix_root(root)


class ConstraintsNotSatisfied(Exception):
    """ProvingError::ConstraintsNotSatisfied (crates/stwo/src/prover/mod.rs:137-140): the
    composition value rebuilt from the sampled composition coordinates differs from the
    constraints evaluated on the sampled trace values at the OODS point."""


def _from_partial_evals(e):
    """SecureField::from_partial_evals (crates/stwo/src/core/fields/qm31.rs): e0 + e1 i + e2 u +
    e3 iu, each e_t a QM31."""
    out = [0, 0, 0, 0]
    for t in range(4):
        basis = [0, 0, 0, 0]
        basis[t] = 1
        out = F.qadd(out, F.qmul(e[t], basis))
    return out


def _doubled_x(x, n):
    """x-coordinate of a point doubled n times (x -> 2x^2 - 1 needs only x)."""
    for _ in range(n):
        x = F.double_x(x)
    return x


def composition_oods_eval(comp_samples, oods_point, max_log_degree_bound):
    """StarkProof::extract_composition_oods_eval (crates/stwo/src/core/proof.rs:27-53): the eight
    sampled coordinate values of the split composition polynomial give left + x * right, with x
    the x-coordinate of the OODS point doubled max_log_degree_bound - 1 times (the split_at_mid
    factor)."""
    left = _from_partial_evals(comp_samples[:4])
    right = _from_partial_evals(comp_samples[4:])
    x = _doubled_x(list(oods_point[:4]), max_log_degree_bound - 1)
    return F.qadd(left, F.qmul(x, right))


def _qmul_rows(x, y):
    """QM31 products of the rows of two (n, 4) uint64 arrays of canonical words (u^2 = 2 + i)."""
    def cmul(a0, a1, b0, b1):
        return (a0 * b0 % P + P - a1 * b1 % P) % P, (a0 * b1 % P + a1 * b0 % P) % P
    ac0, ac1 = cmul(x[:, 0], x[:, 1], y[:, 0], y[:, 1])
    bd0, bd1 = cmul(x[:, 2], x[:, 3], y[:, 2], y[:, 3])
    ad0, ad1 = cmul(x[:, 0], x[:, 1], y[:, 2], y[:, 3])
    bc0, bc1 = cmul(x[:, 2], x[:, 3], y[:, 0], y[:, 1])
    r0 = (ac0 + 2 * bd0 + P - bd1) % P    # + (2 + i) * bd
    r1 = (ac1 + bd0 + 2 * bd1) % P
    return np.stack([r0, r1, (ad0 + bc0) % P, (ad1 + bc1) % P], axis=1)


def wide_fib_constraints_at_point(mask_values, alpha, oods_point, log_n, powers=None):
    """Components::eval_composition_polynomial_at_point (crates/stwo/src/core/air/components.rs)
    for one WideFibonacciEval component (crates/examples/src/wide_fibonacci/mod.rs:87-98):
    the constraints c_{k+2} - (c_k^2 + c_{k+1}^2) in order, accumulated by Horner's rule in alpha
    (PointEvaluationAccumulator, crates/stwo/src/core/air/accumulation.rs:18-33), i.e. constraint
    j weighted by alpha^(m-1-j) for m constraints, each divided by the vanishing polynomial of the
    trace coset CanonicCoset(log_n) at the point (crates/constraint-framework/src/component.rs:
    220-243).  powers[j] = alpha^j when given (the prover already holds them)."""
    coset_initial = 1 << (31 - (log_n + 1))  # CanonicCoset(n).coset = Coset::odds(n)
    dinv = F.qinv(F.coset_vanishing_q(coset_initial, log_n, oods_point))
    v = np.asarray(mask_values, dtype=np.uint64).reshape(-1, 4)
    m = len(v) - 2
    if m <= 0:
        return [0, 0, 0, 0]
    if powers is None:
        pw = [[1, 0, 0, 0]]
        for _ in range(m - 1):
            pw.append(F.qmul(pw[-1], alpha))
        powers = pw
    w = np.asarray(powers, dtype=np.uint64).reshape(-1, 4)[:m][::-1]   # alpha^(m-1-j)
    a, b, c = v[:m], v[1:m + 1], v[2:]
    cons = (c + 2 * P - _qmul_rows(a, a) - _qmul_rows(b, b)) % P
    terms = _qmul_rows(w, cons)
    total = [int(t) for t in terms.sum(axis=0) % P]
    return F.qmul(dinv, total)


def check_wide_fib_oods(sampled_values, alpha, oods_point, log_n, log_blowup, powers=None):
    """The sanity check closing prove_ex (crates/stwo/src/prover/mod.rs:123-132): raises
    ConstraintsNotSatisfied when the composition value at the OODS point, rebuilt from the
    composition samples, differs from the constraints evaluated on the trace samples.
    sampled_values: [preprocessed, trace (one value per column), composition (8 values)]."""
    max_log_degree_bound = (log_n + log_blowup) - log_blowup  # the lifting size minus the blowup
    lhs = composition_oods_eval([c[0] for c in sampled_values[2]], oods_point, max_log_degree_bound)
    rhs = wide_fib_constraints_at_point([c[0] for c in sampled_values[1]], alpha, oods_point, log_n, powers)
    if [int(v) for v in lhs] != [int(v) for v in rhs]:
        raise ConstraintsNotSatisfied("Constraints not satisfied.")


class _Background:
    """fn(*args) on a daemon thread; result() joins it and re-raises what it raised."""

    def __init__(self, fn, *args):
        self._exc = None

        def run():
            try:
                fn(*args)
            except BaseException as e:  # handed to the caller of result()
                self._exc = e
        self._t = threading.Thread(target=run, daemon=True)
        self._t.start()

    def result(self):
        self._t.join()
        if self._exc is not None:
            raise self._exc


class CommitmentTree:
    """CommitmentTreeProver: coefficient columns, their evaluations on the blown-up domain, and
    the lifted Merkle tree over the evaluations (layers root-first)."""

    def __init__(self, coeffs, evals, layers, root, lifting_log_size):
        self.coeffs, self.evals, self.layers = coeffs, evals, layers
        self.root, self.lifting_log_size = root, lifting_log_size

    def decommit(self, backend, positions):
        """CommitmentTreeProver::decommit (crates/stwo/src/prover/pcs/mod.rs:392-398) ->
        MerkleProverLifted::decommit (crates/stwo/src/prover/vcs_lifted/prover.rs:97-162):
        queried values per column at every position (duplicates kept), and the hash witness."""
        if self.lifting_log_size == 0:
            assert not len(self.evals), "a tree of height 0 holds no columns"
            return np.zeros((0, len(positions)), np.uint32), np.zeros((0, 32), np.uint8)
        return tree_decommit(backend, self.evals, self.layers, positions)  # one round trip


class CommitmentSchemeProver:
    """CommitmentSchemeProver over the HIP backend (crates/stwo/src/prover/pcs/mod.rs:48-315)."""

    def __init__(self, backend, twiddles, fri, trace_lifting_log_size, preprocessed_lifting_log_size):
        self.be, self.tw = backend, twiddles
        self.fri = dict(fri)
        self.trace_lifting_log_size = trace_lifting_log_size
        self.preprocessed_lifting_log_size = preprocessed_lifting_log_size
        self.trees = []

    def lifting_log_size(self, tree_index):
        """PcsConfig::lifting_log_size (crates/stwo/src/core/pcs/mod.rs:70-76)."""
        return self.preprocessed_lifting_log_size if tree_index == 0 else self.trace_lifting_log_size

    def _commit(self, coeffs, evals, channel):
        L = self.lifting_log_size(len(self.trees))
        if not len(evals):
            assert L == 0, "lifting_log_size must be 0 when no columns are committed"
            # the empty tree (root = the hash of nothing) is the same for every proof: committed
            # once per backend, then reused
            empty = getattr(self.be, "_empty_tree", None)
            if empty is None:
                empty = self.be._empty_tree = self.be.commit_lifted([], 0)
            layers, root = empty
        else:
            layers, root = self.be.commit_lifted(evals, L, want_root=not isinstance(channel, DeviceChannel))
        _mix_root(channel, layers, root)
        self.trees.append(CommitmentTree(coeffs, evals, layers, root, L))
        return self.trees[-1]

    def commit_evals(self, evals, channel):
        """TreeBuilder::extend_evals + commit (crates/stwo/src/prover/pcs/mod.rs:325-334, 81-93):
        interpolate, evaluate on the blown-up domain, commit.  One batched sh_commit_trace when
        the tree height is the largest column's blown-up size (the usual case)."""
        b = self.fri["log_blowup_factor"]
        L = self.lifting_log_size(len(self.trees))
        logs = np.asarray(_logs(evals), np.uint32)
        if len(evals) and L == int(logs.max()) + b:
            if isinstance(channel, DeviceChannel) and getattr(self, "graph_commit", False):
                tree = self._commit_captured(evals, logs, b, L, channel)
                if tree is not None:
                    return tree
            # every coefficient and every LDE column of the tree in one pooled allocation each
            coeffs = ColumnSet(self.be.ctx, logs)
            ldes = ColumnSet(self.be.ctx, logs + b)
            layers = HashLayerSet(self.be.ctx, L)
            _, _, _, root = self.be.commit_trace(evals, b, self.tw, coeffs=coeffs, ldes=ldes, layers=layers,
                                                 want_root=not isinstance(channel, DeviceChannel))
            _mix_root(channel, layers, root)
            self.trees.append(CommitmentTree(coeffs, ldes, layers, root, L))
            return self.trees[-1]
        coeffs = []
        for e in evals:
            c = BaseColumn(self.be.ctx, e.log_size)
            c.copy_from(e)
            coeffs.append(c)
        if coeffs:
            self.be.interpolate_columns(coeffs, self.tw)
        return self.commit_polys(coeffs, channel)

    def _commit_captured(self, evals, logs, b, L, channel):
        """The batched commit as a replayed hipGraph (sh_commit_trace_graph_create) when the same
        evaluation buffers are committed again (a prover run over and over on resident traces):
        the graph owns the tree's coefficient, LDE and layer buffers and is kept on the backend
        (one per backend).  The second commit of a set of buffers captures the graph (whose warm-up
        run is that commit); later ones replay it.  Returns None to have the caller commit eagerly:
        on the first commit of a set of buffers, and when the cached graph's buffers still belong
        to a live tree of another proof."""
        key = (tuple(int(p) for p in _ptr_array(evals)), tuple(int(x) for x in logs), b, id(self.tw))
        cached = getattr(self.be, "_commit_graph", None)
        in_use = cached is not None and cached["tree"] is not None and cached["tree"]() is not None
        if cached is not None and cached["key"] == key:
            if in_use:
                return None
            cached["graph"].launch()
        else:
            seen = self.be.__dict__.setdefault("_commit_seen", [])
            if key not in seen or in_use:
                seen.append(key)
                del seen[:-4]
                return None
            if cached is not None:
                cached["graph"].free()
            coeffs = ColumnSet(self.be.ctx, logs)
            ldes = ColumnSet(self.be.ctx, logs + b)
            layers = HashLayerSet(self.be.ctx, L)
            graph = self.be.commit_trace_graph(evals, b, self.tw, coeffs, ldes, layers)
            # the graph reads the evaluations at whatever buffers have these addresses when it is
            # replayed (a replay happens only for a commit of buffers at the same addresses); it
            # keeps the twiddle tree its kernels address, and its own output buffers
            graph.buffers = None
            cached = self.be._commit_graph = {"key": key, "graph": graph, "coeffs": coeffs, "ldes": ldes,
                                              "layers": layers, "tw": self.tw, "tree": None}
        _mix_root(channel, cached["layers"], None)
        tree = CommitmentTree(cached["coeffs"], cached["ldes"], cached["layers"], None, L)
        cached["tree"] = weakref.ref(tree)
        self.trees.append(tree)
        return tree

    def commit_polys(self, coeffs, channel):
        """TreeBuilder::extend_polys + commit: evaluate with the blowup, then commit."""
        b = self.fri["log_blowup_factor"]
        ldes = self.be.evaluate_polynomials(coeffs, b, self.tw) if coeffs else []
        return self._commit(coeffs, ldes, channel)

    def prove_values(self, sampled_points, channel, oods_proof_of_work, check_samples=None):
        """prove_values (crates/stwo/src/prover/pcs/mod.rs:224-315). sampled_points[tree][col] is
        a list of QM31 circle points (8 words). check_samples, when given, is called with the
        sampled values (per tree and column) while the quotient kernels run; an exception from it
        ends the proof.  Returns (felts of the CommitmentSchemeProof as a uint64 array, the proof
        as a dict)."""
        L = self.trees[-1].lifting_log_size
        b = self.fri["log_blowup_factor"]
        ctx = self.be.ctx
        # compute_samples (crates/stwo/src/prover/pcs/mod.rs:176-222): coefficient evaluation at the
        # point doubled down to the column's own domain.  Samples are numbered tree-major, column,
        # point (the order of the channel mix and of the quotient batches); one batched evaluation
        # per distinct (point, doubling count) over every column that samples it.
        counts, flat_pts = [], []
        for tree, pts_tree in zip(self.trees, sampled_points):
            assert len(pts_tree) == len(tree.evals), "one point list per committed column"
            counts.extend(len(p) for p in pts_tree)
            flat_pts.extend(p for pts in pts_tree for p in pts)
        counts = np.asarray(counts, np.int64)
        n_samples = len(flat_pts)
        coeff_ptrs = np.concatenate([_ptr_array(t.coeffs) for t in self.trees])
        coeff_logs = np.concatenate([np.asarray(_logs(t.coeffs), np.uint32) for t in self.trees])
        eval_ptrs = np.concatenate([_ptr_array(t.evals) for t in self.trees])
        eval_logs = np.concatenate([np.asarray(_logs(t.evals), np.uint32) for t in self.trees])
        values = np.zeros((n_samples, 4), np.uint32)
        if n_samples:
            col_of = np.repeat(np.arange(len(counts)), counts)          # column of each sample
            ids = np.fromiter((id(p) for p in flat_pts), np.int64, n_samples)
            _, first, inv = np.unique(ids, return_index=True, return_inverse=True)
            # points equal in value but given as different objects share one evaluation
            by_value, remap = {}, np.empty(len(first), np.int64)
            for k, i in enumerate(first):
                remap[k] = by_value.setdefault(tuple(int(x) for x in flat_pts[i]), len(by_value))
            uniq = list(by_value)
            pt_of = remap[inv]                                            # point of each sample
            dbl = L - eval_logs[col_of].astype(np.int64)                  # doublings of each sample
            keys, job_of = np.unique(pt_of * 64 + dbl, return_inverse=True)
            for j, key in enumerate(keys):
                u, d = divmod(int(key), 64)
                sids = np.nonzero(job_of == j)[0]
                cols = col_of[sids]
                q = list(F.qpoint_repeated_double(list(uniq[u]), d))
                res = self.be.eval_at_points(ColumnRefs(ctx, coeff_ptrs[cols], coeff_logs[cols]), [q], as_array=True)
                values[sids] = res[:, 0, :]
            points = np.asarray(uniq, dtype=np.uint32).reshape(-1, 8)[pt_of]
        else:
            points = np.zeros((0, 8), np.uint32)
        channel.mix_felts(values)
        random_coeff = [int(x) for x in channel.draw_secure_felt()]
        offsets = np.zeros(len(counts) + 1, np.uint32)
        np.cumsum(counts, out=offsets[1:])
        quotients = fri_quotients_arrays(self.be, self.tw, ColumnRefs(ctx, eval_ptrs, eval_logs), offsets, points,
                                         values, random_coeff, L, b)
        # sampled values per tree and column (the proof's nested vectors), and their array blocks
        vl = values.tolist()
        sampled_values, blocks, c = [], [], 0
        for pts_tree in sampled_points:
            n = len(pts_tree)
            o = offsets[c:c + n + 1].tolist()
            sampled_values.append([vl[o[i]:o[i + 1]] for i in range(n)])
            blocks.append((counts[c:c + n].tolist(), values[o[0]:o[-1]]))
            c += n
        checker = None
        if check_samples is not None:
            # on a host thread while the FRI driver runs (the ctypes calls release the GIL);
            # its verdict is collected once the proof is assembled
            checker = _Background(check_samples, sampled_values)
        fri_felts, queries = fri_prove(self.be, self.tw, channel, self.fri_params(), quotients, as_array=True)
        quotients.free()
        return self._decommit_and_serialize(fri_felts, queries, sampled_values, blocks, oods_proof_of_work, checker)

    def prove_values_dpoint_enqueue(self, dch, ws, device_check=None):
        """The device half of prove_values_dpoint, enqueued without waiting: the samples, the
        device check, their mix, the DEEP coefficient, the DEEP quotient and the FRI commit phase
        (with its proof of work).  Returns the pending state for prove_values_dpoint_finish."""
        L = self.trees[-1].lifting_log_size
        b = self.fri["log_blowup_factor"]
        ctx = self.be.ctx
        trees = [t for t in self.trees if len(t.evals)]
        coeff_ptrs = np.concatenate([_ptr_array(t.coeffs) for t in trees])
        coeff_logs = np.concatenate([np.asarray(_logs(t.coeffs), np.uint32) for t in trees])
        eval_ptrs = np.concatenate([_ptr_array(t.evals) for t in trees])
        eval_logs = np.concatenate([np.asarray(_logs(t.evals), np.uint32) for t in trees])
        n_samples = len(eval_ptrs)
        assert n_samples == ws.n_samples
        # roots the host has not seen (mixed on the device) come back with the values
        unknown = [i for i, t in enumerate(self.trees) if t.root is None]
        assert len(unknown) <= ws.n_roots
        for k, i in enumerate(unknown):
            ctx.check(ctx.lib.sh_copy(ctx.h, ws.ptr(ws.ROOTS + 8 * k), self.trees[i].layers[0].ptr, 32), "sh_copy")
        # compute_samples at the point doubled down to each column's domain
        self.be.eval_at_dpoint(ColumnRefs(ctx, coeff_ptrs, coeff_logs), L - eval_logs, ws.ptr(ws.Z),
                               ws.ptr(ws.VALS))
        if device_check is not None:
            device_check()
        dch.mix_words(ws.ptr(ws.VALS), 4 * n_samples)
        dch.draw_secure_felts(1, ws.ptr(ws.RC))
        quotients = fri_quotients_dpoint(self.be, self.tw, ColumnRefs(ctx, eval_ptrs, eval_logs), ws.ptr(ws.Z),
                                         ws.ptr(ws.VALS), ws.ptr(ws.RC), L, b)
        fri_run = fri_commit_dch(self.be, self.tw, dch, self.fri_params(), quotients,
                                 extra=(ws.ptr(0), ws.download_words))
        return {"unknown": unknown, "quotients": quotients, "fri_run": fri_run}

    def prove_values_dpoint_finish(self, channel, ws, pending, oods_proof_of_work=None, check_samples=None,
                                   device_check=False):
        """The host half of prove_values_dpoint: waits for the FRI commit phase's download, then
        the decommitment (one gather for the FRI witnesses and the trees' openings) and the
        felts.  May run again on the same pending state after a graph replays the device half."""
        fri_felts, queries, host, openings = pending["fri_run"].decommit(
            channel, trees=[(t.evals, t.layers, i == 0) for i, t in enumerate(self.trees)])
        ws.read(host)
        for k, i in enumerate(pending["unknown"]):
            self.trees[i].root = host[ws.ROOTS + 8 * k:ws.ROOTS + 8 * k + 8].tobytes()
        values = ws.values
        vl = values.tolist()
        sampled_values, blocks, c = [], [], 0
        for n in (len(t.evals) for t in self.trees):
            sampled_values.append([[vl[c + i]] for i in range(n)])
            blocks.append(([1] * n, values[c:c + n]))
            c += n
        checker = _Background(check_samples, sampled_values, ws) if check_samples is not None else None
        if oods_proof_of_work is None:
            oods_proof_of_work = ws.nonce
        out = self._decommit_and_serialize(fri_felts, queries, sampled_values, blocks, oods_proof_of_work, checker,
                                           openings)
        if device_check and ws.check_failed:
            raise ConstraintsNotSatisfied("Constraints not satisfied.")
        return out

    @staticmethod
    def free_pending(pending):
        pending["fri_run"].free()
        pending["quotients"].free()

    def prove_values_dpoint(self, dch, channel, ws, oods_proof_of_work=None, check_samples=None,
                            device_check=None):
        """prove_values when every committed column is sampled once, at one point that the device
        channel dch drew into ws (a ProofState): the samples are evaluated, mixed and combined
        into the DEEP quotient on the device, and the FRI driver's commit-phase download brings
        back the sampled values with the proof's other device-drawn words; the host channel
        `channel` receives the final state.  check_samples as in prove_values (called with
        (sampled values, the downloaded ProofState)); device_check, when given, is called once the
        sampled values are on the device to enqueue a check that writes ws's CHECK words, and the
        proof raises ConstraintsNotSatisfied when that check's flag is set.  Returns (felts,
        proof dict)."""
        pending = self.prove_values_dpoint_enqueue(dch, ws, device_check)
        try:
            return self.prove_values_dpoint_finish(channel, ws, pending, oods_proof_of_work, check_samples,
                                                   device_check is not None)
        finally:
            self.free_pending(pending)

    def fri_params(self):
        b = self.fri["log_blowup_factor"]
        return FriParams(self.fri["pow_bits"], b, self.fri["log_last_layer_degree_bound"], self.fri["n_queries"],
                         self.fri["fold_step"])

    def _decommit_and_serialize(self, fri_felts, queries, sampled_values, blocks, oods_proof_of_work, checker,
                                openings=None):
        """The query phase of prove_values (crates/stwo/src/prover/pcs/mod.rs:286-315) after the FRI
        proof: every tree decommitted at the queries (unless the FRI driver opened them already:
        openings, one (values, witness) per tree), the check's verdict, the proof's felts."""
        if openings is None:
            L = self.trees[-1].lifting_log_size
            pp_height = self.trees[0].lifting_log_size
            pp_positions = prepare_preprocessed_query_positions(queries, L, pp_height)
            openings = [tree.decommit(self.be, pp_positions if i == 0 else queries)
                        for i, tree in enumerate(self.trees)]
        queried_values = [qv for qv, _ in openings]
        decommitments = [wit for _, wit in openings]
        if checker is not None:
            checker.result()  # re-raises the check's exception (ConstraintsNotSatisfied)
        proof = {"fri_params": dict(self.fri), "commitments": [t.root for t in self.trees],
                 "oods_proof_of_work": oods_proof_of_work, "sampled_values": sampled_values,
                 "sampled_value_blocks": blocks,
                 "decommitments": decommitments, "queried_values": queried_values, "fri_felts": fri_felts,
                 "query_positions": queries}
        return serialize_commitment_scheme_proof(proof, as_array=True), proof


class ProofState:
    """Device words a proof's device channel fills, read back in one download with the FRI
    commit phase: the composition coefficient, the OODS point, the DEEP random coefficient, the
    OODS proof-of-work nonce, the roots mixed on the device, the sampled values; then, not read
    back, the constraint powers."""
    ALPHA, Z, RC, CTL, CHECK, ROOTS = 0, 4, 12, 16, 20, 32

    def __init__(self, ctx, n_samples, n_powers, n_roots=2):
        self.n_samples, self.n_roots = n_samples, n_roots
        self.VALS = self.ROOTS + 8 * n_roots
        self.POW = self.VALS + 4 * n_samples
        self.download_words = self.POW
        self.buf = DeviceBuffer(ctx, 4 * (self.POW + 4 * max(1, n_powers)))
        self.host = None

    def ptr(self, word):
        return self.buf.ptr + 4 * word

    def read(self, host):
        self.host = host

    @property
    def alpha(self):
        return [int(x) for x in self.host[self.ALPHA:self.ALPHA + 4]]

    @property
    def oods_point(self):
        return [int(x) for x in self.host[self.Z:self.Z + 8]]

    @property
    def nonce(self):
        return int(self.host[self.CTL]) | (int(self.host[self.CTL + 1]) << 32)

    @property
    def check_failed(self):
        return bool(self.host[self.CHECK])

    @property
    def values(self):
        return self.host[self.VALS:self.VALS + 4 * self.n_samples].reshape(-1, 4)

    def free(self):
        self.buf.free()


def prepare_preprocessed_query_positions(query_positions, max_log_size, pp_max_log_size):
    """crates/stwo/src/core/pcs/utils.rs:185-203."""
    if pp_max_log_size == 0:
        return []
    if max_log_size < pp_max_log_size:
        return [((p >> 1) << (pp_max_log_size - max_log_size + 1)) + (p & 1) for p in query_positions]
    return [((p >> (max_log_size - pp_max_log_size + 1)) << 1) + (p & 1) for p in query_positions]


def serialize_commitment_scheme_proof(proof, as_array=False):
    """CairoSerialize for CommitmentSchemeProof (crates/cairo-serialize/src/serialize.rs:121-143):
    PcsConfig (its FriParams only), commitments, oods_proof_of_work, sampled_values,
    decommitments, queried_values, fri_proof; vectors are length-prefixed, a hash is 8 LE u32,
    a QM31 4 felts.  Returns a list of ints, or a uint64 array when as_array."""
    parts = []
    head = []  # small scalars gathered into one array per run

    def flush():
        if head:
            parts.append(np.asarray(head, dtype=np.uint64))
            head.clear()

    def hashes(hs):
        head.append(len(hs))
        if len(hs):
            flush()
            raw = np.ascontiguousarray(hs, np.uint8).reshape(-1) if isinstance(hs, np.ndarray) else \
                np.frombuffer(b"".join(hs), np.uint8)
            parts.append(raw.view("<u4").astype(np.uint64))

    f = proof["fri_params"]
    head.extend([f["pow_bits"], f["log_blowup_factor"], f["log_last_layer_degree_bound"], f["n_queries"],
                 f["fold_step"]])
    hashes(proof["commitments"])
    head.extend([proof["oods_proof_of_work"], len(proof["sampled_values"])])
    blocks = proof.get("sampled_value_blocks")
    for ti, tree in enumerate(proof["sampled_values"]):
        head.append(len(tree))
        cnt = blocks[ti][0] if blocks else None
        if cnt and len(set(cnt)) == 1 and cnt[0]:
            # every column has the same number of samples: one block [count, 4 * count felts] per column
            flush()
            k = cnt[0]
            block = np.empty((len(cnt), 1 + 4 * k), np.uint64)
            block[:, 0] = k
            block[:, 1:] = np.asarray(blocks[ti][1], np.uint64).reshape(len(cnt), 4 * k)
            parts.append(block.ravel())
            continue
        for col in tree:
            head.append(len(col))
            for v in col:
                head.extend(int(x) for x in v)
    head.append(len(proof["decommitments"]))
    for wit in proof["decommitments"]:
        hashes(wit)
    head.append(len(proof["queried_values"]))
    for tree in proof["queried_values"]:
        tree = np.asarray(tree, dtype=np.uint64).reshape(len(tree), -1) if len(tree) else np.zeros((0, 0), np.uint64)
        head.append(tree.shape[0])
        if tree.shape[0]:
            flush()
            block = np.empty((tree.shape[0], tree.shape[1] + 1), np.uint64)
            block[:, 0] = tree.shape[1]
            block[:, 1:] = tree
            parts.append(block.ravel())
    flush()
    parts.append(np.asarray(proof["fri_felts"], dtype=np.uint64))
    out = np.concatenate(parts)
    return out if as_array else out.tolist()


def start_wide_fibonacci(backend, twiddles, trace, log_n, fri, channel=None, committed_trace=None,
                         device_channel=None, graph_commit=None):
    """First phase of test_wide_fib_prove_with_blake's prover (crates/examples/src/wide_fibonacci/
    mod.rs:183-210): config with an empty, unlifted preprocessed tree; commit it and the trace.
    committed_trace may carry a CommitmentTree built elsewhere (batched across proofs); its root is
    mixed exactly as a fresh commit would be.  With the device channel (DEVICE_CHANNEL unless
    device_channel says otherwise) a fresh trace's root is mixed on the device, and the host
    channel is behind until the FRI commit phase returns; a trace committed again from the same
    buffers then replays a captured graph of its commit (GRAPH_COMMITS unless graph_commit says
    otherwise).  Returns (scheme, channel)."""
    b = fri["log_blowup_factor"]
    channel = channel or Channel(backend.lib)
    scheme = CommitmentSchemeProver(backend, twiddles, fri, log_n + b, 0)
    scheme.device_channel = DEVICE_CHANNEL if device_channel is None else bool(device_channel)
    scheme.graph_commit = GRAPH_COMMITS if graph_commit is None else bool(graph_commit)
    scheme.dch = None
    scheme.commit_evals([], channel)                       # preprocessed trace (empty; root cached)
    if committed_trace is None:
        if scheme.device_channel:
            scheme.dch = DeviceChannel(backend, channel)
            scheme.commit_evals(trace, scheme.dch)         # trace, root mixed on the device
        else:
            scheme.commit_evals(trace, channel)            # trace
    else:
        assert committed_trace.lifting_log_size == log_n + b
        channel.mix_root(committed_trace.root)
        scheme.trees.append(committed_trace)
    return scheme, channel


def finish_wide_fibonacci(scheme, channel, n_cols, log_n):
    """prove_ex (crates/stwo/src/prover/mod.rs:38-135) after the trace commitment: the composition
    polynomial (evaluated on the committed evaluations, interpolated, split at the middle,
    committed as 8 coordinate polynomials), the OODS proof of work and point, prove_values.
    Runs on scheme.be (which may be another context than the one that committed the trace)."""
    if getattr(scheme, "device_channel", False):
        return _finish_wide_fibonacci_device(scheme, channel, n_cols, log_n)
    backend, twiddles, fri = scheme.be, scheme.tw, scheme.fri
    ctx = backend.ctx
    b = fri["log_blowup_factor"]
    trace_tree = scheme.trees[1]
    # compute_composition_polynomial (crates/stwo/src/prover/air/component_prover.rs:103-131)
    alpha = [int(x) for x in channel.draw_secure_felt()]
    m = n_cols - 2
    powers = backend.generate_secure_powers(alpha, m)[::-1]  # ColumnAccumulator powers, reversed
    eval_log = log_n + 1                                   # max_constraint_log_degree_bound
    log_expansion = b - 1                                  # EvaluationMode::infer: SubDomain
    assert log_expansion == 0, "only log_blowup_factor 1 (SubDomain {log_expansion: 0}) is driven here"
    acc = SecureColumnByCoords(ctx, eval_log)
    for c in acc.columns:
        ctx.check(ctx.lib.sh_memset(ctx.h, c.ptr, 0, c.nbytes), "memset")
    backend.wide_fibonacci_composition(trace_tree.evals, log_n, eval_log, powers, acc, log_expansion)
    backend.interpolate_columns(acc.columns, twiddles)    # DomainEvaluationAccumulator::finalize
    halves = [backend.split_at_mid(c) for c in acc.columns]
    comp_polys = [h[0] for h in halves] + [h[1] for h in halves]
    scheme.commit_polys(comp_polys, channel)
    # OODS proof of work and point (crates/stwo/src/prover/mod.rs:74-81)
    oods_pow = backend.grind(channel.digest, OODS_POW_BITS)
    channel.mix_u64(oods_pow)
    z = [int(x) for x in channel.draw_random_point()]
    sampled_points = [[], [[z] for _ in range(n_cols)], [[z] for _ in range(8)]]
    # the sanity check closing prove_ex (crates/stwo/src/prover/mod.rs:123-132), run on the host
    # while the quotient kernels run (it reads only the sampled values)
    felts, proof = scheme.prove_values(
        sampled_points, channel, oods_pow,
        check_samples=lambda sv: check_wide_fib_oods(sv, alpha, z, log_n, b, powers[::-1]))
    proof["composition_random_coeff"] = alpha
    proof["oods_point"] = z
    return felts, proof


class _ProofRun:
    """A device-channel proof after its device work was enqueued (_finish_device_enqueue): the
    scheme with its trees, the device channel, the ProofState words and the pending DEEP/FRI
    state.  finish() waits for the device and completes the proof on the host."""

    def __init__(self, scheme, dch, own_dch, ws, pending, n_cols, log_n):
        self.scheme, self.dch, self.own_dch, self.ws, self.pending = scheme, dch, own_dch, ws, pending
        self.n_cols, self.log_n = n_cols, log_n

    def finish(self, channel):
        scheme, ws = self.scheme, self.ws
        felts, proof = scheme.prove_values_dpoint_finish(channel, ws, self.pending, device_check=True)
        proof["composition_random_coeff"] = ws.alpha
        proof["oods_point"] = ws.oods_point
        return felts, proof

    def free(self, trees=False):
        """Releases the proof's device buffers (with trees: the committed trees' too, except the
        shared empty preprocessed tree)."""
        CommitmentSchemeProver.free_pending(self.pending)
        self.ws.free()
        if self.own_dch:
            self.dch.free()
        if trees:
            empty = getattr(self.scheme.be, "_empty_tree", (None,))[0]
            for t in self.scheme.trees:
                if t.layers is empty:
                    continue
                for part in (t.coeffs, t.evals, t.layers):
                    if hasattr(part, "free"):
                        part.free()
                    else:
                        for x in part:
                            x.free()


def _finish_device_enqueue(scheme, channel, n_cols, log_n, own_dch=None):
    """The device work of finish_wide_fibonacci with the channel in device memory, enqueued only:
    the composition coefficient and its powers, the composition polynomial and its commitment,
    the OODS proof of work and point, the samples and the closing check, the DEEP quotient and
    the FRI commit phase.  Returns a _ProofRun."""
    backend, twiddles, fri = scheme.be, scheme.tw, scheme.fri
    ctx = backend.ctx
    b = fri["log_blowup_factor"]
    trace_tree = scheme.trees[1]
    # the trace committed elsewhere (batched) was mixed on the host: the device channel starts here
    if scheme.dch is not None and scheme.dch.be is backend:
        dch, own = scheme.dch, False
    else:
        dch, own = DeviceChannel(backend, channel), True
    if own_dch is not None:
        own = own_dch
    m = n_cols - 2
    ws = ProofState(ctx, n_cols + 8, m)
    # compute_composition_polynomial (crates/stwo/src/prover/air/component_prover.rs:103-131)
    dch.draw_secure_felts(1, ws.ptr(ws.ALPHA))
    backend.secure_powers_dev(ws.ptr(ws.ALPHA), m, True, ws.ptr(ws.POW))  # ColumnAccumulator order
    eval_log = log_n + 1
    log_expansion = b - 1
    assert log_expansion == 0, "only log_blowup_factor 1 (SubDomain {log_expansion: 0}) is driven here"
    acc = SecureColumnByCoords(ctx, eval_log)
    for c in acc.columns:
        ctx.check(ctx.lib.sh_memset(ctx.h, c.ptr, 0, c.nbytes), "memset")
    backend.wide_fibonacci_composition_dev(trace_tree.evals, log_n, eval_log, ws.ptr(ws.POW), acc, log_expansion)
    backend.interpolate_columns(acc.columns, twiddles)
    halves = [backend.split_at_mid(c) for c in acc.columns]   # each acc column becomes a left half
    scheme.commit_polys([h[0] for h in halves] + [h[1] for h in halves], dch)
    # OODS proof of work and point (crates/stwo/src/prover/mod.rs:74-81)
    dch.grind(OODS_POW_BITS, ws.ptr(ws.CTL))
    dch.draw_random_point(ws.ptr(ws.Z))
    # the closing sanity check runs on the device, on the sampled values where they are
    pending = scheme.prove_values_dpoint_enqueue(dch, ws, device_check=lambda: backend.wide_fibonacci_oods_check(
        ws.ptr(ws.VALS), n_cols, ws.ptr(ws.POW), ws.ptr(ws.Z), log_n, ws.ptr(ws.CHECK)))
    return _ProofRun(scheme, dch, own, ws, pending, n_cols, log_n)


def _finish_wide_fibonacci_device(scheme, channel, n_cols, log_n):
    """finish_wide_fibonacci with the channel in device memory: the composition coefficient and
    its powers, the composition root, the OODS proof of work and point, the sampled values and the
    DEEP coefficient are drawn, mixed and used on the device; the host reads them back with the
    FRI commit phase.  Same proof, byte for byte."""
    run = _finish_device_enqueue(scheme, channel, n_cols, log_n)
    try:
        return run.finish(channel)
    finally:
        run.free()


class _ProofGraph:
    """A whole device-channel proof of one trace (fixed buffers and shapes) as one hipGraph: the
    batched trace commit, the composition and its commit, the OODS proof of work and samples, the
    DEEP quotient and the FRI commit phase with its proof of work.  The buffers the graph writes
    (trees, quotient, FRI layers, ProofState words) are the graph's own; the trace's evaluation
    buffers are read wherever they are, with their contents at launch time.  A launch is preceded
    by loading the host channel's state into the graph's device channel and followed by the host
    half of the proof (_ProofRun.finish)."""

    def __init__(self, backend, twiddles, trace, log_n, fri, key):
        self.be, self.tw, self.trace, self.log_n, self.fri, self.key = backend, twiddles, trace, log_n, fri, key
        self.graph = Graph(backend.ctx)
        self.run = None

    def _enqueue(self, channel, load):
        b = self.fri["log_blowup_factor"]
        scheme = CommitmentSchemeProver(self.be, self.tw, self.fri, self.log_n + b, 0)
        scheme.device_channel, scheme.graph_commit = True, False
        scheme.commit_evals([], channel)                   # the cached empty preprocessed tree (host mix)
        scheme.dch = DeviceChannel(self.be, channel, load=load)
        scheme.commit_evals(self.trace, scheme.dch)
        return _finish_device_enqueue(scheme, channel, len(self.trace), self.log_n, own_dch=True)

    def build(self, channel):
        """Proves once eagerly on the graph's resources (the warm-up; its proof is returned), then
        captures the same device work."""
        if getattr(self.be, "_empty_tree", None) is None:   # shared by every proof: not the graph's
            self.be._empty_tree = self.be.commit_lifted([], 0)
        with self.graph.warmup():
            run = self._enqueue(channel, load=True)
        try:
            out = run.finish(channel)
        finally:
            run.free(trees=True)   # the warm-up's blocks, free again for the capture
        with self.graph.capture():
            self.run = self._enqueue(Channel(self.be.lib), load=False)
        return out

    def start(self, channel):
        """The device half of one proof: the host mixes the empty preprocessed root, loads the
        channel into the device and launches the graph (no wait)."""
        layers, root = self.be._empty_tree
        _mix_root(channel, layers, root)
        self.run.dch.load(channel)
        self.graph.launch()

    def finish(self, channel):
        """The host half of the proof that start() launched: waits for the replay, then the
        decommitment and the felts.  Returns (felts, proof dict)."""
        # the host waits for the replay before enqueueing the decommitment's download: a
        # device-to-host copy enqueued behind a replay still in flight stalls the replay partway
        # (measured: ~1.2 ms of idle device, 10.2 instead of 8.8 ms per proof)
        self.be.ctx.sync()
        return self.run.finish(channel)

    def prove(self, channel):
        """One proof by a graph launch (start, then finish)."""
        self.start(channel)
        return self.finish(channel)

    def free(self):
        if self.run is not None:
            self.run.free(trees=True)
            self.run = None
        self.graph.free()


def _prove_with_graph(backend, twiddles, trace, log_n, fri, channel):
    """prove_wide_fibonacci through a captured _ProofGraph when this trace's buffers were proved
    before (the first proof of a set of buffers returns None: the caller proves eagerly; the
    second builds the graph, one per backend; later ones replay it)."""
    key = (tuple(int(p) for p in _ptr_array(trace)), tuple(int(x) for x in _logs(trace)), log_n,
           tuple(sorted(fri.items())), id(twiddles))
    cached = getattr(backend, "_proof_graph", None)
    if cached is not None and cached.key == key:
        return cached.prove(channel)
    seen = backend.__dict__.setdefault("_proof_seen", [])
    if key not in seen:
        seen.append(key)
        del seen[:-4]
        return None
    if cached is not None:
        backend._proof_graph = None
        cached.free()
    pg = _ProofGraph(backend, twiddles, trace, log_n, fri, key)
    try:
        out = pg.build(channel)
    except BaseException:
        pg.free()
        raise
    backend._proof_graph = pg
    return out


def prove_wide_fibonacci(backend, twiddles, trace, log_n, fri, channel=None, device_channel=None,
                         graph_commit=None):
    """test_wide_fib_prove_with_blake's prover side (crates/examples/src/wide_fibonacci/mod.rs:
    183-220) through prove_ex: empty preprocessed tree, the trace tree, then
    finish_wide_fibonacci.  trace: device columns (BaseColumn, bit-reversed evaluations on
    CanonicCoset(log_n)).  device_channel: keep the channel in device memory (default
    DEVICE_CHANNEL); graph_commit: with the device channel and a proof of work the device search
    covers (pow_bits <= 32), a trace proved again from the same buffers is proved by replaying a
    captured graph of the proof's whole device work (_ProofGraph; default GRAPH_COMMITS).
    Returns (felts, proof dict, the channel in its final state)."""
    assert fri["log_blowup_factor"] >= 1, "the wide-Fibonacci constraints need log_blowup_factor >= 1"
    dc = DEVICE_CHANNEL if device_channel is None else bool(device_channel)
    gc = GRAPH_COMMITS if graph_commit is None else bool(graph_commit)
    if dc and gc and fri["pow_bits"] <= 32:
        channel = channel or Channel(backend.lib)
        out = _prove_with_graph(backend, twiddles, trace, log_n, fri, channel)
        if out is not None:
            return out[0], out[1], channel
    scheme, channel = start_wide_fibonacci(backend, twiddles, trace, log_n, fri, channel,
                                           device_channel=dc, graph_commit=False)
    felts, proof = finish_wide_fibonacci(scheme, channel, len(trace), log_n)
    return felts, proof, channel


class ProofPipeline:
    """Many proofs of one wide-Fibonacci trace (fixed buffers), several in flight: proof i runs on
    backends[i % len(backends)], each backend (a context of its own, with its own stream) replaying
    its own captured _ProofGraph, and the host finishes proof i (decommitment, felts) while the
    device work of the proofs after it runs.  The contexts share the device and their streams run
    concurrently, so the narrow stretches of one proof (Merkle tops, FRI folds, channel hashing,
    the proofs of work's last waves) overlap the wide passes of the next.  Each proof is byte for
    byte the one prove_wide_fibonacci makes with the same channel (the channel objects receive
    their final states).  Building runs one warm-up proof per backend."""

    def __init__(self, backends, trace, log_n, fri):
        assert fri["pow_bits"] <= 32, "the proof graph covers device proofs of work of up to 32 bits"
        assert fri["log_blowup_factor"] >= 1, "the wide-Fibonacci constraints need log_blowup_factor >= 1"
        self.graphs = []
        try:
            for be in backends:
                if getattr(be, "_empty_tree", None) is None:
                    be._empty_tree = be.commit_lifted([], 0)
                tw = be.precompute_canonic_twiddles(log_n + fri["log_blowup_factor"])
                pg = _ProofGraph(be, tw, trace, log_n, fri, None)
                self.graphs.append(pg)
                pg.build(Channel(be.lib))
        except BaseException:
            self.free()
            raise

    def prove(self, channels):
        """Proves once per channel (a fresh Channel for each None), in order, with up to
        len(backends) proofs in flight.  Returns [(felts, proof dict)], one per channel."""
        channels = [c if c is not None else Channel(self.graphs[0].be.lib) for c in channels]
        k = len(self.graphs)
        out = []
        try:
            for i, ch in enumerate(channels):
                if i >= k:   # the slot's previous proof is finished before its graph runs again
                    out.append(self.graphs[(i - k) % k].finish(channels[i - k]))
                self.graphs[i % k].start(ch)
            for i in range(max(0, len(channels) - k), len(channels)):
                out.append(self.graphs[i % k].finish(channels[i]))
        except BaseException:
            # a proof that failed (ConstraintsNotSatisfied, a device error) leaves the others of
            # this call unfinished: wait for them so that the pipeline is idle again
            for pg in self.graphs:
                try:
                    pg.be.ctx.sync()
                except Exception:
                    pass
            raise
        return out

    def free(self):
        for pg in self.graphs:
            pg.free()
        self.graphs = []


def commit_traces_batched(backend, twiddles, traces, log_sizes, log_blowup):
    """Trace commitments of several proofs at once: one interpolation and one evaluation call
    over every column of every proof (one launch chain per column size, whichever proof a column
    belongs to), then each proof's own lifted Merkle tree.  The coefficient columns are the
    interpolated trace buffers themselves.  Returns one CommitmentTree per proof, identical to
    what commit_evals would build for it alone."""
    cols = [c for t in traces for c in t]
    backend.interpolate_columns(cols, twiddles)
    ldes = backend.evaluate_polynomials(cols, log_blowup, twiddles)
    trees, k = [], 0
    for t, lg in zip(traces, log_sizes):
        mine = ldes[k:k + len(t)]
        k += len(t)
        layers, root = backend.commit_lifted(mine, lg + log_blowup)
        trees.append(CommitmentTree(list(t), mine, layers, root, lg + log_blowup))
    return trees
