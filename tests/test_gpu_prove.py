"""End-to-end proofs on the device: the wide-Fibonacci AIR proved through the commitment
scheme (trace commit, composition, OODS sampling, DEEP quotients, FRI, decommitment) with the
proof written as CairoSerde felts.  Small cases are byte-identical to the oracle's prover; every
proof is accepted by the oracle's restatement of the reference verifier, and a tampered proof
is rejected."""
import numpy as np
import pytest

import oracle as O
from oracle import protocol as PR
from proving_amd.backend import BaseColumn
from proving_amd.prover import ConstraintsNotSatisfied, finish_wide_fibonacci, prove_wide_fibonacci, start_wide_fibonacci
from proving_amd.workloads import WIDE_FIB, wide_fib_trace

pytestmark = pytest.mark.gpu
P = O.P


def params(pow_bits=4, last=0, n_queries=3, fold_step=1):
    return {"pow_bits": pow_bits, "log_blowup_factor": 1, "log_last_layer_degree_bound": last,
            "n_queries": n_queries, "fold_step": fold_step}


@pytest.mark.parametrize("n_cols,log_n,fri", [
    (4, 4, params()), (12, 6, params(6, 1, 9)), (32, 10, params(8, 2, 20)), (64, 13, params(10, 0, 70)),
    # the fold step of the reference's Cairo proofs (packed FRI leaves)
    (16, 9, params(6, 0, 30, 3)), (8, 7, params(4, 1, 12, 2))])
@pytest.mark.parametrize("device_channel", [True, False], ids=["device_channel", "host_channel"])
def test_wide_fib_proof_matches_oracle_and_verifies(ctx, backend, n_cols, log_n, fri, device_channel):
    """The proof with the channel in device memory and with the host channel: both byte-identical
    to the oracle's."""
    trace = wide_fib_trace(n_cols, log_n, 100 + n_cols)
    tw, itw = O.canonic_twiddles(log_n + 1)
    exp_felts, _ = PR.wide_fib_prove(trace, log_n, fri, tw, itw)
    tree = backend.precompute_canonic_twiddles(log_n + 1)
    cols = [BaseColumn.from_cpu(ctx, c) for c in trace]
    felts, proof, _ = prove_wide_fibonacci(backend, tree, cols, log_n, fri, device_channel=device_channel)
    assert felts.tolist() == exp_felts
    assert PR.wide_fib_verify(felts.tolist(), n_cols, log_n, fri)


def test_device_channel_final_state_and_proof_fields(ctx, backend):
    """The device-channel proof leaves the caller's channel where the host-channel proof does, and
    reports the same composition coefficient, OODS point, nonce and sampled values."""
    n_cols, log_n, fri = 12, 8, params(6, 1, 9)
    trace = wide_fib_trace(n_cols, log_n, 77)
    tree = backend.precompute_canonic_twiddles(log_n + 1)
    out = {}
    for dc in (True, False):
        cols = [BaseColumn.from_cpu(ctx, c) for c in trace]
        felts, proof, ch = prove_wide_fibonacci(backend, tree, cols, log_n, fri, device_channel=dc)
        out[dc] = (felts.tolist(), ch.digest, ch.s.n_draws, proof)
    (fd, dd, nd, pd), (fh, dh, nh, ph) = out[True], out[False]
    assert fd == fh and (dd, nd) == (dh, nh)
    for k in ("composition_random_coeff", "oods_point", "oods_proof_of_work", "sampled_values", "commitments",
              "query_positions"):
        assert pd[k] == ph[k], k


def test_wide_fib_bad_trace_rejected(ctx, backend):
    """A trace that breaks one transition constraint: the prover itself refuses it
    (ProvingError::ConstraintsNotSatisfied, crates/stwo/src/prover/mod.rs:123-132)."""
    n_cols, log_n = 16, 8
    trace = wide_fib_trace(n_cols, log_n, 5)
    trace[7][100] = (int(trace[7][100]) + 1) % P
    fri = params()
    tree = backend.precompute_canonic_twiddles(log_n + 1)
    for device_channel in (True, False):
        with pytest.raises(ConstraintsNotSatisfied):
            prove_wide_fibonacci(backend, tree, [BaseColumn.from_cpu(ctx, c) for c in trace], log_n, fri,
                                 device_channel=device_channel)


def test_wide_fib_full_size_verifies(ctx, backend):
    """256 columns x 2^20 with FriParams::new(26, 0, 1, 70, 1) (the production defaults): the
    proof from the device verifies with the oracle verifier."""
    w = WIDE_FIB
    trace = wide_fib_trace(w["n_cols"], w["log_size"], w["seed"])
    tree = backend.precompute_canonic_twiddles(w["log_size"] + 1)
    cols = [BaseColumn.from_cpu(ctx, c) for c in trace]
    del trace
    felts, proof, _ = prove_wide_fibonacci(backend, tree, cols, w["log_size"], w["fri"])
    assert PR.wide_fib_verify(felts.tolist(), w["n_cols"], w["log_size"], w["fri"])


def test_repeated_proofs_replay_the_proof_graph(ctx, backend):
    """Proving a trace again from the same buffers: the first proof runs eagerly, the second is
    the warm-up of a captured graph of the proof's whole device work (then captured), later ones
    replay it; every proof is byte-equal to the oracle's and leaves the channel where the eager
    proof does, also after the trace buffers are overwritten with another trace and with other
    proofs (other buffers, the host channel) between the replays."""
    n_cols, log_n, fri = 12, 9, params(6, 1, 9)
    tw, itw = O.canonic_twiddles(log_n + 1)
    tree = backend.precompute_canonic_twiddles(log_n + 1)
    traces = [wide_fib_trace(n_cols, log_n, s) for s in (31, 32)]
    exp = [PR.wide_fib_prove(t, log_n, fri, tw, itw)[0] for t in traces]
    cols = [BaseColumn.from_cpu(ctx, c) for c in traces[0]]
    digests = set()
    for k in range(4):
        felts, _, ch = prove_wide_fibonacci(backend, tree, cols, log_n, fri, graph_commit=True)
        assert felts.tolist() == exp[0], f"proof {k}"
        digests.add((ch.digest, ch.s.n_draws))
    assert len(digests) == 1
    assert backend._proof_graph is not None
    other = [BaseColumn.from_cpu(ctx, c) for c in traces[1]]
    felts, _, _ = prove_wide_fibonacci(backend, tree, other, log_n, fri, graph_commit=True)
    assert felts.tolist() == exp[1]
    felts, _, _ = prove_wide_fibonacci(backend, tree, other, log_n, fri, device_channel=False)
    assert felts.tolist() == exp[1]
    felts, _, _ = prove_wide_fibonacci(backend, tree, cols, log_n, fri, graph_commit=True)
    assert felts.tolist() == exp[0]
    # new contents in the same buffers: the replayed graph reads them
    for c, h in zip(cols, traces[1]):
        c.upload(h)
    felts, _, _ = prove_wide_fibonacci(backend, tree, cols, log_n, fri, graph_commit=True)
    assert felts.tolist() == exp[1]
    # a trace that breaks a constraint, proved by the replay: the device check still fires
    bad = [np.array(h) for h in traces[1]]
    bad[5][17] = (int(bad[5][17]) + 1) % P
    for c, h in zip(cols, bad):
        c.upload(h)
    with pytest.raises(ConstraintsNotSatisfied):
        prove_wide_fibonacci(backend, tree, cols, log_n, fri, graph_commit=True)
    backend._proof_graph.free()
    backend._proof_graph = None


def test_start_wide_fibonacci_replays_the_captured_commit(ctx, backend):
    """The trace commit of start_wide_fibonacci proved again from the same buffers: the first
    commits eagerly, the second captures the batched commit as a graph, later ones replay it, and
    a proof started while another's tree from that graph is alive commits eagerly; every proof is
    byte-equal to the oracle's."""
    n_cols, log_n, fri = 12, 9, params(6, 1, 9)
    tw, itw = O.canonic_twiddles(log_n + 1)
    tree = backend.precompute_canonic_twiddles(log_n + 1)
    trace = wide_fib_trace(n_cols, log_n, 41)
    exp = PR.wide_fib_prove(trace, log_n, fri, tw, itw)[0]
    cols = [BaseColumn.from_cpu(ctx, c) for c in trace]
    for k in range(4):
        scheme, ch = start_wide_fibonacci(backend, tree, cols, log_n, fri, graph_commit=True)
        felts, _ = finish_wide_fibonacci(scheme, ch, n_cols, log_n)
        assert felts.tolist() == exp, f"proof {k}"
        del scheme
    assert backend._commit_graph is not None
    # a proof in progress holds the graph's tree: another proof of the same buffers commits eagerly
    scheme1, ch1 = start_wide_fibonacci(backend, tree, cols, log_n, fri, graph_commit=True)
    scheme2, ch2 = start_wide_fibonacci(backend, tree, cols, log_n, fri, graph_commit=True)
    felts2, _ = finish_wide_fibonacci(scheme2, ch2, n_cols, log_n)
    felts1, _ = finish_wide_fibonacci(scheme1, ch1, n_cols, log_n)
    assert felts1.tolist() == exp and felts2.tolist() == exp


@pytest.mark.parametrize("n_ctx", [2, 3])
def test_proof_pipeline_contexts(ctx, backend, n_ctx):
    """ProofPipeline: proofs of one trace with n_ctx in flight, on n_ctx contexts each replaying its
    own proof graph.  With fresh channels every proof is the oracle's; with channels that start
    from other states each proof and final channel state equal the host-channel prover's for the
    same starting state."""
    from proving_amd.backend import Channel, Context, HipBackend
    from proving_amd.prover import ProofPipeline
    n_cols, log_n, fri = 12, 9, params(6, 1, 9)
    tw, itw = O.canonic_twiddles(log_n + 1)
    trace = wide_fib_trace(n_cols, log_n, 33)
    exp = PR.wide_fib_prove(trace, log_n, fri, tw, itw)[0]
    cols = [BaseColumn.from_cpu(ctx, c) for c in trace]
    tree = backend.precompute_canonic_twiddles(log_n + 1)
    extra = [Context(0) for _ in range(n_ctx - 1)]
    try:
        pipe = ProofPipeline([backend] + [HipBackend(c) for c in extra], cols, log_n, fri)
        try:
            for felts, _ in pipe.prove([None] * 5):
                assert felts.tolist() == exp
            chans = []
            for i in range(5):
                c = Channel(backend.lib)
                c.mix_u64(1000 + i)
                chans.append(c)
            got = pipe.prove(chans)
            for i, ((felts, _), c) in enumerate(zip(got, chans)):
                h = Channel(backend.lib)
                h.mix_u64(1000 + i)
                want, _, h = prove_wide_fibonacci(backend, tree, cols, log_n, fri, channel=h, device_channel=False)
                assert np.array_equal(felts, want), f"proof {i}"
                assert bytes(c.s.digest) == bytes(h.s.digest) and c.s.n_draws == h.s.n_draws, f"channel {i}"
            assert len({bytes(c.s.digest) for c in chans}) == 5
        finally:
            pipe.free()
    finally:
        for c in extra:
            c.close()


def test_proof_pipeline_bad_trace_then_recovers(ctx, backend):
    """A trace that breaks a constraint, proved through ProofPipeline: the proof raises
    ConstraintsNotSatisfied, the pipeline is left idle, and once the buffers hold a good trace
    again its proofs are the oracle's."""
    from proving_amd.backend import Context, HipBackend
    from proving_amd.prover import ProofPipeline
    n_cols, log_n, fri = 12, 9, params(6, 1, 9)
    tw, itw = O.canonic_twiddles(log_n + 1)
    good = wide_fib_trace(n_cols, log_n, 35)
    exp = PR.wide_fib_prove(good, log_n, fri, tw, itw)[0]
    bad = [np.array(c) for c in good]
    bad[4][33] = (int(bad[4][33]) + 1) % P
    cols = [BaseColumn.from_cpu(ctx, c) for c in good]
    ctx2 = Context(0)
    try:
        pipe = ProofPipeline([backend, HipBackend(ctx2)], cols, log_n, fri)
        try:
            for c, h in zip(cols, bad):
                c.upload(h)
            with pytest.raises(ConstraintsNotSatisfied):
                pipe.prove([None] * 3)
            for c, h in zip(cols, good):
                c.upload(h)
            for felts, _ in pipe.prove([None] * 3):
                assert felts.tolist() == exp
        finally:
            pipe.free()
    finally:
        ctx2.close()
