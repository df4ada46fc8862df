"""The proof-serving loop: a mix of requests proved with cross-request trace commits and
overlapping worker contexts gives, for every request, exactly the proof made alone; the oracle
verifier accepts them."""
import numpy as np
import pytest

import oracle as O
from oracle import protocol as PR
from proving_amd.backend import Context, HipBackend
from proving_amd.serving import ProofServer, device_bytes, make_requests, prove_sequential

pytestmark = pytest.mark.gpu

FRI = {"pow_bits": 6, "log_blowup_factor": 1, "log_last_layer_degree_bound": 1, "n_queries": 8, "fold_step": 1}


@pytest.mark.parametrize("whole_proofs", [False, True], ids=["batched_commits", "whole_proofs"])
def test_each_proof_is_the_same_alone_and_in_the_mix(whole_proofs):
    reqs = make_requests(7, 40, col_choices=(4, 16, 32), log_range=(6, 11))
    # a budget that fits about two requests forces admissions to wait for completions
    budget = 2 * max(device_bytes(r.n_cols, r.log_n) for r in reqs)
    server = ProofServer(12, FRI, n_workers=2, hbm_budget=budget, whole_proofs=whole_proofs)
    try:
        mixed = server.serve(reqs, realtime=False)
    finally:
        server.close()
    alone = make_requests(7, 40, col_choices=(4, 16, 32), log_range=(6, 11))
    ctx = Context(0)
    be = HipBackend(ctx)
    tw = be.precompute_canonic_twiddles(12)
    prove_sequential(be, tw, alone, FRI, realtime=False)
    for m, a in zip(mixed, alone):
        assert np.array_equal(m.felts, a.felts), f"request {m.index} differs"
        assert PR.wide_fib_verify(m.felts.tolist(), m.n_cols, m.log_n, FRI)
    ctx.close()


@pytest.mark.parametrize("cap_shapes", [64, 1], ids=["every_shape_cached", "one_shape_cached"])
def test_repeated_shapes_replay_their_proof_graphs(cap_shapes):
    """Requests of a few shapes, each seen several times: a worker proves the first request of a
    shape eagerly, builds the shape's proof graph on the second and replays it afterwards (with a
    cache of one shape, shapes are dropped and rebuilt as they alternate).  Every proof equals the
    same request proved alone and verifies."""
    reqs = make_requests(12, 41, col_choices=(4, 16), log_range=(6, 7))
    shapes = {(r.n_cols, r.log_n) for r in reqs}
    assert len(shapes) < len(reqs)
    cap = cap_shapes * max(device_bytes(c, lg) for c, lg in shapes)
    server = ProofServer(12, FRI, n_workers=2, hbm_budget=1 << 34, graph_cache_bytes=cap)
    try:
        mixed = server.serve(reqs, realtime=False)
    finally:
        server.close()
    alone = make_requests(12, 41, col_choices=(4, 16), log_range=(6, 7))
    ctx = Context(0)
    be = HipBackend(ctx)
    tw = be.precompute_canonic_twiddles(12)
    prove_sequential(be, tw, alone, FRI, realtime=False)
    for m, a in zip(mixed, alone):
        assert np.array_equal(m.felts, a.felts), f"request {m.index} differs"
        assert PR.wide_fib_verify(m.felts.tolist(), m.n_cols, m.log_n, FRI)
    ctx.close()


# The serving workload itself: column counts from {32, 64, 128, 256}, 2^14..2^20 rows, seed 4,
# production FriParams::new(26, 0, 1, 70, 1) (crates/prover/src/prover.rs:385-396).
C5_FRI = {"pow_bits": 26, "log_blowup_factor": 1, "log_last_layer_degree_bound": 0, "n_queries": 70,
          "fold_step": 1}
_C5_CACHE = {}


def _c5_alone():
    """The six requests proved one at a time on a plain context (computed once per module)."""
    if "alone" not in _C5_CACHE:
        alone = make_requests(6, 4)
        ctx = Context(0)
        be = HipBackend(ctx)
        tw = be.precompute_canonic_twiddles(21)
        prove_sequential(be, tw, alone, C5_FRI, realtime=False)
        tw.free()
        ctx.close()
        _C5_CACHE["alone"] = alone
    return _C5_CACHE["alone"]


def _wide_fib_columns(r):
    cols = [r.c0, r.c1]
    x, y = r.c0.astype(np.uint64), r.c1.astype(np.uint64)
    for _ in range(2, r.n_cols):
        z = (x * x % PR.P + y * y % PR.P) % PR.P
        cols.append(z.astype(np.uint32))
        x, y = y, z
    return cols


@pytest.mark.parametrize("whole_proofs", [False, True], ids=["batched_commits", "whole_proofs"])
def test_serving_workload_proofs_match_alone_and_oracle(whole_proofs):
    """Six requests of the serving workload (make_requests(6, 4): 128 x 2^20, 256 x 2^17,
    256 x 2^20, 256 x 2^14, 64 x 2^18, 64 x 2^16) through the server with an HBM budget that
    admits one of the largest requests at a time: every proof equals the same request proved
    alone and is accepted by the oracle verifier; the two smallest (4M cells each) are
    byte-equal to the oracle prover."""
    reqs = make_requests(6, 4)
    assert any(r.n_cols == 256 and r.log_n == 20 for r in reqs) and any(r.log_n == 14 for r in reqs)
    budget = device_bytes(256, 20) + device_bytes(64, 16)
    print(f"serving 6 requests, whole_proofs={whole_proofs}, budget {budget >> 20} MiB", flush=True)
    server = ProofServer(21, C5_FRI, n_workers=2, hbm_budget=budget, whole_proofs=whole_proofs)
    try:
        mixed = server.serve(reqs, realtime=False)
    finally:
        server.close()
    alone = _c5_alone()
    for m, a in zip(mixed, alone):
        assert np.array_equal(m.felts, a.felts), f"request {m.index} differs from its proof made alone"
        assert PR.wide_fib_verify(m.felts.tolist(), m.n_cols, m.log_n, C5_FRI)
    if whole_proofs:
        return  # the oracle comparison below is the same for both modes (proofs are equal)
    smallest = sorted(mixed, key=lambda r: (r.n_cols << r.log_n, r.index))[:2]
    for r in smallest:
        print(f"oracle prover: request {r.index} ({r.n_cols} x 2^{r.log_n})", flush=True)
        tw, itw = O.canonic_twiddles(r.log_n + 1)
        exp, _ = PR.wide_fib_prove(_wide_fib_columns(r), r.log_n, C5_FRI, tw, itw)
        assert r.felts.tolist() == exp, f"request {r.index} differs from the oracle prover"
