"""The proof-serving loop: a mix of requests proved with cross-request trace commits and
overlapping worker contexts gives, for every request, exactly the proof made alone; the oracle
verifier accepts them."""
import numpy as np
import pytest

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
