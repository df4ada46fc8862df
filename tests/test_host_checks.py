"""Host-side checks of the prover that need no device: the ConstraintsNotSatisfied sanity check
(crates/stwo/src/prover/mod.rs:123-132) on proofs made by the oracle's prover, honest and broken."""
import pytest

import oracle as O
from oracle import protocol as PR
from proving_amd.prover import ConstraintsNotSatisfied, check_wide_fib_oods
from proving_amd.workloads import wide_fib_trace

P = O.P
FRI = {"pow_bits": 4, "log_blowup_factor": 1, "log_last_layer_degree_bound": 0, "n_queries": 3, "fold_step": 1}


def _oracle_proof(n_cols, log_n, seed, bad_row=None):
    """The oracle prover's proof, and the composition coefficient and OODS point replayed from its
    commitments the way the verifier draws them (crates/stwo/src/core/verifier.rs:19-60)."""
    trace = wide_fib_trace(n_cols, log_n, seed)
    if bad_row is not None:
        trace[n_cols // 2][bad_row] = (int(trace[n_cols // 2][bad_row]) + 1) % P
    tw, itw = O.canonic_twiddles(log_n + 1)
    felts, _ = PR.wide_fib_prove(trace, log_n, FRI, tw, itw)
    proof = PR.deserialize_pcs_proof(felts)
    ch = O.Channel()
    ch.mix_hash(proof["commitments"][0])
    ch.mix_hash(proof["commitments"][1])
    alpha = [int(v) for v in ch.draw_secure_felt()]
    ch.mix_hash(proof["commitments"][2])
    ch.mix_u64(proof["oods_proof_of_work"])
    z = [int(v) for v in ch.random_point()]
    return proof["sampled_values"], alpha, z


@pytest.mark.parametrize("n_cols,log_n", [(4, 4), (9, 5), (16, 6)])
def test_honest_proof_passes_the_oods_check(n_cols, log_n):
    sv, alpha, z = _oracle_proof(n_cols, log_n, 21)
    check_wide_fib_oods(sv, alpha, z, log_n, 1)


def test_bad_trace_fails_the_oods_check():
    sv, alpha, z = _oracle_proof(12, 5, 22, bad_row=7)
    with pytest.raises(ConstraintsNotSatisfied):
        check_wide_fib_oods(sv, alpha, z, 5, 1)


@pytest.mark.parametrize("which", ["trace", "composition", "alpha", "point"])
def test_any_wrong_input_fails_the_oods_check(which):
    """A changed trace sample, composition sample, coefficient or point each breaks the equation
    (a dropped term, a wrong power order or a wrong vanishing coset would as well)."""
    sv, alpha, z = _oracle_proof(8, 5, 23)
    sv = [[list(map(list, col)) for col in tree] for tree in sv]
    if which == "trace":
        sv[1][3][0][0] = (sv[1][3][0][0] + 1) % P
    elif which == "composition":
        sv[2][5][0][2] = (sv[2][5][0][2] + 1) % P
    elif which == "alpha":
        alpha = [(alpha[0] + 1) % P] + alpha[1:]
    else:
        # the canonic coset's vanishing polynomial reads only x (its shift is zero)
        z = [(z[0] + 1) % P] + z[1:]
    with pytest.raises(ConstraintsNotSatisfied):
        check_wide_fib_oods(sv, alpha, z, 5, 1)
