"""The benchmark's CPU baseline (oracle/simd_port.c, a multi-threaded AVX-512 restatement of the
SimdBackend commit path) against the scalar oracle: the same lifted Merkle root for mixed column
sizes, blowups 1 and 2, a lifting height above the largest column, and any thread count.  The
root hashes every LDE word, so a wrong twiddle, layer order, scale or lifting map changes it."""
import numpy as np
import pytest

import oracle as O
from proving_amd.workloads import splitmix64_columns

pytestmark = pytest.mark.skipif(not O.simd_port_available(), reason="this host has no AVX-512F")

CASES = [
    ([4], 1, None),                                   # in-register layers only
    ([4, 5], 1, None),
    ([5] * 3 + [6] * 17 + [9] * 2, 1, None),          # a pending block carried across groups
    ([4] * 20 + [12] * 5 + [13] * 3, 1, None),        # block passes and one radix-2 pass above
    ([8] * 5, 2, None),                               # blowup 4: two replicated layers
    ([6] * 5 + [8] * 3, 1, 11),                       # states lifted to a taller tree at the end
    ([13] * 16, 1, None),                             # exactly one full block per row
    ([12] * 33 + [14] * 16 + [16] * 3, 1, None),
]


_ROOTS = {}


def _oracle_root(host, logs, b, L):
    key = (tuple(logs), b, L, tuple(int(c.sum()) for c in host))
    if key in _ROOTS:
        return _ROOTS[key]
    tw, itw = O.canonic_twiddles(L)
    ldes = [O.evaluate(tw, O.interpolate(itw, c, lg), lg + b) for c, lg in zip(host, logs)]
    _ROOTS[key] = O.merkle_commit_lifted(ldes, L)[0][0].tobytes()
    return _ROOTS[key]


@pytest.mark.parametrize("logs,b,L", CASES, ids=[f"case{i}" for i in range(len(CASES))])
@pytest.mark.parametrize("threads", [1, 3])
def test_root_equals_oracle(logs, b, L, threads):
    L = L or max(logs) + b
    host = splitmix64_columns(7, logs)
    got, times = O.simd_port_commit(host, logs, b, L, threads)
    assert got == _oracle_root(host, logs, b, L)
    assert all(t >= 0 for t in times)


def test_changed_value_changes_root():
    logs = [6] * 4 + [9] * 3
    host = splitmix64_columns(3, logs)
    a, _ = O.simd_port_commit(host, logs, 1, 10, 2)
    host[5] = host[5].copy()
    host[5][17] = (int(host[5][17]) + 1) % O.P
    b, _ = O.simd_port_commit(host, logs, 1, 10, 2)
    assert a != b and b == _oracle_root(host, logs, 1, 10)


@pytest.mark.parametrize("logs,b,L", [([3], 1, 4), ([6, 5], 1, 7), ([6], 1, 6), ([6], 1, 29)])
def test_rejects_unsupported_shapes(logs, b, L):
    host = [np.zeros(1 << lg, np.uint32) for lg in logs]
    with pytest.raises(ValueError):
        O.simd_port_commit(host, logs, b, L, 1)
