"""Synthetic inputs and the benchmark/parity workload shapes.

The generator is shared by the HIP path, the tests and the CPU baseline so
that all of them see identical columns: SplitMix64 seeded with the workload
seed, v = next_u64() >> 33, rejecting v = 2^31 - 1 (uniform in [0, 2^31 - 2]),
filled column-major in the listed column order.  Columns are bit-reversed
evaluations on CanonicCoset(log_size), as a trace generator produces them.
"""
import numpy as np

P = (1 << 31) - 1
_M64 = (1 << 64) - 1
_GAMMA = 0x9E3779B97F4A7C15

# Plain commit: 64 columns x 2^14, seed 0, log_blowup 1.
PLAIN_COMMIT = {"name": "plain_commit_64x2^14", "logs": [14] * 64, "seed": 0, "log_blowup": 1}
# Batched uneven commit: 512 columns (256 @ 2^12, 128 @ 2^16, 96 @ 2^18, 32 @ 2^20), seed 1.
UNEVEN_COMMIT = {"name": "batched_uneven_commit_512cols_log12-20",
                 "logs": [12] * 256 + [16] * 128 + [18] * 96 + [20] * 32, "seed": 1, "log_blowup": 1}


def splitmix64_columns(seed, log_sizes):
    cols = []
    state = seed & _M64
    for lg in log_sizes:
        n = 1 << lg
        out = np.empty(n, np.uint32)
        i = 0
        while i < n:
            k = n - i
            idx = np.arange(1, k + 1, dtype=np.uint64)
            with np.errstate(over="ignore"):
                z = np.uint64(state) + idx * np.uint64(_GAMMA)
                z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
                z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
                z = z ^ (z >> np.uint64(31))
            v = (z >> np.uint64(33)).astype(np.uint32)
            ok = np.nonzero(v != P)[0][: n - i]
            out[i:i + len(ok)] = v[ok]
            i += len(ok)
            consumed = int(ok[-1]) + 1 if i == n else k
            state = (state + consumed * _GAMMA) & _M64
        cols.append(out)
    return cols


def cells(log_sizes):
    return sum(1 << lg for lg in log_sizes)


# FRI + DEEP phase: 200 columns x 2^21 (LDE 2^22), seed 2, 3 out-of-domain points from
# Blake2sChannel::default() + mix_u64(2), FriParams{pow 0, blowup 1, last layer 3, 32 queries}.
FRI_DEEP = {"name": "fri_deep_200x2^21", "n_cols": 200, "log_size": 21, "seed": 2, "n_points": 3,
            "fri": {"pow_bits": 0, "log_blowup_factor": 1, "log_last_layer_degree_bound": 3,
                    "n_queries": 32, "fold_step": 1}}
# Full prove of the wide-Fibonacci AIR: 256 columns x 2^20, seed 3, FriParams::new(26, 0, 1, 70, 1).
WIDE_FIB = {"name": "wide_fib_256x2^20", "n_cols": 256, "log_size": 20, "seed": 3,
            "fri": {"pow_bits": 26, "log_blowup_factor": 1, "log_last_layer_degree_bound": 0,
                    "n_queries": 70, "fold_step": 1}}


def wide_fib_trace(n_cols, log_size, seed):
    """generate_trace (crates/examples/src/wide_fibonacci/mod.rs:27-48) with row inputs
    (c0, c1) drawn from the shared generator (two columns of 2^log_size): c_k = c_{k-2}^2 +
    c_{k-1}^2 for k >= 2.  Returns n_cols uint32 columns (bit-reversed evaluations on
    CanonicCoset(log_size))."""
    a, b = splitmix64_columns(seed, [log_size, log_size])
    cols = [a, b]
    x, y = a.astype(np.uint64), b.astype(np.uint64)
    for _ in range(2, n_cols):
        z = (x * x % P + y * y % P) % P
        cols.append(z.astype(np.uint32))
        x, y = y, z
    return cols
