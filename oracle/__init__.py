"""ctypes wrapper over the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle is the checker the HIP backend is
compared against.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg import this module; the product path never does.

All arrays are numpy ``uint32`` (canonical M31 words) or ``uint8`` (hashes).
QM31 values are 4 words, QM31 circle points 8 words (x then y); a QM31 column
is a ``(4, n)`` array (coordinates, as ``SecureColumnByCoords``).
"""
import ctypes
import os
import subprocess

import numpy as np

P = (1 << 31) - 1
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u32p = ctypes.POINTER(ctypes.c_uint32)
u8p = ctypes.POINTER(ctypes.c_uint8)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.or_grind.restype = ctypes.c_uint64
        _lib.or_m31_mul.restype = ctypes.c_uint32
        _lib.or_m31_inv.restype = ctypes.c_uint32
        _lib.or_m31_add.restype = ctypes.c_uint32
        _lib.or_m31_reduce.restype = ctypes.c_uint32
        _lib.or_m31_reduce.argtypes = [ctypes.c_uint64]
        _lib.or_verify_pow_nonce.restype = ctypes.c_int
    return _lib


def _p32(a):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u32p)


def _p8(a):
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u8p)


def _w(vals):
    return np.ascontiguousarray(np.asarray(vals, dtype=np.uint32))


def _ptr_array(arrs):
    return (ctypes.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])


# ---- fields / circle -------------------------------------------------------
def m31_mul(a, b):
    return lib().or_m31_mul(a, b)


def m31_add(a, b):
    return lib().or_m31_add(a, b)


def m31_reduce(v):
    return lib().or_m31_reduce(v)


def m31_inv(a):
    return lib().or_m31_inv(a)


def qm31_mul(a, b):
    out = np.zeros(4, np.uint32)
    lib().or_qm31_mul(_p32(_w(a)), _p32(_w(b)), _p32(out))
    return out


def qm31_inv(a):
    out = np.zeros(4, np.uint32)
    lib().or_qm31_inv(_p32(_w(a)), _p32(out))
    return out


def index_to_point(index):
    out = np.zeros(2, np.uint32)
    lib().or_index_to_point(ctypes.c_uint32(index), _p32(out))
    return out


def subgroup_gen(log_size):
    return 1 << (31 - log_size)


def canonic_half_initial(log_size):
    """Initial index of CanonicCoset(log_size).half_coset = half_odds(log_size - 1)."""
    return subgroup_gen(log_size + 1)


def circle_domain_at(half_initial, half_log, i):
    out = np.zeros(2, np.uint32)
    lib().or_circle_domain_at(half_initial, half_log, ctypes.c_uint64(i), _p32(out))
    return out


# ---- twiddles / FFT ----------------------------------------------------------
def precompute_twiddles(initial_index, log_size):
    n = 1 << log_size
    tw = np.zeros(n, np.uint32)
    itw = np.zeros(n, np.uint32)
    lib().or_precompute_twiddles(initial_index, log_size, _p32(tw), _p32(itw))
    return tw, itw


def canonic_twiddles(domain_log):
    """Twiddle tree of CanonicCoset(domain_log).half_coset."""
    return precompute_twiddles(canonic_half_initial(domain_log), domain_log - 1)


def extract_subdomain_twiddles(buf, domain_log, sub_log):
    out = np.zeros(1 << (sub_log - 1), np.uint32)
    lib().or_extract_subdomain_twiddles(_p32(buf), ctypes.c_uint64(len(buf)), domain_log, sub_log,
                                        _p32(out))
    return out


def interpolate(itw, values, log_size, half_initial=None):
    """Interpolate bit-reversed evaluations on CanonicCoset(log_size) (or the circle domain with
    the given half-coset initial index) -> coefficients."""
    if half_initial is None:
        half_initial = canonic_half_initial(log_size)
    v = np.array(values, dtype=np.uint32, copy=True)
    lib().or_interpolate(_p32(itw), ctypes.c_uint64(len(itw)), half_initial, log_size - 1, _p32(v))
    return v


def evaluate(tw, coeffs, domain_log, half_initial=None):
    if half_initial is None:
        half_initial = canonic_half_initial(domain_log)
    c = _w(coeffs)
    coeff_log = int(len(c)).bit_length() - 1
    out = np.zeros(1 << domain_log, np.uint32)
    lib().or_evaluate(_p32(tw), ctypes.c_uint64(len(tw)), half_initial, domain_log - 1, _p32(c),
                      coeff_log, _p32(out))
    return out


def eval_at_point(coeffs, point):
    c = _w(coeffs)
    out = np.zeros(4, np.uint32)
    lib().or_eval_at_point(_p32(c), int(len(c)).bit_length() - 1, _p32(_w(point)), _p32(out))
    return out


def barycentric_weights(log_size, point):
    w = np.zeros((4, 1 << log_size), np.uint32)
    lib().or_barycentric_weights(log_size, _p32(_w(point)), _p32(w[0]), _p32(w[1]), _p32(w[2]),
                                 _p32(w[3]))
    return w


def barycentric_eval(evals, weights):
    e = _w(evals)
    w = np.ascontiguousarray(weights, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().or_barycentric_eval(_p32(e), int(len(e)).bit_length() - 1, _p32(w[0]), _p32(w[1]),
                              _p32(w[2]), _p32(w[3]), _p32(out))
    return out


# ---- hashing / Merkle --------------------------------------------------------
def blake2s(data: bytes) -> bytes:
    buf = np.frombuffer(bytes(data), dtype=np.uint8).copy() if data else np.zeros(1, np.uint8)
    out = np.zeros(32, np.uint8)
    lib().or_blake2s(_p8(buf), ctypes.c_uint64(len(data)), _p8(out))
    return out.tobytes()


def merkle_build_leaves(cols_sorted_asc, lifting_log_size):
    cols = [_w(c) for c in cols_sorted_asc]
    logs = _w([int(len(c)).bit_length() - 1 for c in cols] or [0])
    n_out = 1 if not cols else 1 << lifting_log_size
    out = np.zeros(32 * n_out, np.uint8)
    lib().or_merkle_build_leaves(_ptr_array(cols), _p32(logs), ctypes.c_uint64(len(cols)),
                                 lifting_log_size, _p8(out))
    return out.reshape(n_out, 32)


def merkle_next_layer(prev):
    prev = np.ascontiguousarray(prev, dtype=np.uint8).reshape(-1, 32)
    prev_log = int(prev.shape[0]).bit_length() - 1
    out = np.zeros((prev.shape[0] // 2, 32), np.uint8)
    lib().or_merkle_next_layer(_p8(prev.reshape(-1)), prev_log, _p8(out.reshape(-1)))
    return out


def merkle_commit_lifted(cols, lifting_log_size):
    """MerkleProverLifted::commit (crates/stwo/src/prover/vcs_lifted/prover.rs:35-80): returns
    layers root-first."""
    cols = sorted([_w(c) for c in cols], key=len)  # stable sort ascending by length
    layers = [merkle_build_leaves(cols, lifting_log_size)]
    for _ in range(lifting_log_size if cols else 0):
        layers.append(merkle_next_layer(layers[-1]))
    return layers[::-1]


def commit_on_layer(log_size, prev, cols):
    cols = [_w(c) for c in cols]
    out = np.zeros((1 << log_size, 32), np.uint8)
    prev_p = None if prev is None else _p8(np.ascontiguousarray(prev, dtype=np.uint8).reshape(-1))
    lib().or_commit_on_layer(log_size, prev_p, _ptr_array(cols), ctypes.c_uint64(len(cols)),
                             _p8(out.reshape(-1)))
    return out


def merkle_commit_mixed(cols):
    """MerkleProver::commit (crates/stwo/src/prover/vcs/prover.rs:39-59): columns stable-sorted
    descending by length; returns layers root-first."""
    cols = sorted([_w(c) for c in cols], key=lambda c: -len(c))
    if not cols:
        return [np.frombuffer(blake2s(b""), np.uint8).reshape(1, 32)]
    max_log = int(len(cols[0])).bit_length() - 1
    layers = []
    prev = None
    i = 0
    for log_size in range(max_log, -1, -1):
        layer_cols = []
        while i < len(cols) and len(cols[i]) == (1 << log_size):
            layer_cols.append(cols[i])
            i += 1
        prev = commit_on_layer(log_size, prev, layer_cols)
        layers.append(prev)
    return layers[::-1]


# ---- FRI ---------------------------------------------------------------------
def fold_line(src, line_initial, log_size, alpha):
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.zeros((4, 1 << (log_size - 1)), np.uint32)
    lib().or_fold_line(_ptr_array(list(src)), line_initial, log_size, _p32(_w(alpha)),
                       _ptr_array(list(dst)))
    return dst


def fold_circle_into_line(src, log_size, alpha, half_initial=None):
    if half_initial is None:
        half_initial = canonic_half_initial(log_size)
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.zeros((4, 1 << (log_size - 1)), np.uint32)
    lib().or_fold_circle_into_line(_ptr_array(list(src)), half_initial, log_size - 1,
                                   _p32(_w(alpha)), _ptr_array(list(dst)))
    return dst


def fri_decompose(src):
    """FriOps::decompose (crates/stwo/src/prover/backend/cpu/fri.rs:39-62) with
    decomposition_coefficient (:134-146) for a QM31 evaluation in bit-reversed order, blowup 2:
    lambda = (sum of the first half - sum of the second half) / N, coordinate by coordinate, and
    g = f - lambda on the first half, f + lambda on the second.  Returns (g as (4, N), lambda)."""
    c = np.asarray(src, dtype=np.int64)
    n = c.shape[1]
    h = n // 2
    inv_n = m31_inv(n)
    lam = [int((c[k, :h].sum() - c[k, h:].sum()) % P) * inv_n % P for k in range(4)]
    g = c.copy()
    col = np.array(lam, np.int64)[:, None]
    g[:, :h] -= col
    g[:, h:] += col
    return (g % P).astype(np.uint32), lam


# ---- quotients / accumulation ---------------------------------------------------
def line_coeffs(point, value, alpha):
    out = np.zeros(12, np.uint32)
    lib().or_line_coeffs(_p32(_w(point)), _p32(_w(value)), _p32(_w(alpha)), _p32(out))
    return out


def accumulate_numerators(cols, col_idx, coeffs, log_blowup):
    """One sample batch: coeffs is (n_terms, 12) (a, b, c)."""
    cols = [_w(c) for c in cols]
    col_log = int(len(cols[0])).bit_length() - 1
    out = np.zeros((4, 1 << (col_log - log_blowup)), np.uint32)
    fl = np.zeros(4, np.uint32)
    co = np.ascontiguousarray(coeffs, dtype=np.uint32).reshape(-1)
    lib().or_accumulate_numerators(_ptr_array(cols), col_log, log_blowup, _p32(_w(col_idx)),
                                   _p32(co), ctypes.c_uint64(len(col_idx)), _ptr_array(list(out)),
                                   _p32(fl))
    return out, fl


def compute_quotients_and_combine(partials, points, first_linear, lifting_log_size, log_blowup,
                                  tw, itw):
    parts = []
    logs = []
    for pcol in partials:
        pcol = np.ascontiguousarray(pcol, dtype=np.uint32)
        parts.extend(list(pcol))
        logs.append(int(pcol.shape[1]).bit_length() - 1)
    out = np.zeros((4, 1 << lifting_log_size), np.uint32)
    pts = _w(np.concatenate([_w(p) for p in points]))
    fls = _w(np.concatenate([_w(f) for f in first_linear]))
    lib().or_compute_quotients_and_combine(_ptr_array(parts), _p32(_w(logs)), _p32(pts), _p32(fls),
                                           ctypes.c_uint64(len(partials)), lifting_log_size,
                                           log_blowup, _p32(tw), _p32(itw),
                                           ctypes.c_uint64(len(tw)), _ptr_array(list(out)))
    return out


def lift_and_accumulate(cols):
    cols = [np.ascontiguousarray(c, dtype=np.uint32) for c in cols]
    flat = []
    for c in cols:
        flat.extend(list(c))
    logs = _w([int(c.shape[1]).bit_length() - 1 for c in cols])
    out = np.zeros((4, cols[-1].shape[1]), np.uint32)
    lib().or_lift_and_accumulate(_ptr_array(flat), _p32(logs), ctypes.c_uint64(len(cols)),
                                 _ptr_array(list(out)))
    return out


# ---- proof of work / channel -------------------------------------------------------
def grind(digest: bytes, n_bits):
    d = np.frombuffer(digest, np.uint8).copy()
    return int(lib().or_grind(_p8(d), n_bits))


def verify_pow_nonce(digest: bytes, n_bits, nonce):
    d = np.frombuffer(digest, np.uint8).copy()
    return bool(lib().or_verify_pow_nonce(_p8(d), n_bits, ctypes.c_uint64(nonce)))


class _ChannelStruct(ctypes.Structure):
    _fields_ = [("digest", ctypes.c_uint8 * 32), ("n_draws", ctypes.c_uint32)]


class Channel:
    """Blake2sChannel (crates/stwo/src/core/channel/blake2s.rs:16-117)."""

    def __init__(self):
        self.s = _ChannelStruct()

    @property
    def digest(self) -> bytes:
        return bytes(self.s.digest)

    @property
    def n_draws(self):
        return self.s.n_draws

    def mix_u32s(self, words):
        w = _w(list(words) or [0])
        lib().or_channel_mix_u32s(ctypes.byref(self.s), _p32(w), ctypes.c_uint64(len(words)))

    def mix_u64(self, v):
        lib().or_channel_mix_u64(ctypes.byref(self.s), ctypes.c_uint64(v))

    def mix_felts(self, felts):
        f = _w(np.asarray(felts, dtype=np.uint32).reshape(-1))
        lib().or_channel_mix_felts(ctypes.byref(self.s), _p32(f), ctypes.c_uint64(len(f) // 4))

    def mix_hash(self, root: bytes):
        r = np.frombuffer(bytes(root), np.uint8).copy()
        lib().or_channel_mix_hash(ctypes.byref(self.s), _p8(r))

    def draw_u32s(self):
        out = np.zeros(8, np.uint32)
        lib().or_channel_draw_u32s(ctypes.byref(self.s), _p32(out))
        return out

    def draw_secure_felt(self):
        out = np.zeros(4, np.uint32)
        lib().or_channel_draw_secure_felt(ctypes.byref(self.s), _p32(out))
        return out

    def random_point(self):
        out = np.zeros(8, np.uint32)
        lib().or_channel_random_point(ctypes.byref(self.s), _p32(out))
        return out


# ---- synthetic inputs ----------------------------------------------------------------
def splitmix64_columns(seed, log_sizes):
    """The shared input generator (proving_amd.workloads), re-exported for the tests."""
    from proving_amd.workloads import splitmix64_columns as gen
    return gen(seed, log_sizes)


# ---------------------------------------------------------------- SimdBackend restatement
_SIMD_PATH = os.path.join(_HERE, "libsimdport.so")
_simd = None


def simd_port_available():
    """True when libsimdport.so is built and this CPU has AVX-512F."""
    global _simd
    if _simd is None:
        if not os.path.exists(_SIMD_PATH):
            build()
        _simd = ctypes.CDLL(_SIMD_PATH)
        _simd.sp_commit_lifted.restype = ctypes.c_int
    return bool(_simd.sp_cpu_has_avx512())


def _aligned_copy(a):
    """uint32 copy of a whose data starts on a 64-byte boundary."""
    a = np.asarray(a, dtype=np.uint32)
    raw = np.empty(a.size + 16, np.uint32)
    off = (-raw.ctypes.data % 64) // 4
    out = raw[off:off + a.size]
    out[:] = a
    return out


def simd_port_commit(cols, logs, log_blowup, lifting_log_size, nthreads, reps=1):
    """The trace commitment (interpolate -> LDE -> lifted Merkle root) by the multi-threaded
    AVX-512 restatement of the SimdBackend path.  cols: bit-reversed evaluations, logs ascending
    and >= 4.  Returns (root bytes, (fft_s, leaves_s, layers_s)) of the fastest of reps commits."""
    assert simd_port_available(), "this CPU has no AVX-512F"
    ok = (len(logs) > 0 and 5 <= lifting_log_size <= 28 and all(lg >= 4 for lg in logs)
          and list(logs) == sorted(logs) and max(logs) + log_blowup <= lifting_log_size)
    if not ok:
        raise ValueError("simd_port_commit: logs must be ascending, >= 4, and fit the tree height")
    arrs = [_aligned_copy(c) for c in cols]
    ptrs = (u32p * len(arrs))(*[a.ctypes.data_as(u32p) for a in arrs])
    lg = np.array(logs, np.uint32)
    tw, itw = canonic_twiddles(lifting_log_size)
    root = np.zeros(32, np.uint8)
    times = (ctypes.c_double * 3)()
    rc = _simd.sp_commit_lifted(ptrs, _p32(lg), len(arrs), log_blowup, lifting_log_size, _p32(tw),
                                _p32(itw), ctypes.c_uint64(len(tw)), nthreads, reps, _p8(root), times)
    if rc != 0:
        raise ValueError(f"sp_commit_lifted rejected the input ({rc})")
    return root.tobytes(), tuple(times)
