"""Host-side mirror of the Stwo backend trait family over the HIP C-ABI.

``HipBackend`` exposes the reference's trait methods under their reference
names (``PolyOps``, ``MerkleOpsLifted``, ``MerkleOps``, ``FriOps``,
``QuotientOps``, ``AccumulationOps``, ``GrindOps``; see
crates/stwo/src/prover/backend/mod.rs:18-66) on device-resident columns.
Argument meaning and error behaviour follow the reference: where it would
panic (``assert!``) the C-ABI returns SH_ERR_INVALID and this module raises
``HipError``.
"""
import contextlib
import ctypes
import json
import os
import sys

import numpy as np

from ._lib import HipError, load

P = (1 << 31) - 1


def _ptrs(bufs):
    if isinstance(bufs, (ColumnRefs, HashLayerSet)):
        return bufs.c_ptrs
    arr = (ctypes.c_void_p * max(1, len(bufs)))(*[b.ptr for b in bufs])
    return arr


def _logs(cols):
    """Log sizes of a list of columns or of a ColumnSet (its cached array)."""
    if isinstance(cols, ColumnRefs):
        return cols.log_sizes
    return [c.log_size for c in cols]


def _ptr_array(cols):
    """Device addresses of a list of columns or of a ColumnSet, as a uint64 array."""
    if isinstance(cols, ColumnRefs):
        return cols.ptrs
    return np.array([c.ptr for c in cols], np.uint64)


def _u32s(vals):
    a = np.asarray(vals, dtype=np.uint32).ravel()
    if a.size == 0:
        a = np.zeros(1, np.uint32)
    return a.ctypes.data_as(ctypes.c_void_p)  # the pointer object keeps the array alive


class Context:
    """A device, its stream and staging state (one per host thread)."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        rc = self.lib.sh_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise HipError(f"sh_ctx_create({device}) failed with code {rc}")
        self.h = h
        self._closed = False

    def check(self, rc, what=""):
        if rc != 0:
            msg = self.lib.sh_last_error(self.h).decode(errors="replace")
            raise HipError(f"{what}: code {rc}: {msg}")

    def close(self):
        if not self._closed:
            self.lib.sh_ctx_destroy(self.h)
            self._closed = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self):
        self.check(self.lib.sh_sync(self.h), "sh_sync")

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def profile(self, on=True):
        self.check(self.lib.sh_profile_enable(self.h, 1 if on else 0), "profile")

    def profile_report(self):
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        self.check(self.lib.sh_profile_report(self.h, buf, len(buf)), "profile_report")
        return json.loads(buf.value.decode())

    def timer_start(self):
        self.check(self.lib.sh_timer_start(self.h), "timer_start")

    def timer_stop(self):
        ms = ctypes.c_float()
        self.check(self.lib.sh_timer_stop(self.h, ctypes.byref(ms)), "timer_stop")
        return ms.value


class DeviceBuffer:
    """An owned device allocation."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        ctx.check(ctx.lib.sh_alloc(ctx.h, self.nbytes, ctypes.byref(p)), "sh_alloc")
        self.ptr = p.value

    def free(self):
        if self.ptr is not None and not self.ctx._closed:
            self.ctx.lib.sh_free(self.ctx.h, ctypes.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        # back to the context's pool (stream-ordered reuse, no device sync)
        try:
            self.free()
        except Exception:
            pass

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx.check(self.ctx.lib.sh_upload(self.ctx.h, self.ptr, arr.ctypes.data, arr.nbytes),
                       "sh_upload")

    def download(self, dtype=np.uint32, count=None):
        item = np.dtype(dtype).itemsize
        n = self.nbytes // item if count is None else count
        out = np.empty(n, dtype)
        self.ctx.check(self.ctx.lib.sh_download(self.ctx.h, out.ctypes.data, self.ptr, n * item),
                       "sh_download")
        return out


class BaseColumn(DeviceBuffer):
    """Col<HipBackend, BaseField>: 2^log_size canonical M31 words on the device."""

    def __init__(self, ctx, log_size):
        super().__init__(ctx, 4 << log_size)
        self.log_size = log_size

    def __len__(self):
        return 1 << self.log_size

    @classmethod
    def from_cpu(cls, ctx, values):
        v = np.ascontiguousarray(values, dtype=np.uint32)
        n = len(v)
        assert n and n & (n - 1) == 0, "column length must be a power of two"
        c = cls(ctx, n.bit_length() - 1)
        c.upload(v)
        return c

    def to_cpu(self):
        return self.download(np.uint32, len(self))

    def copy_from(self, other):
        assert other.nbytes <= self.nbytes
        self.ctx.check(self.ctx.lib.sh_copy(self.ctx.h, self.ptr, other.ptr, other.nbytes), "sh_copy")
        return self

    def bit_reverse(self):
        self.ctx.check(self.ctx.lib.sh_bit_reverse_column(self.ctx.h, self.ptr, self.log_size),
                       "bit_reverse_column")


class ColumnView(BaseColumn):
    """A column inside a ColumnSet's allocation.  It borrows the memory: free() only drops the
    pointer, and the allocation is released once neither the set nor any view refers to it.
    A view holds the allocation, not the set, so a set and its views form no reference cycle
    (which would hold the memory until the cyclic garbage collector runs)."""

    def __init__(self, owner, ptr, log_size):
        self.ctx = owner.ctx
        self.nbytes = 4 << log_size
        self.ptr = ptr
        self.log_size = log_size
        self._buf = owner.buf

    def free(self):
        self.ptr = None

    def __del__(self):
        pass


class ColumnRefs:
    """Device columns given by address and log size (uint64 / uint32 arrays), not owned: the
    batched entry points take them without building per-column objects.  The caller keeps the
    columns alive."""

    def __init__(self, ctx, ptrs, log_sizes):
        self.ctx = ctx
        self.ptrs = np.ascontiguousarray(ptrs, dtype=np.uint64).reshape(-1)
        self.log_sizes = np.ascontiguousarray(log_sizes, dtype=np.uint32).reshape(-1)
        assert len(self.ptrs) == len(self.log_sizes)
        n = len(self.ptrs)
        self.c_ptrs = (ctypes.c_void_p * max(1, n)).from_buffer(self.ptrs if n else np.zeros(1, np.uint64))

    def __len__(self):
        return len(self.log_sizes)


class ColumnSet(ColumnRefs):
    """Many device columns in one pooled allocation (the role of BaseColumnPool,
    crates/stwo/src/prover/mempool.rs:14-58, for a whole tree's columns): one sh_alloc and one
    sh_free for all of them, and the pointer and log-size arrays that the batched entry points take,
    built once.  Indexing and iteration give ColumnView objects (created on first use).  A view's
    log_size must not change (split_at_mid works on owned columns only)."""

    def __init__(self, ctx, log_sizes):
        self.ctx = ctx
        self.log_sizes = np.ascontiguousarray(log_sizes, dtype=np.uint32).reshape(-1)
        n = len(self.log_sizes)
        sizes = np.left_shift(np.uint64(4), self.log_sizes.astype(np.uint64))
        sizes = (sizes + np.uint64(255)) // np.uint64(256) * np.uint64(256)  # 256-byte aligned columns
        offs = np.zeros(n, np.uint64)
        if n > 1:
            np.cumsum(sizes[:-1], out=offs[1:])
        self.buf = DeviceBuffer(ctx, int(sizes.sum()) if n else 256)
        super().__init__(ctx, offs + np.uint64(self.buf.ptr), self.log_sizes)
        self._views = None

    def _all(self):
        if self._views is None:
            self._views = [ColumnView(self, int(p), int(lg)) for p, lg in zip(self.ptrs, self.log_sizes)]
        return self._views

    def __len__(self):
        return len(self.log_sizes)

    def __getitem__(self, i):
        return self._all()[i]

    def __iter__(self):
        return iter(self._all())

    def __add__(self, other):
        return list(self) + list(other)

    def __radd__(self, other):
        return list(other) + list(self)

    def free(self):
        """Returns the whole allocation to the context pool (every view becomes invalid)."""
        self.buf.free()
        self._views = None


class SecureColumnByCoords:
    """QM31 column stored as 4 coordinate columns (crates/stwo/src/prover/secure_column.rs:13-63)."""

    def __init__(self, ctx, log_size, columns=None):
        self.ctx = ctx
        self.log_size = log_size
        self.columns = columns or [BaseColumn(ctx, log_size) for _ in range(4)]

    @classmethod
    def from_cpu(cls, ctx, values):
        v = np.ascontiguousarray(values, dtype=np.uint32)
        assert v.shape[0] == 4
        return cls(ctx, v.shape[1].bit_length() - 1, [BaseColumn.from_cpu(ctx, v[c]) for c in range(4)])

    def to_cpu(self):
        return np.stack([c.to_cpu() for c in self.columns])

    def __len__(self):
        return 1 << self.log_size

    def free(self):
        for c in self.columns:
            c.free()


class HashLayer(DeviceBuffer):
    """Col<HipBackend, Blake2sHash>: 2^log_size 32-byte hashes."""

    def __init__(self, ctx, log_size):
        super().__init__(ctx, 32 << log_size)
        self.log_size = log_size

    def to_cpu(self):
        return self.download(np.uint8, 32 << self.log_size).reshape(-1, 32)


class HashLayerView(HashLayer):
    """A layer inside a HashLayerSet's allocation (borrowed, like ColumnView)."""

    def __init__(self, owner, ptr, log_size):
        self.ctx = owner.ctx
        self.nbytes = 32 << log_size
        self.ptr = ptr
        self.log_size = log_size
        self._buf = owner.buf  # the allocation, not the set: no reference cycle

    def free(self):
        self.ptr = None

    def __del__(self):
        pass


class HashLayerSet:
    """The layers of a Merkle tree of height L (layer k: 2^k hashes, root-first indexing) in one
    pooled allocation; indexing gives HashLayerView objects."""

    def __init__(self, ctx, L):
        self.ctx = ctx
        self.L = L
        sizes = [max(256, 32 << k) for k in range(L + 1)]
        offs = np.concatenate([[0], np.cumsum(sizes[:-1])]).astype(np.uint64)
        self.buf = DeviceBuffer(ctx, int(sum(sizes)))
        self.ptrs = np.ascontiguousarray(offs + np.uint64(self.buf.ptr))
        self.c_ptrs = (ctypes.c_void_p * (L + 1)).from_buffer(self.ptrs)
        self._views = [HashLayerView(self, int(p), k) for k, p in enumerate(self.ptrs)]

    def __len__(self):
        return self.L + 1

    def __getitem__(self, i):
        return self._views[i]

    def __iter__(self):
        return iter(self._views)

    def free(self):
        self.buf.free()


class TwiddleTree:
    """TwiddleTree<HipBackend> (crates/stwo/src/prover/poly/twiddles.rs:10-15)."""

    def __init__(self, ctx, root_initial_index, root_log_size):
        self.ctx = ctx
        self.root_initial_index = root_initial_index
        self.root_log_size = root_log_size
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.sh_twiddles_create(ctx.h, root_initial_index, root_log_size, ctypes.byref(h)),
                  "precompute_twiddles")
        self.h = h

    def to_cpu(self):
        n = 1 << self.root_log_size
        tw = np.empty(n, np.uint32)
        itw = np.empty(n, np.uint32)
        self.ctx.check(self.ctx.lib.sh_twiddles_download(self.ctx.h, self.h, tw.ctypes.data,
                                                         itw.ctypes.data), "twiddles_download")
        return tw, itw

    def free(self):
        if self.h is not None and not self.ctx._closed:
            self.ctx.lib.sh_twiddles_destroy(self.ctx.h, self.h)
        self.h = None


class CommitGraph:
    """A captured batched commitment; launch() replays it on the context stream."""

    def __init__(self, ctx, h, layers):
        self.ctx, self.h, self.layers = ctx, h, layers

    def launch(self):
        self.ctx.check(self.ctx.lib.sh_graph_launch(self.ctx.h, self.h), "graph_launch")

    def root(self):
        self.ctx.sync()
        return self.layers[0].to_cpu()[0].tobytes()

    def free(self):
        if self.h is not None and not self.ctx._closed:
            self.ctx.lib.sh_graph_destroy(self.ctx.h, self.h)
        self.h = None


class Graph:
    """A captured sequence of entry points (sh_graph_create and friends): the calls made inside
    warmup() run eagerly on the graph's own staging ring, scratch and device arena; the same calls
    made inside capture() (with the warm-up's buffers freed again first) are recorded, and
    launch() replays them.  Buffers allocated inside either belong to the graph's arena: free()
    releases them, so objects holding them must be dropped first."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.sh_graph_create(ctx.h, ctypes.byref(h)), "graph_create")
        self.h = h.value
        self.captured = False

    def _phase(self, begin, end, what):
        @contextlib.contextmanager
        def scope():
            self.ctx.check(begin(self.ctx.h, self.h), what + "_begin")
            try:
                yield self
            except BaseException:
                end(self.ctx.h, self.h)  # close the phase; the first error is the one reported
                raise
            self.ctx.check(end(self.ctx.h, self.h), what + "_end")
        return scope()

    def warmup(self):
        return self._phase(self.ctx.lib.sh_graph_warmup_begin, self.ctx.lib.sh_graph_warmup_end, "graph_warmup")

    def capture(self):
        lib = self.ctx.lib

        def end(ctx_h, h):
            rc = lib.sh_graph_capture_end(ctx_h, h)
            self.captured = rc == 0
            return rc
        return self._phase(lib.sh_graph_capture_begin, end, "graph_capture")

    def launch(self):
        self.ctx.check(self.ctx.lib.sh_graph_launch(self.ctx.h, self.h), "graph_launch")

    def free(self):
        if self.h and not self.ctx._closed:
            self.ctx.check(self.ctx.lib.sh_graph_destroy(self.ctx.h, self.h), "graph_destroy")
        self.h = None


def canonic_half_initial(log_size):
    """CanonicCoset(log_size).half_coset initial index (half_odds(log_size - 1))."""
    return 1 << (31 - (log_size + 1))


class HipBackend:
    """The MI355X backend: one method per reference trait method."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.lib = ctx.lib
        self.h = ctx.h

    def _chk(self, rc, what):
        self.ctx.check(rc, what)

    def release_graphs(self):
        """Frees the graphs the prover keeps on this backend (a proof graph, a commit graph) and
        the buffers they own; call before closing the context."""
        pg = self.__dict__.pop("_proof_graph", None)
        if pg is not None:
            pg.free()
        cg = self.__dict__.pop("_commit_graph", None)
        if cg is not None:
            cg["graph"].free()

    # ---- PolyOps (crates/stwo/src/prover/poly/circle/ops.rs:18-165) -------------------------
    def precompute_twiddles(self, root_initial_index, root_log_size):
        return TwiddleTree(self.ctx, root_initial_index, root_log_size)

    def precompute_canonic_twiddles(self, domain_log):
        """Twiddles of CanonicCoset(domain_log).half_coset."""
        return self.precompute_twiddles(canonic_half_initial(domain_log), domain_log - 1)

    def interpolate(self, col, twiddles):
        """In place: bit-reversed evaluations on CanonicCoset(col.log_size) -> coefficients."""
        self._chk(self.lib.sh_interpolate(self.h, twiddles.h, col.ptr, col.log_size), "interpolate")
        return col

    def interpolate_columns(self, cols, twiddles):
        self._chk(self.lib.sh_interpolate_columns(self.h, twiddles.h, _ptrs(cols),
                                                  _u32s(_logs(cols)), len(cols)),
                  "interpolate_columns")
        return cols

    def evaluate(self, coeffs, domain_log, twiddles, out=None):
        out = out or BaseColumn(self.ctx, domain_log)
        self._chk(self.lib.sh_evaluate(self.h, twiddles.h, coeffs.ptr, coeffs.log_size, out.ptr,
                                       domain_log), "evaluate")
        return out

    def evaluate_polynomials(self, polys, log_blowup, twiddles, outs=None):
        outs = outs if outs is not None else ColumnSet(self.ctx, np.asarray(_logs(polys), np.uint32) + log_blowup)
        self._chk(self.lib.sh_evaluate_polynomials(self.h, twiddles.h, _ptrs(polys),
                                                   _u32s(_logs(polys)), len(polys),
                                                   log_blowup, _ptrs(outs)), "evaluate_polynomials")
        return outs

    # ---- MerkleOpsLifted (crates/stwo/src/prover/vcs_lifted/ops.rs:10-20) ----------------------
    def build_leaves(self, cols_sorted_asc, lifting_log_size):
        out = HashLayer(self.ctx, lifting_log_size if cols_sorted_asc else 0)
        self._chk(self.lib.sh_merkle_build_leaves(self.h, _ptrs(cols_sorted_asc),
                                                  _u32s(_logs(cols_sorted_asc)),
                                                  len(cols_sorted_asc), lifting_log_size, out.ptr),
                  "build_leaves")
        return out

    def build_next_layer(self, prev):
        out = HashLayer(self.ctx, prev.log_size - 1)
        self._chk(self.lib.sh_merkle_build_next_layer(self.h, prev.ptr, prev.log_size, out.ptr),
                  "build_next_layer")
        return out

    def commit_lifted(self, cols, lifting_log_size, layers=None, want_root=True):
        """MerkleProverLifted::commit: returns (layers root-first, root bytes or None)."""
        layers = layers if layers is not None else HashLayerSet(self.ctx, lifting_log_size)
        root = (ctypes.c_uint8 * 32)() if want_root else None
        self._chk(self.lib.sh_merkle_commit_lifted(self.h, _ptrs(cols), _u32s(_logs(cols)),
                                                   len(cols), lifting_log_size, _ptrs(layers), root),
                  "commit_lifted")
        return layers, (bytes(root) if want_root else None)

    # ---- MerkleOps (crates/stwo/src/prover/vcs/ops.rs:8-29) ------------------------------------
    def commit_on_layer(self, log_size, prev, cols, out=None):
        out = out if out is not None else HashLayer(self.ctx, log_size)
        self._chk(self.lib.sh_merkle_commit_on_layer(self.h, log_size, prev.ptr if prev else None,
                                                     _ptrs(cols), len(cols), out.ptr), "commit_on_layer")
        return out

    def commit_mixed(self, cols):
        max_log = max(c.log_size for c in cols) if cols else 0
        layers = [HashLayer(self.ctx, k) for k in range(max_log + 1)]
        root = (ctypes.c_uint8 * 32)()
        self._chk(self.lib.sh_merkle_commit_mixed(self.h, _ptrs(cols), _u32s(_logs(cols)),
                                                  len(cols), _ptrs(layers), root), "commit_mixed")
        return layers, bytes(root)

    def commit_trace_graph(self, evals, log_blowup, twiddles, coeffs, ldes, layers, mixed=False):
        """Captures commit_trace (mixed: its mixed-degree tree variant) on these buffers as a HIP
        graph; returns a CommitGraph."""
        h = ctypes.c_void_p()
        fn = self.lib.sh_commit_trace_mixed_graph_create if mixed else self.lib.sh_commit_trace_graph_create
        self._chk(fn(self.h, twiddles.h, _ptrs(evals), _ptrs(coeffs), _u32s(_logs(evals)), len(evals),
                     log_blowup, _ptrs(ldes), _ptrs(layers), ctypes.byref(h)), "commit_trace_graph_create")
        g = CommitGraph(self.ctx, h, layers)
        g.buffers = (list(evals), list(coeffs), list(ldes), twiddles)  # the graph reads them on every launch
        return g

    # ---- pass-shape tuning (no reference counterpart) --------------------------------------------
    def autotune_fft(self, log_sizes, log_blowup, twiddles):
        self._chk(self.lib.sh_autotune_fft(self.h, twiddles.h, _u32s(log_sizes), len(log_sizes), log_blowup),
                  "autotune_fft")

    def load_fft_tuning(self, path=None):
        """Pass shapes per transform log size from the table proving_amd.autotune writes
        (fft_tuning_gfx950.json beside this file by default); returns False when there is none."""
        path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_tuning_gfx950.json")
        if not os.path.exists(path):
            return False
        with open(path) as f:
            table = json.load(f)["shapes"]
        for m, t in table.items():
            self.set_fft_tuning(int(m), *[int(x) for x in t])
        return True

    def fft_tuning(self, log_size):
        out = (ctypes.c_uint32 * 4)()
        self._chk(self.lib.sh_get_fft_tuning(self.h, log_size, out), "get_fft_tuning")
        return tuple(out)

    def set_fft_tuning(self, log_size, block_log, stride_layers, r_block, r_stride):
        self._chk(self.lib.sh_set_fft_tuning(self.h, log_size, block_log, stride_layers, r_block, r_stride),
                  "set_fft_tuning")

    def set_fft_fusion(self, mode):
        """Interpolation + extension in commit_trace: 0 separate transforms, 1 the interpolation's
        last strided pass fused with the extension's first where their plans share those layers
        (default), 2 the extension planned with the interpolation's layer split (always fused)."""
        self._chk(self.lib.sh_set_fft_fusion(self.h, mode), "set_fft_fusion")

    # ---- batched trace commitment (crates/stwo/src/prover/pcs/mod.rs:325-386) -------------------
    def commit_trace(self, evals, log_blowup, twiddles, coeffs=None, ldes=None, layers=None,
                     want_root=True, mixed=False):
        """Batched interpolate + LDE + Merkle commitment (sh_commit_trace); mixed=True builds the
        mixed-degree tree (commit_on_layer) instead of the lifted one (sh_commit_trace_mixed)."""
        logs = np.asarray(_logs(evals), np.uint32)
        coeffs = coeffs if coeffs is not None else ColumnSet(self.ctx, logs)
        ldes = ldes if ldes is not None else ColumnSet(self.ctx, logs + log_blowup)
        L = int(logs.max()) + log_blowup
        layers = layers if layers is not None else HashLayerSet(self.ctx, L)
        root = (ctypes.c_uint8 * 32)() if want_root else None
        fn = self.lib.sh_commit_trace_mixed if mixed else self.lib.sh_commit_trace
        self._chk(fn(self.h, twiddles.h, _ptrs(evals), _ptrs(coeffs), _u32s(_logs(evals)), len(evals), log_blowup,
                     _ptrs(ldes), _ptrs(layers), root), "commit_trace")
        return coeffs, ldes, layers, (bytes(root) if want_root else None)


def _u32arr(vals):
    a = np.ascontiguousarray(np.asarray(vals, dtype=np.uint32).reshape(-1))
    return a


class _Ops:
    """Trait methods beyond PolyOps / Merkle (bound onto HipBackend below)."""

    # ---- PolyOps::split_at_mid (crates/stwo/src/prover/poly/circle/ops.rs:145-165) ------------
    def split_at_mid(self, poly):
        """Consumes poly (its buffer becomes the left half); returns (left, right)."""
        right = BaseColumn(self.ctx, poly.log_size - 1)
        self._chk(self.lib.sh_split_at_mid(self.h, poly.ptr, poly.log_size, poly.ptr, right.ptr), "split_at_mid")
        poly.log_size -= 1
        return poly, right

    # ---- PolyOps: OOD sampling (crates/stwo/src/prover/poly/circle/ops.rs:44-80) ---------------
    def eval_at_point(self, coeffs, point):
        out = (ctypes.c_uint32 * 4)()
        pt = _u32s(list(point))
        self._chk(self.lib.sh_eval_at_point(self.h, coeffs.ptr, coeffs.log_size, pt, out), "eval_at_point")
        return np.array(out, np.uint32)

    def eval_at_points(self, coeffs, points, as_array=False):
        """eval_at_point of every polynomial at every point, one batched call; returns
        out[i][j] = coeffs[i] at points[j] (QM31 as 4 ints; a (n, k, 4) uint32 array when
        as_array)."""
        n, k = len(coeffs), len(points)
        out = np.empty(max(1, 4 * n * k), np.uint32)
        self._chk(self.lib.sh_eval_at_points(self.h, _ptrs(coeffs), _u32s(_logs(coeffs)), n,
                                             _u32s(points), k, out.ctypes.data), "eval_at_points")
        out = out[:4 * n * k].reshape(n, k, 4)
        return out if as_array else out.tolist()

    def barycentric_weights(self, log_size, point):
        w = SecureColumnByCoords(self.ctx, log_size)
        self._chk(self.lib.sh_barycentric_weights(self.h, log_size, _u32s(list(point)), _ptrs(w.columns)),
                  "barycentric_weights")
        return w

    def barycentric_eval_at_point(self, evals, weights):
        out = (ctypes.c_uint32 * 4)()
        self._chk(self.lib.sh_barycentric_eval_at_point(self.h, evals.ptr, evals.log_size,
                                                        _ptrs(weights.columns), out), "barycentric_eval")
        return np.array(out, np.uint32)

    # ---- FriOps (crates/stwo/src/prover/fri.rs:27-76) -----------------------------------------
    def fold_line(self, src, line_initial_index, alphas, twiddles, dst=None):
        f = len(alphas)
        dst = dst or SecureColumnByCoords(self.ctx, src.log_size - f)
        a = _u32s([w for al in alphas for w in al])
        self._chk(self.lib.sh_fold_line(self.h, twiddles.h, _ptrs(src.columns), line_initial_index,
                                        src.log_size, a, f, _ptrs(dst.columns)), "fold_line")
        return dst

    def fold_circle_into_line(self, src, alpha, twiddles, dst=None):
        dst = dst or SecureColumnByCoords(self.ctx, src.log_size - 1)
        self._chk(self.lib.sh_fold_circle_into_line(self.h, twiddles.h, _ptrs(src.columns), src.log_size,
                                                    _u32s(list(alpha)), _ptrs(dst.columns)),
                  "fold_circle_into_line")
        return dst

    def decompose(self, src):
        dst = SecureColumnByCoords(self.ctx, src.log_size)
        lam = (ctypes.c_uint32 * 4)()
        self._chk(self.lib.sh_fri_decompose(self.h, _ptrs(src.columns), src.log_size, _ptrs(dst.columns), lam),
                  "decompose")
        return dst, np.array(lam, np.uint32)

    # ---- QuotientOps (crates/stwo/src/prover/pcs/quotient_ops.rs:20-53) -----------------------
    def accumulate_numerators(self, cols, sample_batches, log_blowup):
        """sample_batches: list of (point[8], [(col_idx, value[4], random_coeff[4]), ...]).
        Returns a list of (point, partial SecureColumnByCoords, first_linear[4])."""
        col_log = cols[0].log_size
        pts, offs, idx, vals, rc = [], [0], [], [], []
        for point, terms in sample_batches:
            pts.extend(int(v) for v in point)
            for ci, v, r in terms:
                idx.append(ci)
                vals.extend(int(x) for x in v)
                rc.extend(int(x) for x in r)
            offs.append(len(idx))
        outs = [SecureColumnByCoords(self.ctx, col_log - log_blowup) for _ in sample_batches]
        fl = (ctypes.c_uint32 * (4 * max(1, len(sample_batches))))()
        self._chk(self.lib.sh_accumulate_numerators(
            self.h, _ptrs(cols), len(cols), col_log, log_blowup, _u32s(pts), _u32s(offs), _u32s(idx),
            _u32s(vals), _u32s(rc), len(sample_batches), _ptrs([c for o in outs for c in o.columns]), fl),
            "accumulate_numerators")
        fl = np.array(fl, np.uint32)
        return [(np.asarray(p[0], np.uint32), o, fl[4 * i:4 * i + 4])
                for i, (p, o) in enumerate(zip(sample_batches, outs))]

    def compute_quotients_and_combine(self, accumulations, lifting_log_size, log_blowup, twiddles, out=None):
        """accumulations: list of (point[8], partial SecureColumnByCoords, first_linear[4])."""
        out = out or SecureColumnByCoords(self.ctx, lifting_log_size)
        parts = [c for _, p, _ in accumulations for c in p.columns]
        logs = [p.log_size for _, p, _ in accumulations]
        pts = [int(v) for pt, _, _ in accumulations for v in pt]
        fls = [int(v) for _, _, f in accumulations for v in f]
        self._chk(self.lib.sh_compute_quotients_and_combine(
            self.h, twiddles.h, _ptrs(parts), _u32s(logs), _u32s(pts), _u32s(fls), len(accumulations),
            lifting_log_size, log_blowup, _ptrs(out.columns)), "compute_quotients_and_combine")
        return out

    # ---- AccumulationOps (crates/stwo/src/prover/air/accumulation.rs:183-201) -----------------
    def lift_and_accumulate(self, cols):
        if not cols:
            return None
        out = SecureColumnByCoords(self.ctx, cols[-1].log_size)
        self._chk(self.lib.sh_lift_and_accumulate(self.h, _ptrs([c for col in cols for c in col.columns]),
                                                  _u32s(_logs(cols)), len(cols),
                                                  _ptrs(out.columns)), "lift_and_accumulate")
        return out

    def accumulate(self, column, other):
        self._chk(self.lib.sh_accumulate(self.h, _ptrs(column.columns), _ptrs(other.columns), column.log_size),
                  "accumulate")
        return column

    def generate_secure_powers(self, felt, n_powers):
        """AccumulationOps::generate_secure_powers (crates/stwo/src/prover/air/accumulation.rs:
        188-196): [1, felt, felt^2, ...] as an (n_powers, 4) uint32 array."""
        out = np.empty((max(n_powers, 0), 4), np.uint32)
        self.lib.sh_generate_secure_powers(_u32s([int(x) for x in felt]), n_powers, out.ctypes.data)
        return out

    # ---- ComponentProver of the wide-Fibonacci AIR ------------------------------------------
    def wide_fibonacci_composition(self, cols, trace_log, eval_log, powers, accum, log_expansion=0):
        """evaluate_constraint_quotients_on_domain (crates/constraint-framework/src/prover/
        component_prover.rs:118-193) for WideFibonacciEval: adds the constraint quotients of the
        trace columns (committed evaluations) into accum (SecureColumnByCoords of 2^eval_log);
        powers[j] is constraint j's random-coefficient power."""
        flat = np.ascontiguousarray(np.asarray(powers, dtype=np.uint32).reshape(-1))
        self._chk(self.lib.sh_wide_fibonacci_composition(self.h, _ptrs(cols), len(cols), trace_log, eval_log,
                                                         log_expansion, flat.ctypes.data, _ptrs(accum.columns)),
                  "wide_fibonacci_composition")
        return accum

    def wide_fibonacci_trace(self, c0, c1, n_cols, out=None):
        """generate_trace (crates/examples/src/wide_fibonacci/mod.rs:27-48) on the device from the
        two input columns; returns n_cols BaseColumns (out: existing columns to write)."""
        cols = out if out is not None else [BaseColumn(self.ctx, c0.log_size) for _ in range(n_cols)]
        assert len(cols) == n_cols and all(c.log_size == c0.log_size for c in cols)
        self._chk(self.lib.sh_wide_fibonacci_trace(self.h, c0.ptr, c1.ptr, c0.log_size, n_cols, _ptrs(cols)),
                  "wide_fibonacci_trace")
        return cols

    def secure_powers_dev(self, felt_ptr, n_powers, reversed_, out_ptr):
        """generate_secure_powers of a device felt into device memory (reversed: felt^(n-1-k))."""
        self._chk(self.lib.sh_secure_powers_dev(self.h, felt_ptr, n_powers, int(bool(reversed_)), out_ptr),
                  "secure_powers_dev")

    def wide_fibonacci_composition_dev(self, cols, trace_log, eval_log, powers_ptr, accum, log_expansion=0):
        """wide_fibonacci_composition with the constraint powers in device memory."""
        self._chk(self.lib.sh_wide_fibonacci_composition_dev(self.h, _ptrs(cols), len(cols), trace_log, eval_log,
                                                             log_expansion, powers_ptr, _ptrs(accum.columns)),
                  "wide_fibonacci_composition_dev")
        return accum

    def wide_fibonacci_oods_check(self, values_ptr, n_cols, powers_rev_ptr, point_ptr, log_n, out_ptr):
        """prove_ex's closing sanity check on device data (9 words at out_ptr: mismatch flag, the
        composition value, the constraint value)."""
        self._chk(self.lib.sh_wide_fibonacci_oods_check(self.h, values_ptr, n_cols, powers_rev_ptr, point_ptr,
                                                        log_n, out_ptr), "wide_fibonacci_oods_check")

    def eval_at_dpoint(self, coeffs, doublings, point_ptr, values_ptr):
        """compute_samples at one device point: column i at the point doubled doublings[i] times
        into values_ptr[4i..4i+4] (device)."""
        dbl = np.ascontiguousarray(doublings, dtype=np.uint32)
        self._chk(self.lib.sh_eval_at_dpoint(self.h, _ptrs(coeffs), _u32s(_logs(coeffs)), dbl.ctypes.data,
                                             len(coeffs), point_ptr, values_ptr), "eval_at_dpoint")

    # ---- GrindOps (crates/stwo/src/core/proof_of_work.rs:10-14) -------------------------------
    def grind(self, digest: bytes, pow_bits):
        d = (ctypes.c_uint8 * 32)(*digest)
        nonce = ctypes.c_uint64()
        self._chk(self.lib.sh_grind(self.h, d, pow_bits, ctypes.byref(nonce)), "grind")
        return nonce.value


for _name, _fn in list(vars(_Ops).items()):
    if not _name.startswith("__"):
        setattr(HipBackend, _name, _fn)



class ChannelState(ctypes.Structure):
    """sh_channel: Blake2sChannel state (crates/stwo/src/core/channel/blake2s.rs:16-20)."""
    _fields_ = [("digest", ctypes.c_uint8 * 32), ("n_draws", ctypes.c_uint32)]


class Channel:
    """Blake2sChannel over the C++ host implementation."""

    def __init__(self, lib=None):
        self.lib = lib or load()
        self.s = ChannelState()

    @property
    def digest(self):
        return bytes(self.s.digest)

    def mix_u32s(self, words):
        self.lib.sh_channel_mix_u32s(ctypes.byref(self.s), _u32s(list(words)), len(words))

    def mix_u64(self, v):
        self.lib.sh_channel_mix_u64(ctypes.byref(self.s), v)

    def mix_felts(self, felts):
        """felts: QM31 values as 4-int sequences, or a (m, 4) uint32 array."""
        if isinstance(felts, np.ndarray):
            w = np.ascontiguousarray(felts, dtype=np.uint32).reshape(-1)
        else:
            w = np.asarray([int(x) for f in felts for x in f], dtype=np.uint32)
        self.lib.sh_channel_mix_felts(ctypes.byref(self.s), _u32s(w), len(w) // 4)

    def mix_root(self, root: bytes):
        self.lib.sh_channel_mix_root(ctypes.byref(self.s), (ctypes.c_uint8 * 32)(*root))

    def draw_secure_felt(self):
        out = (ctypes.c_uint32 * 4)()
        self.lib.sh_channel_draw_secure_felt(ctypes.byref(self.s), out)
        return np.array(out, np.uint32)

    def draw_random_point(self):
        out = (ctypes.c_uint32 * 8)()
        self.lib.sh_channel_draw_random_point(ctypes.byref(self.s), out)
        return np.array(out, np.uint32)

    def draw_queries(self, log_domain_size, n):
        out = (ctypes.c_uint32 * max(1, n))()
        self.lib.sh_channel_draw_queries(ctypes.byref(self.s), log_domain_size, n, out)
        return list(out)[:n]


class DeviceChannel:
    """Blake2sChannel held in device memory (16 words: the digest, then n_draws), advanced by
    stream-ordered launches on the backend's context: a proof mixes its roots, nonces and sampled
    values and draws its randomness without the host waiting (dchannel.hip)."""

    def __init__(self, backend, channel=None, load=True):
        self.be = backend
        self.buf = DeviceBuffer(backend.ctx, 64)
        if load:
            self.load(channel)

    def load(self, channel=None):
        """Sets the device state to the host channel's (a fresh channel when None)."""
        st = channel.s if channel is not None else ChannelState()
        self.be._chk(self.be.lib.sh_dchannel_load(self.be.h, self.buf.ptr, st.digest, st.n_draws), "dchannel_load")

    @property
    def ptr(self):
        return self.buf.ptr

    def mix_words(self, ptr, n_words):
        """mix_u32s / mix_felts / mix_root over n_words device words at ptr."""
        self.be._chk(self.be.lib.sh_dchannel_mix(self.be.h, self.buf.ptr, ptr, n_words), "dchannel_mix")

    def draw_secure_felts(self, n, out_ptr):
        self.be._chk(self.be.lib.sh_dchannel_draw_secure_felts(self.be.h, self.buf.ptr, n, out_ptr),
                     "dchannel_draw_secure_felts")

    def draw_random_point(self, out_ptr):
        self.be._chk(self.be.lib.sh_dchannel_draw_random_point(self.be.h, self.buf.ptr, out_ptr),
                     "dchannel_draw_random_point")

    def grind(self, pow_bits, ctl_ptr):
        """GrindOps::grind + mix_u64(nonce); the nonce lands at ctl_ptr (2 words, little-endian)."""
        self.be._chk(self.be.lib.sh_grind_dchannel(self.be.h, self.buf.ptr, pow_bits, ctl_ptr), "grind_dchannel")

    def state(self):
        """(digest bytes, n_draws), read back (waits for the stream)."""
        w = self.buf.download(np.uint32, 16)
        return w[:8].tobytes(), int(w[8])

    def free(self):
        self.buf.free()


class FriParams(ctypes.Structure):
    """FriParams (crates/stwo/src/core/fri.rs:31-43)."""
    _fields_ = [("pow_bits", ctypes.c_uint32), ("log_blowup_factor", ctypes.c_uint32),
                ("log_last_layer_degree_bound", ctypes.c_uint32), ("n_queries", ctypes.c_uint32),
                ("fold_step", ctypes.c_uint32)]


def merkle_decommit(backend, layers, positions):
    """Hash witness of MerkleProverLifted::decommit (crates/stwo/src/prover/vcs_lifted/prover.rs:
    97-162); layers root-first as returned by commit_lifted; positions sorted and deduplicated."""
    L = len(layers) - 1
    cap = max(1, len(positions) * max(1, L))
    wit = np.empty(32 * cap, np.uint8)
    nw = ctypes.c_uint64()
    backend._chk(backend.lib.sh_merkle_decommit(backend.h, _ptrs(layers), L, _u32s(positions), len(positions),
                                                wit.ctypes.data, cap, ctypes.byref(nw)), "merkle_decommit")
    raw = wit[:32 * nw.value].tobytes()
    return [raw[32 * i:32 * i + 32] for i in range(nw.value)]


def tree_decommit(backend, cols, layers, positions):
    """CommitmentTreeProver::decommit (crates/stwo/src/prover/pcs/mod.rs:392-398) in one round
    trip: the queried values (n_cols, len(positions)) at the lifted rows of the positions (in the
    given order), and the hash witness of the sorted, deduplicated positions as a (n, 32) uint8
    array.  layers: root-first, as commit_lifted returns them."""
    L = len(layers) - 1
    pos = np.ascontiguousarray(positions, dtype=np.uint32).reshape(-1)
    n, k = len(cols), len(pos)
    values = np.empty((n, k), np.uint32)
    cap = max(1, k * max(1, L))
    wit = np.empty((cap, 32), np.uint8)
    nw = ctypes.c_uint64()
    backend._chk(backend.lib.sh_tree_decommit(backend.h, _ptrs(cols), _u32s(_logs(cols)), n,
                                              _ptrs(layers) if L >= 0 else None, max(L, 0),
                                              pos.ctypes.data if k else None, k,
                                              values.ctypes.data if n and k else None, wit.ctypes.data, cap,
                                              ctypes.byref(nw)), "tree_decommit")
    return values, wit[:nw.value]


def gather_ptrs(backend, ptrs, rows):
    """out[i] = word rows[i] of the device column at address ptrs[i] (numpy arrays), one launch."""
    ptrs = np.ascontiguousarray(ptrs, dtype=np.uint64)
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    out = np.empty(max(1, len(rows)), np.uint32)
    backend._chk(backend.lib.sh_gather(backend.h, ptrs.ctypes.data, rows.ctypes.data, len(rows), out.ctypes.data),
                 "gather")
    return out[:len(rows)]


def gather(backend, cols, rows):
    """out[i] = cols[i].at(rows[i]) in one launch."""
    out = np.empty(max(1, len(rows)), np.uint32)
    backend._chk(backend.lib.sh_gather(backend.h, _ptrs(cols), _u32s(rows), len(rows), out.ctypes.data), "gather")
    return out[:len(rows)].tolist()


def fri_prove(backend, twiddles, channel, params, column, cap=1 << 22, as_array=False):
    """FriProver::commit + decommit (crates/stwo/src/prover/fri.rs:113-284) on a device QM31
    column; returns (proof felts, sorted query positions); the felts as a uint64 array when
    as_array, else as a list."""
    buf = getattr(backend, "_fri_felt_buf", None)
    if buf is None or len(buf) < cap:
        # kept on the backend: a fresh buffer would be paged in again on every proof
        buf = backend._fri_felt_buf = np.empty(cap, np.uint64)
    felts = buf
    nf = ctypes.c_uint64()
    qp = np.empty(max(1, params.n_queries), np.uint32)
    nq = ctypes.c_uint64()
    backend._chk(backend.lib.sh_fri_prove(backend.h, twiddles.h, ctypes.byref(channel.s), ctypes.byref(params),
                                          _ptrs(column.columns), column.log_size, felts.ctypes.data, cap,
                                          ctypes.byref(nf), qp.ctypes.data, ctypes.byref(nq)), "fri_prove")
    out = felts[:nf.value].copy() if as_array else felts[:nf.value].tolist()
    return out, qp[:nq.value].tolist()


class TreeOpening(ctypes.Structure):
    """sh_tree_opening: a committed tree opened at the proof's queries by sh_fri_prove_dch."""
    _fields_ = [("d_cols", ctypes.c_void_p), ("log_sizes", ctypes.c_void_p), ("n_cols", ctypes.c_uint64),
                ("d_layers", ctypes.c_void_p), ("L", ctypes.c_uint32), ("preprocessed", ctypes.c_uint32),
                ("values", ctypes.c_void_p), ("witness", ctypes.c_void_p), ("witness_cap", ctypes.c_uint64),
                ("n_positions", ctypes.c_uint64), ("n_witness", ctypes.c_uint64)]


def _tree_openings(trees, n_queries):
    """sh_tree_opening records (and the host arrays they point at) for trees given as
    (cols, layers, preprocessed)."""
    ops = (TreeOpening * max(1, len(trees)))()
    keep = []
    for o, (cols, layers, pre) in zip(ops, trees):
        L = len(layers) - 1
        n = len(cols)
        values = np.empty((n, n_queries), np.uint32)
        cap_w = max(1, n_queries * max(1, L))
        wit = np.empty((cap_w, 32), np.uint8)
        ptrs, logs = _ptrs(cols), _u32s(_logs(cols))
        lp = _ptrs(layers) if L >= 0 else None
        keep.append((values, wit, ptrs, logs, lp))
        o.d_cols = ctypes.cast(ptrs, ctypes.c_void_p) if n else None
        o.log_sizes = ctypes.cast(logs, ctypes.c_void_p) if n else None
        o.n_cols = n
        o.d_layers = ctypes.cast(lp, ctypes.c_void_p) if lp is not None else None
        o.L = max(L, 0)
        o.preprocessed = int(bool(pre))
        o.values = values.ctypes.data
        o.witness = wit.ctypes.data
        o.witness_cap = cap_w
    return ops, keep


def _openings(ops, keep):
    openings = []
    for o, (values, wit, *_) in zip(ops, keep):
        k = o.n_positions
        v = np.ascontiguousarray(values.reshape(-1)[:o.n_cols * k].reshape(o.n_cols, k))
        openings.append((v, wit[:o.n_witness]))
    return openings


def _felt_buf(backend, cap):
    buf = getattr(backend, "_fri_felt_buf", None)
    if buf is None or len(buf) < cap:
        # kept on the backend: a fresh buffer would be paged in again on every proof
        buf = backend._fri_felt_buf = np.empty(cap, np.uint64)
    return buf


def fri_prove_dch(backend, twiddles, dchannel, channel_out, params, column, extra=None, trees=(), cap=1 << 22):
    """fri_prove on the device channel: channel_out (Channel) receives the final state.  extra:
    (device pointer, n words) read back with the commit phase's download.  trees: (cols, layers,
    preprocessed) of committed trees to open at the queries in the same gather.  Returns (felts as
    a uint64 array, sorted query positions, the extra words, [(values (n_cols, n_pos), witness
    (n, 32) uint8) per tree])."""
    buf = _felt_buf(backend, cap)
    nf = ctypes.c_uint64()
    nqmax = max(1, params.n_queries)
    qp = np.empty(nqmax, np.uint32)
    nq = ctypes.c_uint64()
    e_ptr, e_n = extra if extra is not None else (None, 0)
    h_extra = np.empty(max(1, e_n), np.uint32)
    ops, keep = _tree_openings(trees, nqmax)
    backend._chk(backend.lib.sh_fri_prove_dch(backend.h, twiddles.h, dchannel.ptr, ctypes.byref(channel_out.s),
                                              e_ptr, e_n, h_extra.ctypes.data, ops, len(trees),
                                              ctypes.byref(params), _ptrs(column.columns), column.log_size,
                                              buf.ctypes.data, cap, ctypes.byref(nf), qp.ctypes.data,
                                              ctypes.byref(nq)), "fri_prove_dch")
    return buf[:nf.value].copy(), qp[:nq.value].tolist(), h_extra[:e_n], _openings(ops, keep)


class FriRun:
    """One FRI proof between its commit phase, enqueued by fri_commit_dch (sh_fri_commit_dch;
    recordable into a Graph), and its decommitment."""

    def __init__(self, backend, h, params, n_extra):
        self.be, self.h, self.params, self.n_extra = backend, h, params, n_extra

    def decommit(self, channel_out, trees=(), cap=1 << 22):
        """Waits for the commit phase and finishes the proof (sh_fri_decommit_dch): returns what
        fri_prove_dch returns.  Callable again after a graph replays the commit phase."""
        be = self.be
        buf = _felt_buf(be, cap)
        nf = ctypes.c_uint64()
        nqmax = max(1, self.params.n_queries)
        qp = np.empty(nqmax, np.uint32)
        nq = ctypes.c_uint64()
        h_extra = np.empty(max(1, self.n_extra), np.uint32)
        ops, keep = _tree_openings(trees, nqmax)
        be._chk(be.lib.sh_fri_decommit_dch(be.h, self.h, ctypes.byref(channel_out.s), h_extra.ctypes.data, ops,
                                           len(trees), buf.ctypes.data, cap, ctypes.byref(nf), qp.ctypes.data,
                                           ctypes.byref(nq)), "fri_decommit_dch")
        return buf[:nf.value].copy(), qp[:nq.value].tolist(), h_extra[:self.n_extra], _openings(ops, keep)

    def free(self):
        if self.h and not self.be.ctx._closed:
            self.be._chk(self.be.lib.sh_fri_run_free(self.be.h, self.h), "fri_run_free")
        self.h = None

    def __del__(self):
        # at interpreter exit the context and the library may already be gone: nothing is freed
        if sys.is_finalizing():
            return
        try:
            self.free()
        except Exception:
            pass


def fri_commit_dch(backend, twiddles, dchannel, params, column, extra=None):
    """The commit half of fri_prove_dch (sh_fri_commit_dch): enqueued only; returns a FriRun."""
    e_ptr, e_n = extra if extra is not None else (None, 0)
    h = ctypes.c_void_p()
    backend._chk(backend.lib.sh_fri_commit_dch(backend.h, twiddles.h, dchannel.ptr, e_ptr, e_n,
                                               ctypes.byref(params), _ptrs(column.columns), column.log_size,
                                               ctypes.byref(h)), "fri_commit_dch")
    return FriRun(backend, h.value, params, e_n)


def fri_quotients_dpoint(backend, twiddles, columns, point_ptr, values_ptr, random_coeff_ptr, lifting_log_size,
                         log_blowup, out=None):
    """fri_quotients when every column has one sample at the one device point point_ptr, its
    value at values_ptr[4i..], and the random coefficient in device memory."""
    out = out or SecureColumnByCoords(backend.ctx, lifting_log_size)
    backend._chk(backend.lib.sh_fri_quotients_dpoint(
        backend.h, twiddles.h, _ptrs(columns), _u32s(_logs(columns)), len(columns), point_ptr, values_ptr,
        random_coeff_ptr, lifting_log_size, log_blowup, _ptrs(out.columns)), "fri_quotients_dpoint")
    return out


def fri_quotients(backend, twiddles, columns, samples, random_coeff, lifting_log_size, log_blowup, out=None):
    """compute_fri_quotients (crates/stwo/src/prover/pcs/quotient_ops.rs:86-160).
    columns: committed evaluations (BaseColumn) in tree/column order; samples[i]: list of
    (point[8], value[4]) for column i."""
    offs = np.zeros(len(samples) + 1, np.uint32)
    np.cumsum([len(smp) for smp in samples], out=offs[1:])
    rows = np.array([[*pt, *v] for smp in samples for pt, v in smp], dtype=np.uint64).reshape(-1, 12)
    return fri_quotients_arrays(backend, twiddles, columns, offs, rows[:, :8], rows[:, 8:], random_coeff,
                                lifting_log_size, log_blowup, out)


def fri_quotients_arrays(backend, twiddles, columns, offsets, points, values, random_coeff, lifting_log_size,
                         log_blowup, out=None):
    """fri_quotients with the samples as arrays: column i's samples are rows offsets[i] ..
    offsets[i + 1] - 1 of points ((m, 8) words, x then y) and values ((m, 4) words)."""
    out = out or SecureColumnByCoords(backend.ctx, lifting_log_size)
    offs = np.ascontiguousarray(offsets, dtype=np.uint32)
    assert len(offs) == len(columns) + 1
    pts = np.ascontiguousarray(points, dtype=np.uint32).reshape(-1, 8)
    vals = np.ascontiguousarray(values, dtype=np.uint32).reshape(-1, 4)
    backend._chk(backend.lib.sh_fri_quotients(
        backend.h, twiddles.h, _ptrs(columns), _u32s(_logs(columns)), len(columns),
        offs.ctypes.data, pts.ctypes.data if len(pts) else None, vals.ctypes.data if len(vals) else None,
        _u32s(list(random_coeff)), lifting_log_size, log_blowup, _ptrs(out.columns)), "fri_quotients")
    return out
