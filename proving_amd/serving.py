"""A proof-serving loop on one GPU: continuous admission of wide-Fibonacci prove requests.

Requests arrive over time (a Poisson stream in the benchmark) and are admitted as worker slots
and the HBM budget allow; the proofs of different requests overlap on the device, and each is
the same bytes whether it is made alone or in the mix.  The default schedule (ProofServer's
"lanes") runs whole proofs on worker contexts, each idle worker taking the oldest waiting request,
so two proofs share the device whenever two are waiting; the "batched" schedule commits the traces of every
admitted batch together (one interpolation and evaluation launch chain per column size across
requests) and finishes the proofs on worker contexts.  Twiddle trees stay resident.

Worker contexts are independent (own stream and buffer pool); in the batched schedule, buffers
committed on the scheduler's context are handed to a worker after the scheduler's stream is
synchronised, and are released only after the worker has synchronised on its last use.
"""
import collections
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from .backend import BaseColumn, Channel, Context, HipBackend
from .prover import (_ProofGraph, commit_traces_batched, finish_wide_fibonacci, prove_wide_fibonacci,
                     start_wide_fibonacci)


@dataclass
class ProveRequest:
    """One prove request: the wide-Fibonacci AIR with n_cols columns of 2^log_n rows whose two
    input columns (c0, c1) the request carries (the rest of the trace is generated on device)."""
    index: int
    n_cols: int
    log_n: int
    c0: np.ndarray
    c1: np.ndarray
    arrival: float = 0.0           # seconds after the stream starts
    admitted: float = None
    done: float = None
    felts: np.ndarray = field(default=None, repr=False)  # CairoSerde felts of the proof


def device_bytes(n_cols, log_n, log_blowup=1):
    """Upper estimate of a request's device footprint: trace / coefficients (4 B per cell), LDE
    (4 B per blown-up cell), the lifted tree, composition (4 QM31 columns on the doubled domain,
    8 committed coordinate columns and their tree), quotients and FRI layers."""
    cells = n_cols << log_n
    L = log_n + log_blowup
    tree = 32 * (2 << L)
    comp = (16 << (log_n + 1)) + 8 * (4 << L) + tree
    fri = 2 * (16 << L) + 2 * tree
    return 4 * cells + 4 * (cells << log_blowup) + tree + comp + fri + (8 << log_n)


class ShapeGraphs:
    """One worker's proof graphs, one per request shape (columns, log rows) it has proved, each with
    its own input and trace buffers: a request of a shape seen before has its two input columns
    uploaded and its trace generated into that shape's buffers and is proved by a replay of the
    shape's captured proof (_ProofGraph); the first request of a shape is proved eagerly, the
    second builds the graph.  Shapes beyond `cap` bytes (device_bytes estimates) are dropped least
    recently used first; a shape larger than the cap is not cached (prove returns None).  The
    proofs are the same bytes as prove_wide_fibonacci's."""

    def __init__(self, backend, twiddles, fri, cap):
        self.be, self.tw, self.fri, self.cap = backend, twiddles, dict(fri), cap
        self.entries = collections.OrderedDict()
        self.bytes = 0

    def _drop(self, key):
        e = self.entries.pop(key)
        self.bytes -= e["bytes"]
        if e["pg"] is not None:
            e["pg"].free()
        for c in e["trace"] + [e["c0"], e["c1"]]:
            c.free()

    def prove(self, req):
        key = (req.n_cols, req.log_n)
        e = self.entries.get(key)
        if e is None:
            cost = device_bytes(req.n_cols, req.log_n, self.fri["log_blowup_factor"])
            if cost > self.cap:
                return None
            while self.entries and self.bytes + cost > self.cap:
                self._drop(next(iter(self.entries)))
            ctx = self.be.ctx
            e = {"c0": BaseColumn(ctx, req.log_n), "c1": BaseColumn(ctx, req.log_n),
                 "trace": [BaseColumn(ctx, req.log_n) for _ in range(req.n_cols)], "pg": None, "seen": False,
                 "bytes": cost}
            self.entries[key] = e
            self.bytes += cost
        self.entries.move_to_end(key)
        e["c0"].upload(req.c0)
        e["c1"].upload(req.c1)
        self.be.wide_fibonacci_trace(e["c0"], e["c1"], req.n_cols, out=e["trace"])
        if e["pg"] is not None:
            return e["pg"].prove(Channel(self.be.lib))[0]
        if not e["seen"]:
            e["seen"] = True
            return prove_wide_fibonacci(self.be, self.tw, e["trace"], req.log_n, self.fri, graph_commit=False)[0]
        pg = _ProofGraph(self.be, self.tw, e["trace"], req.log_n, self.fri, key)
        try:
            felts = pg.build(Channel(self.be.lib))[0]
        except BaseException:
            pg.free()
            raise
        e["pg"] = pg
        return felts

    def free(self):
        while self.entries:
            self._drop(next(iter(self.entries)))


class ProofServer:
    """The request loop.  Two scheduling modes, both admitting by the HBM budget and keeping the
    twiddle tree resident:

    * "lanes" (default): each worker context proves whole requests, and an idle worker takes the
      oldest waiting request of any size (first come, first served over n_workers lanes).  Two
      proofs then share the device whenever two requests wait, and one proof's host phases are
      covered by the other's kernels.  (Measured against a main lane in arrival order plus a lane for small
      requests only, and against big proofs running alone: at 80% and 105% of capacity this had
      the lowest 99th percentile and the same throughput.)  One stream per worker (a context
      creates its helper stream only for multi-size commits), so the process stays within the
      device's hardware queues.
    * "batched": the scheduler commits the traces of each admitted batch together on its own
      context (one interpolation and evaluation launch chain per column size across requests)
      and the workers finish the proofs, shortest job first with aging.

    whole_proofs=True selects "lanes", False "batched" (the names the tests use)."""

    def __init__(self, twiddle_log, fri, n_workers=2, hbm_budget=64 << 30, max_in_flight=None, max_wait=0.05,
                 whole_proofs=True, graph_cache_bytes=24 << 30):
        self.fri = dict(fri)
        # lanes: each worker keeps a proof graph per request shape (ShapeGraphs), within this many
        # bytes of device memory beside the budget (0: every proof eager)
        self.graph_cache_bytes = graph_cache_bytes
        self.mode = "lanes" if whole_proofs else "batched"
        self.whole_proofs = whole_proofs
        self.twiddle_log = twiddle_log
        self.budget = hbm_budget
        self.n_workers = n_workers
        self.max_in_flight = max_in_flight or (n_workers if whole_proofs else 2 * n_workers)
        self.max_wait = max_wait
        self._lock = threading.Lock()
        self._reserved = 0
        self._in_flight = 0
        self._errors = []
        self._t0 = None
        self.ctx = self.be = self.tw = None
        if self.mode == "batched":
            self.ctx = Context(0)
            self.be = HipBackend(self.ctx)
            self.be.load_fft_tuning()  # the committed pass-shape table, when there is one
            self.tw = self.be.precompute_canonic_twiddles(twiddle_log)   # resident for every request
        self._queues = [queue.Queue() for _ in range(n_workers)]
        self._busy = [False] * n_workers
        self._ready = threading.Semaphore(0)
        self._workers = [threading.Thread(target=self._worker, args=(k,), daemon=True) for k in range(n_workers)]
        for w in self._workers:
            w.start()
        for _ in range(n_workers):  # every worker's context and twiddles exist before serving starts
            while not self._ready.acquire(timeout=1.0):
                self._check_workers()

    # ---- worker: whole proofs (lanes) or the rest of each proof (batched) ---------------------
    def _worker(self, k):
        ctx = Context(0)
        be = HipBackend(ctx)
        be.load_fft_tuning()
        tw = be.precompute_canonic_twiddles(self.twiddle_log) if self.mode == "lanes" else None
        graphs = ShapeGraphs(be, tw, self.fri, self.graph_cache_bytes) \
            if self.mode == "lanes" and self.graph_cache_bytes > 0 else None
        ctx.sync()
        self._ready.release()
        while True:
            item = self._queues[k].get()
            if item is None:
                if graphs is not None:
                    graphs.free()
                if tw is not None:
                    tw.free()
                be.release_graphs()
                ctx.close()
                return
            req, scheme, channel, cost = item
            try:
                if scheme is None:              # the whole proof on this worker's context
                    felts = graphs.prove(req) if graphs is not None else None
                    if felts is None:
                        c0, c1 = BaseColumn.from_cpu(ctx, req.c0), BaseColumn.from_cpu(ctx, req.c1)
                        trace = be.wide_fibonacci_trace(c0, c1, req.n_cols)
                        felts = prove_wide_fibonacci(be, tw, trace, req.log_n, self.fri, graph_commit=False)[0]
                    req.felts = felts
                else:
                    scheme.be = be              # continue the proof on this worker's context
                    req.felts, _ = finish_wide_fibonacci(scheme, channel, req.n_cols, req.log_n)
                ctx.sync()
            except Exception as e:  # surfaced by serve()
                self._errors.append((req.index, repr(e)))
            finally:
                if scheme is not None:
                    # kernels still queued on this worker's stream may read the request's buffers
                    # (after a failure the success-path sync above did not run): wait for them
                    # before the buffers go back to the pool they came from
                    try:
                        ctx.sync()
                    except Exception:
                        pass
                    scheme.trees = None         # release this request's buffers (back to the pools)
                req.done = time.perf_counter() - self._t0
                with self._lock:
                    self._reserved -= cost
                    self._in_flight -= 1
                    self._busy[k] = False

    def _check_workers(self):
        """A worker thread that died would leave its requests in flight forever."""
        if not all(w.is_alive() for w in self._workers):
            raise RuntimeError(f"a worker thread stopped; errors so far: {self._errors[:3]}")

    def close(self):
        for q in self._queues:
            q.put(None)
        for w in self._workers:
            w.join()
        if self.tw is not None:
            self.tw.free()
        if self.ctx is not None:
            self.ctx.close()

    # ---- scheduling ----------------------------------------------------------------------------
    def _cells(self, r):
        return r.n_cols << r.log_n

    def _dispatch_lanes(self, waiting, now):
        """Hands the oldest waiting requests to idle workers, as the HBM budget allows.  Returns
        the dispatched."""
        b = self.fri["log_blowup_factor"]
        out = []
        with self._lock:
            for k in range(self.n_workers):
                if self._busy[k] or not waiting:
                    continue
                r = waiting[0]
                cost = device_bytes(r.n_cols, r.log_n, b)
                if self._reserved + cost > self.budget and self._in_flight > 0:
                    break  # the oldest waits for memory; nothing overtakes it
                self._reserved += cost
                self._in_flight += 1
                self._busy[k] = True
                r.admitted = now
                waiting.pop(0)
                out.append((k, r, cost))
        for k, r, cost in out:
            self._queues[k].put((r, None, None, cost))
        return out

    def _admit_batched(self, batch):
        """Generate and commit the traces of newly admitted requests together, then hand each to
        the workers (round robin)."""
        b = self.fri["log_blowup_factor"]
        traces = []
        for r in batch:
            c0, c1 = BaseColumn.from_cpu(self.ctx, r.c0), BaseColumn.from_cpu(self.ctx, r.c1)
            traces.append(self.be.wide_fibonacci_trace(c0, c1, r.n_cols))
        trees = commit_traces_batched(self.be, self.tw, traces, [r.log_n for r in batch], b)
        self.ctx.sync()  # the workers' contexts read these buffers
        for r, t in zip(batch, trees):
            scheme, channel = start_wide_fibonacci(self.be, self.tw, None, r.log_n, self.fri, committed_trace=t)
            k = r.index % self.n_workers
            self._queues[k].put((r, scheme, channel, device_bytes(r.n_cols, r.log_n, b)))

    def serve(self, requests, realtime=True):
        """Runs the request stream; arrivals are honoured in wall-clock time when realtime.
        Returns the requests with their timestamps and proof felts."""
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting = []  # arrived, not admitted, in arrival order
        self._t0 = time.perf_counter()
        i = 0
        b = self.fri["log_blowup_factor"]
        while i < len(pending) or waiting:
            now = time.perf_counter() - self._t0
            while i < len(pending) and (pending[i].arrival <= now or not realtime):
                waiting.append(pending[i])
                i += 1
            if self.mode == "lanes":
                if not self._dispatch_lanes(waiting, now):
                    self._check_workers()
                    time.sleep(0.0002)
                continue
            # batched: shortest job first (by trace cells); a request waiting longer than
            # max_wait goes first (no starvation)
            waiting.sort(key=lambda r: (now - r.arrival < self.max_wait, self._cells(r), r.arrival))
            batch = []
            with self._lock:
                while waiting:
                    r = waiting[0]
                    cost = device_bytes(r.n_cols, r.log_n, b)
                    if self._in_flight >= self.max_in_flight or (self._reserved + cost > self.budget
                                                                   and self._in_flight > 0):
                        break
                    self._reserved += cost
                    self._in_flight += 1
                    r.admitted = now
                    batch.append(waiting.pop(0))
            if batch:
                self._admit_batched(batch)
            else:
                self._check_workers()
                time.sleep(0.0005)
        while True:
            with self._lock:
                if self._in_flight == 0:
                    break
            self._check_workers()
            time.sleep(0.0002)
        if self._errors:
            raise RuntimeError(f"requests failed: {self._errors[:3]}")
        return requests


def prove_sequential(backend, twiddles, requests, fri, realtime=True):
    """The plain baseline: one request at a time on one context, its own trace commit, in
    arrival order (a request waits for the previous one)."""
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda r: r.arrival):
        if realtime:
            while time.perf_counter() - t0 < r.arrival:
                time.sleep(0.0005)
        r.admitted = time.perf_counter() - t0
        c0, c1 = BaseColumn.from_cpu(backend.ctx, r.c0), BaseColumn.from_cpu(backend.ctx, r.c1)
        trace = backend.wide_fibonacci_trace(c0, c1, r.n_cols)
        r.felts, _, _ = prove_wide_fibonacci(backend, twiddles, trace, r.log_n, fri, graph_commit=False)
        backend.ctx.sync()
        r.done = time.perf_counter() - t0
    return requests


def make_requests(n, seed, col_choices=(32, 64, 128, 256), log_range=(14, 20), rate=None):
    """The serving workload: column count uniform in col_choices, log rows uniform in log_range
    (inclusive), inputs from the shared generator with a per-request seed, Poisson arrivals at
    `rate` requests/s (all at time 0 when rate is None)."""
    from .workloads import splitmix64_columns
    rng = np.random.default_rng(seed)
    reqs, t = [], 0.0
    for i in range(n):
        n_cols = int(rng.choice(col_choices))
        log_n = int(rng.integers(log_range[0], log_range[1] + 1))
        c0, c1 = splitmix64_columns(request_seed(seed, i), [log_n, log_n])
        if rate:
            t += float(rng.exponential(1.0 / rate))
        reqs.append(ProveRequest(i, n_cols, log_n, c0, c1, arrival=t if rate else 0.0))
    return reqs


def request_seed(seed, index):
    """Per-request seed: SplitMix64 finaliser of (seed, index)."""
    z = (seed * 0x9E3779B97F4A7C15 + index + 1) & ((1 << 64) - 1)
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & ((1 << 64) - 1)
    return z ^ (z >> 31)
