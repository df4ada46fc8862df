"""A proof-serving loop on one GPU: continuous batching of wide-Fibonacci prove requests.

Requests arrive over time (a Poisson stream in the benchmark). The scheduler admits waiting
requests shortest first (by trace cells) when the estimated device footprint fits the HBM budget
and a slot is free; every scheduling
tick, the traces of all newly admitted requests are generated on the device and committed
together (one interpolation and one evaluation launch chain per column size across all of
them, then each request's own Merkle tree), and each request then finishes its proof
(composition, OODS, DEEP quotients, FRI, decommitment) on one of several worker contexts, so the
proofs of different requests overlap on the device. The twiddle tree stays resident. A proof is
the same bytes whether it is made alone or in the mix.

Worker contexts are independent (own streams and buffer pools); buffers committed on the
scheduler's context are handed to a worker after the scheduler's stream is synchronised, and
are released only after the worker has synchronised on its last use.
"""
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from .backend import BaseColumn, Channel, Context, HipBackend
from .prover import commit_traces_batched, finish_wide_fibonacci, prove_wide_fibonacci, start_wide_fibonacci


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


class ProofServer:
    """whole_proofs: each worker generates, commits and proves its requests on its own context
    (the proofs of different requests overlap on the device, with no handoff between contexts);
    otherwise the scheduler commits the traces of each admitted batch together on its context
    and the workers finish the proofs."""

    def __init__(self, twiddle_log, fri, n_workers=2, hbm_budget=64 << 30, max_in_flight=None, max_wait=0.05,
                 whole_proofs=False):
        self.fri = dict(fri)
        self.whole_proofs = whole_proofs
        self.ctx = Context(0)
        self.be = HipBackend(self.ctx)
        self.be.load_fft_tuning()  # the committed pass-shape table, when there is one
        self.tw = self.be.precompute_canonic_twiddles(twiddle_log)   # resident for every request
        self.budget = hbm_budget
        self.max_in_flight = max_in_flight or 2 * n_workers
        self.n_workers = n_workers
        self.max_wait = max_wait
        self._work = queue.Queue()
        self._lock = threading.Lock()
        self._reserved = 0
        self._in_flight = 0
        self._errors = []
        self._t0 = None
        self._workers = [threading.Thread(target=self._worker, daemon=True) for _ in range(n_workers)]
        for w in self._workers:
            w.start()

    # ---- worker: the rest of each proof on its own context ---------------------------------
    def _worker(self):
        ctx = Context(0)
        be = HipBackend(ctx)
        be.load_fft_tuning()
        tw = be.precompute_canonic_twiddles(self.tw.root_log_size + 1) if self.whole_proofs else None
        while True:
            item = self._work.get()
            if item is None:
                if tw is not None:
                    tw.free()
                ctx.close()
                return
            req, scheme, channel, cost = item
            try:
                if scheme is None:              # the whole proof on this worker's context
                    c0, c1 = BaseColumn.from_cpu(ctx, req.c0), BaseColumn.from_cpu(ctx, req.c1)
                    trace = be.wide_fibonacci_trace(c0, c1, req.n_cols)
                    req.felts, _, _ = prove_wide_fibonacci(be, tw, trace, req.log_n, self.fri)
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

    def _check_workers(self):
        """A worker thread that died would leave its requests in flight forever."""
        if not all(w.is_alive() for w in self._workers):
            raise RuntimeError(f"a worker thread stopped; errors so far: {self._errors[:3]}")

    def close(self):
        for _ in self._workers:
            self._work.put(None)
        for w in self._workers:
            w.join()
        self.tw.free()
        self.ctx.close()

    # ---- scheduler ---------------------------------------------------------------------------
    def _admit(self, batch):
        """Generate and commit the traces of newly admitted requests together, then hand each to
        the workers."""
        b = self.fri["log_blowup_factor"]
        traces = []
        for r in batch:
            c0, c1 = BaseColumn.from_cpu(self.ctx, r.c0), BaseColumn.from_cpu(self.ctx, r.c1)
            traces.append(self.be.wide_fibonacci_trace(c0, c1, r.n_cols))
        trees = commit_traces_batched(self.be, self.tw, traces, [r.log_n for r in batch], b)
        self.ctx.sync()  # the workers' contexts read these buffers
        for r, t in zip(batch, trees):
            scheme, channel = start_wide_fibonacci(self.be, self.tw, None, r.log_n, self.fri, committed_trace=t)
            self._work.put((r, scheme, channel, device_bytes(r.n_cols, r.log_n, b)))

    def serve(self, requests, realtime=True):
        """Runs the request stream; arrivals are honoured in wall-clock time when realtime.
        Returns the requests with their timestamps and proof felts."""
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting = []  # arrived, not admitted
        self._t0 = time.perf_counter()
        i = 0
        b = self.fri["log_blowup_factor"]
        while i < len(pending) or waiting:
            now = time.perf_counter() - self._t0
            while i < len(pending) and (pending[i].arrival <= now or not realtime):
                waiting.append(pending[i])
                i += 1
            # shortest job first (by trace cells), so a small request does not queue behind a
            # large one; a request waiting longer than max_wait goes first (no starvation)
            waiting.sort(key=lambda r: (now - r.arrival < self.max_wait, r.n_cols << r.log_n, r.arrival))
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
            if batch and self.whole_proofs:
                for r in batch:
                    self._work.put((r, None, None, device_bytes(r.n_cols, r.log_n, b)))
            elif batch:
                self._admit(batch)
            else:
                self._check_workers()
                time.sleep(0.0005)
        while True:
            with self._lock:
                if self._in_flight == 0:
                    break
            self._check_workers()
            time.sleep(0.0005)
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
        r.felts, _, _ = prove_wide_fibonacci(backend, twiddles, trace, r.log_n, fri)
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
