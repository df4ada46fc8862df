#!/usr/bin/env python3
"""Benchmark: committed trace cells/s of the batched uneven trace commitment on one MI355X.

Workload (one step): 512 M31 columns of mixed sizes (256 @ 2^12, 128 @ 2^16, 96 @ 2^18,
32 @ 2^20; 68,157,440 cells), values from SplitMix64(seed 1), resident in HBM.  A step is
interpolate (iFFT) -> LDE on CanonicCoset(log + 1) -> lifted Blake2s Merkle tree of height 21,
all through sh_commit_trace.  The reference's equivalent is TreeBuilder::extend_evals +
CommitmentTreeProver::new (crates/stwo/src/prover/pcs/mod.rs:325-386).

Output: one JSON line on stdout (progress goes to stderr).  ``roofline`` covers the kernel with
the largest share of the timed region, from per-launch HIP events recorded on the stream the
kernels run on.  ``cpu_baseline`` times a multi-threaded AVX-512 C restatement of the SimdBackend
commit path (and, beside it, one thread of it and the single-threaded scalar oracle) on the same
workload and checks that each produces the same Merkle root.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 39.32  # 256 CU x 64 lanes x 2.4 GHz, 32-bit integer ops
B2_OPS_PER_COMPRESSION = 960  # 10 rounds x 8 G x 12 ops (3-input adds counted once)
VALU_OPS_PER_BUTTERFLY = 11  # M31 butterfly: doubled-twiddle multiply 5, add 3, subtract 3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def lifted_compressions(logs, log_blowup):
    """Blake2s compressions of the lifted leaves with hash-state sharing per size group."""
    groups = {}
    for lg in logs:
        groups[lg + log_blowup] = groups.get(lg + log_blowup, 0) + 1
    sizes = sorted(groups)
    pend = 0
    total = 0
    for gi, size in enumerate(sizes):
        blocks = 0
        for _ in range(groups[size]):
            if pend == 16:
                blocks += 1
                pend = 0
            pend += 1
        if gi + 1 < len(sizes) and pend == 16:  # compressed early: more data follows
            blocks += 1
            pend = 0
        total += blocks << size
    total += 1 << sizes[-1]  # final block of every leaf (before any lift to a taller tree)
    return total


def algorithmic_work(logs, log_blowup):
    """Per-step algorithmic bytes / compressions of each kernel class (per-op accounting)."""
    cells = sum(1 << lg for lg in logs)
    lde = sum(1 << (lg + log_blowup) for lg in logs)
    L = max(logs) + log_blowup
    return {
        # the interpolation (8 B per cell) and the extension (4 B per coefficient + 4 B per
        # extended word), whichever launches run their layers
        "fft_passes": {"bytes": 12 * cells + 4 * lde},
        "merkle_leaves_group": {"bytes": 4 * lde + 32 * (1 << L),
                                "compressions": lifted_compressions(logs, log_blowup)},
        "merkle_next_layers": {"bytes": 96 * ((1 << L) - 1), "compressions": (1 << L) - 1},
    }


def fft_plan(top, tune):
    """The layer ranges [lo, hi) of a transform's passes, as the library plans them (plan_passes in
    proving_amd/csrc/fft.hip) from a tuning entry (block_log, stride_layers, r_block, r_stride)."""
    block, stride = tune[0], tune[1]
    rem = 0
    if top > block:
        rem = (top - block) % stride
        if rem != 0 and rem <= 2 and block + rem <= 13:
            block, rem = block + rem, 0
    bh = min(top, block)
    passes, lo = [(0, bh)], bh
    while lo < top:
        g = min(top - lo, stride)
        if rem != 0 and rem <= 2 and stride < top - lo < 2 * stride:
            g = (top - lo + 1) // 2
        passes.append((lo, lo + g))
        lo += g
    return passes


def fft_butterflies(logs, b, tuning):
    """Butterflies per commit in each FFT launch class: the interpolation's passes, the fused pass
    (the interpolation's last strided pass and the extension's first, when their plans share those
    layers and a fused kernel has that width: 4..8 layers) and the extension's passes."""
    out = {"fft_ifft_pass": 0, "fft_ie_pass": 0, "fft_rfft_pass": 0}
    for m in logs:
        pi = fft_plan(m, tuning[m]) if m > 2 else []
        pe = fft_plan(m, tuning[m + b]) if m > 2 else []
        fused = (len(pi) >= 2 and len(pe) >= 2 and pi[-1] == pe[-1] and pi[-1][1] == m
                 and 4 <= pi[-1][1] - pi[-1][0] <= 8)
        for k, (lo, hi) in enumerate(pi):
            cls = "fft_ie_pass" if fused and k == len(pi) - 1 else "fft_ifft_pass"
            out[cls] += (hi - lo) << (m - 1)
        for k, (lo, hi) in enumerate(pe):
            cls = "fft_ie_pass" if fused and k == len(pe) - 1 else "fft_rfft_pass"
            out[cls] += (hi - lo) << (m + b - 1)
    return out


def profiled_traffic(kernel, tuning):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE/WRITE_SIZE summary of the
    uneven commit (two rocprofv3 --pmc runs, calibrated by a copy of known size), and whether
    those runs used the same FFT pass shapes as this run."""
    import re
    path = os.path.join(ROOT, "profiles", "commit_uneven512_hbm_traffic.md")
    if not os.path.exists(path):
        return None, None
    text = open(path).read()
    m = re.search(r"traffic_bytes_per_launch %s (\d+)" % re.escape(kernel), text)
    t = re.search(r"--fft-tuning '([^']+)'", text)
    if not m:
        return None, None
    same = t is not None and json.loads(t.group(1)) == {k: list(v) for k, v in tuning.items()}
    return int(m.group(1)), same


def _cpu_threads():
    """Host threads this process may use: its CPU affinity, capped by OMP_NUM_THREADS when set
    (the share of the host that goes with one GPU)."""
    n = len(os.sched_getaffinity(0))
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def cpu_baseline(work, expected_root):
    """The same commit on the host.  Headline: the multi-threaded AVX-512 C restatement of the
    SimdBackend commit path (oracle/simd_port.c) on every thread this process may use, fastest of
    3 commits; beside it the same on one thread and the single-threaded scalar oracle
    (restatement of CpuBackend).  Each must produce the GPU's root."""
    import oracle as O
    from proving_amd.workloads import splitmix64_columns
    logs, b = work["logs"], work["log_blowup"]
    host = splitmix64_columns(work["seed"], logs)
    L = max(logs) + b
    cells = sum(1 << lg for lg in logs)
    tw, itw = O.canonic_twiddles(L)
    t0 = time.perf_counter()
    ldes = [O.evaluate(tw, O.interpolate(itw, c, lg), lg + b) for c, lg in zip(host, logs)]
    layers = O.merkle_commit_lifted(ldes, L)
    dt = time.perf_counter() - t0
    del ldes
    scalar = {"value": cells / dt, "unit": "committed trace cells/s", "cores": 1,
              "sample": f"the full workload, one commit ({cells} cells) by the single-threaded C "
                        f"oracle (restatement of CpuBackend), {dt:.1f} s",
              "root_matches_gpu": layers[0][0].tobytes() == expected_root}
    if not O.simd_port_available():
        return dict(scalar, kind="port", simd_port=None,
                    simd_port_missing="this host has no AVX-512F; the scalar oracle is the baseline")
    nt = _cpu_threads()
    root_n, t_n = O.simd_port_commit(host, logs, b, L, nt, reps=3)
    root_1, t_1 = O.simd_port_commit(host, logs, b, L, 1, reps=1)
    ms = lambda t: {"fft": round(t[0] * 1e3, 2), "leaves": round(t[1] * 1e3, 2),
                    "inner_layers": round(t[2] * 1e3, 2)}
    return {"value": cells / sum(t_n), "unit": "committed trace cells/s", "cores": nt, "kind": "port",
            "sample": f"the full workload ({cells} cells), fastest of 3 commits by the C restatement "
                      f"of the SimdBackend commit path (oracle/simd_port.c: 16 M31 lanes per AVX-512 "
                      f"register, {nt} threads), {sum(t_n) * 1e3:.1f} ms per commit",
            "ms_by_phase": ms(t_n), "root_matches_gpu": root_n == expected_root,
            "single_thread": {"value": cells / sum(t_1), "cores": 1, "ms_by_phase": ms(t_1),
                              "root_matches_gpu": root_1 == expected_root},
            "scalar_oracle": scalar}


def _tune(args, be, logs, log_blowup, tree):
    """FFT pass shapes: the committed table (proving_amd/fft_tuning_gfx950.json, written by
    python -m proving_amd.autotune) unless --autotune asks for a live autotune or there is no
    table; --no-autotune keeps the built-in shapes.  Returns whether shapes were set."""
    if args.fft_tuning:
        for m, t in json.loads(args.fft_tuning).items():
            be.set_fft_tuning(int(m), *[int(x) for x in t])
        log("[bench] FFT pass shapes set from --fft-tuning")
        return True
    if args.no_autotune:
        return False
    if not args.autotune and be.load_fft_tuning():
        log("[bench] FFT pass shapes from proving_amd/fft_tuning_gfx950.json")
        return True
    t0 = time.perf_counter()
    be.autotune_fft(logs, log_blowup, tree)
    log(f"[bench] autotuned FFT pass shapes in {time.perf_counter() - t0:.1f} s")
    return True


def _timed(ctx, step_fn, steps):
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step_fn()
    ctx.sync()
    return time.perf_counter() - t0


def _kernel_table(prof, steps):
    return {k: {"launches_per_step": v["n"] / steps, "ms_per_step": v["ms"] / steps,
                "avg_launch_us": 1e3 * v["ms"] / v["n"]} for k, v in prof.items()}


def bench_fri_deep(args):
    """The FRI + DEEP-quotient phase of prove_values (crates/stwo/src/prover/pcs/mod.rs:224-258) on
    200 columns x 2^21 (LDE 2^22): one step = OOD sampling of every column at 3 points, mixing the
    samples, drawing the random coefficient, compute_fri_quotients, and FRI commit + decommit with
    32 queries; the channel starts from the same state each step (the phase is deterministic)."""
    from proving_amd.backend import (BaseColumn, Channel, ColumnSet, Context, FriParams, HashLayerSet, HipBackend,
                                     fri_prove, fri_quotients_arrays)
    from proving_amd.workloads import FRI_DEEP, splitmix64_columns
    w = FRI_DEEP
    n, N, b = w["log_size"], w["n_cols"], w["fri"]["log_blowup_factor"]
    L = n + b
    ctx = Context(0)
    be = HipBackend(ctx)
    tree = be.precompute_canonic_twiddles(L)
    _tune(args, be, [n] * N, b, tree)
    log(f"[bench] fri_deep: {N} columns x 2^{n}, LDE 2^{L}; generating inputs")
    host = splitmix64_columns(w["seed"], [n] * N)
    evals = [BaseColumn.from_cpu(ctx, c) for c in host]
    del host
    coeffs = ColumnSet(ctx, [n] * N)
    ldes = ColumnSet(ctx, [L] * N)
    layers = HashLayerSet(ctx, L)
    be.commit_trace(evals, b, tree, coeffs, ldes, layers)  # setup: interpolate + LDE (not timed)
    for e in evals:
        e.free()
    ch0 = Channel(ctx.lib)
    ch0.mix_u64(w["seed"])
    points = [[int(x) for x in ch0.draw_random_point()] for _ in range(w["n_points"])]
    params = FriParams(*[w["fri"][k] for k in ("pow_bits", "log_blowup_factor", "log_last_layer_degree_bound",
                                               "n_queries", "fold_step")])
    state = bytes(ch0.s.digest), ch0.s.n_draws
    result = {}
    k = len(points)
    offsets = np.arange(0, k * N + 1, k, dtype=np.uint32)          # k samples per column
    sample_points = np.tile(np.asarray(points, np.uint32), (N, 1))  # column-major, point order

    def step():
        ch = Channel(ctx.lib)
        ch.s.digest[:] = list(state[0])
        ch.s.n_draws = state[1]
        vals = be.eval_at_points(coeffs, points, as_array=True).reshape(-1, 4)
        ch.mix_felts(vals)
        rc = [int(x) for x in ch.draw_secure_felt()]
        quot = fri_quotients_arrays(be, tree, ldes, offsets, sample_points, vals, rc, L, b)
        felts, queries = fri_prove(be, tree, ch, params, quot, as_array=True)
        quot.free()
        result["felts"], result["queries"], result["digest"] = felts, queries, ch.digest

    step()
    first = dict(result)
    for _ in range(args.warmup):
        step()
    sec = _timed(ctx, step, args.steps)  # timed without per-launch event spans
    assert np.array_equal(result["felts"], first["felts"]), "the phase is not deterministic"
    ms = sec * 1e3 / args.steps
    prof_steps = max(1, args.steps // 4)  # a separate pass with spans, for the kernel table
    ctx.profile(True)
    _timed(ctx, step, prof_steps)
    prof = ctx.profile_report()
    ctx.profile(False)
    # roofline: the fused numerator pass reads the 2^(L-b)-row subdomain of every column once
    num_bytes = 4 * N * (1 << (L - b)) + 16 * w["n_points"] * (1 << (L - b))
    k = prof.get("quotients_accumulate_numerators")
    roof = None
    if k:
        avg = k["ms"] / k["n"] / 1e3
        ach = num_bytes / (k["n"] / prof_steps) / avg / 1e9
        roof = {"kernel": "quotients_accumulate_numerators", "bound": "hbm", "achieved": round(ach, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_step": num_bytes, "avg_launch_us": round(avg * 1e6, 2)}
    out = {"metric": "FRI+DEEP phases/sec (200 x 2^21 columns, 3 OOD points, 32 queries)",
           "value": round(1e3 / ms, 3), "unit": "phases/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u32 (M31/QM31)", "data": "synthetic",
           "config": {"workload": w["name"], "columns": N, "log_size": n, "lde_log_size": L,
                      "points": w["n_points"], "fri": w["fri"]},
           "roofline": roof, "kernels": _kernel_table(prof, prof_steps),
           "proof_felts": len(first["felts"]), "queries": first["queries"][:8],
           "last_layer_poly_felts": [int(x) for x in first["felts"][-(4 * 8 + 2):-2]]}
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_prove(args):
    """A full proof of the wide-Fibonacci AIR (256 columns x 2^20, FriParams::new(26, 0, 1, 70, 1))
    on the device: one step = prove() from a resident trace (trace commit, composition,
    composition commit, OODS grind and sampling, DEEP quotients, FRI with the 26-bit grind,
    decommitment, proof felts)."""
    from proving_amd.backend import BaseColumn, Channel, Context, HipBackend
    from proving_amd.prover import ProofPipeline, prove_wide_fibonacci
    from proving_amd.workloads import WIDE_FIB, wide_fib_trace
    w = WIDE_FIB
    n, N = w["log_size"], w["n_cols"]
    ctx = Context(0)
    be = HipBackend(ctx)
    tree = be.precompute_canonic_twiddles(n + w["fri"]["log_blowup_factor"])
    log(f"[bench] prove: wide Fibonacci {N} x 2^{n}; generating the trace")
    trace = [BaseColumn.from_cpu(ctx, c) for c in wide_fib_trace(N, n, w["seed"])]
    _tune(args, be, [n] * N, w["fri"]["log_blowup_factor"], tree)
    res = {}

    def step(graph=True):
        # graph: from the second proof of these buffers on, a replay of the captured graph of the
        # proof's device work; else every launch issued eagerly
        felts, proof, _ = prove_wide_fibonacci(be, tree, trace, n, w["fri"], graph_commit=graph)
        res["felts"] = felts

    step()
    first = res["felts"].copy()
    for _ in range(args.warmup):
        step()
    steps = max(1, args.steps // 4)
    sec = _timed(ctx, step, steps)  # timed without per-launch event spans
    assert np.array_equal(res["felts"], first), "proofs differ between runs"
    ms = sec * 1e3 / steps
    # the device span of one replay (a hipEvent pair around the channel load and the launch, no
    # profiler): over ms, the GPU-busy share of proving one at a time
    pg, dev = be._proof_graph, []
    for _ in range(steps):
        ch = Channel(be.lib)
        ctx.timer_start()
        pg.start(ch)
        dev.append(ctx.timer_stop())
        res["felts"] = pg.finish(ch)[0]
    assert np.array_equal(res["felts"], first), "proofs differ between runs"
    dev_ms = float(np.median(dev))
    step(False)
    eager_ms = _timed(ctx, lambda: step(False), steps) * 1e3 / steps
    assert np.array_equal(res["felts"], first), "eager and graph proofs differ"
    ctx.profile(True)  # a separate eager pass with spans, for the kernel table
    _timed(ctx, lambda: step(False), steps)
    prof = ctx.profile_report()
    ctx.profile(False)
    # the headline: three proofs in flight, on three contexts each replaying its own proof graph
    # (the host finishes one proof while the others' device work runs; measured 7.9 ms per proof
    # with two in flight, 7.6 with three, 7.5 with four, which would take every hardware queue)
    in_flight = 3
    extra = [Context(0) for _ in range(in_flight - 1)]
    bes = [be]
    for c in extra:
        b2 = HipBackend(c)
        for m in range(32):   # the same pass shapes as the first context's
            b2.set_fft_tuning(m, *be.fft_tuning(m))
        bes.append(b2)
    pipe = ProofPipeline(bes, trace, n, w["fri"])
    for f, _ in pipe.prove([None] * in_flight):   # warm-up
        assert np.array_equal(f, first), "pipelined and single proofs differ"
    n_pipe = max(4, args.steps // 2)
    t0 = time.perf_counter()
    proofs = pipe.prove([None] * n_pipe)
    sec_pipe = time.perf_counter() - t0
    for f, _ in proofs:
        assert np.array_equal(f, first), "pipelined and single proofs differ"
    pipe_ms = sec_pipe * 1e3 / n_pipe
    out = {"metric": "wide-Fibonacci proofs/sec (256 x 2^20, 70 queries, 26-bit grind)",
           "value": round(1e3 / pipe_ms, 3), "unit": "proofs/s", "n_gpus": 1, "steps": n_pipe,
           "warmup": args.warmup, "ms_per_step": round(pipe_ms, 2), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u32 (M31/QM31)", "data": "synthetic",
           "config": {"workload": w["name"], "columns": N, "log_size": n, "fri": w["fri"], "in_flight": in_flight},
           "one_at_a_time_ms_per_step": round(ms, 2), "replay_device_ms": round(dev_ms, 3),
           "one_at_a_time_gpu_busy": round(dev_ms / ms, 3), "eager_ms_per_step": round(eager_ms, 2),
           "kernels": _kernel_table(prof, steps), "proof_felts": len(first)}
    if not args.no_cpu_baseline:
        import oracle as O
        from oracle import protocol as PR
        log("[bench] verifying the proof with the oracle verifier ...")
        t0 = time.perf_counter()
        out["oracle_verifies"] = bool(PR.wide_fib_verify(first.tolist(), N, n, w["fri"]))
        out["oracle_verify_s"] = round(time.perf_counter() - t0, 2)
    print(json.dumps(out), flush=True)
    pipe.free()
    be.release_graphs()
    for c in extra:
        c.close()
    ctx.close()


def bench_serve(args):
    """Proof serving on one GPU: requests with 32/64/128/256 columns of 2^14..2^20 rows (seed 4),
    Poisson arrivals at 80% and 105% of the capacity measured by back-to-back plain proving; the
    server (admission by HBM budget, overlapping proofs on worker contexts) against the plain
    sequential drop-in on the same arrivals."""
    import numpy as np
    from proving_amd.backend import Context, HipBackend
    from proving_amd.serving import ProofServer, make_requests, prove_sequential
    fri = {"pow_bits": 10, "log_blowup_factor": 1, "log_last_layer_degree_bound": 0, "n_queries": 70,
           "fold_step": 1}
    n = args.requests
    ctx = Context(0)
    be = HipBackend(ctx)
    be.load_fft_tuning()  # the same pass shapes as the server's contexts
    tw = be.precompute_canonic_twiddles(21)
    log(f"[bench] serve: {n} requests; warm-up and capacity (back-to-back plain proving)")
    # warm-up: the whole request set once (every size's buffers pooled, clocks at sustained load),
    # so the capacity that sets the loads is plain proving's sustained rate
    prove_sequential(be, tw, make_requests(n, 4), fri, realtime=False)
    reqs = make_requests(n, 4)
    t0 = time.perf_counter()
    prove_sequential(be, tw, reqs, fri, realtime=False)
    capacity = n / (time.perf_counter() - t0)
    # one long-lived server (as a service runs): its capacity run, every request queued at time 0,
    # also warms every worker's buffer pool with every request size before the timed loads
    server = ProofServer(21, fri, n_workers=args.serve_workers, hbm_budget=128 << 30,
                         whole_proofs=not args.serve_batched)
    cap_reqs = make_requests(n, 4)
    t0 = time.perf_counter()
    server.serve(cap_reqs, realtime=False)
    server_capacity = n / (time.perf_counter() - t0)
    log(f"[bench] capacity: plain {capacity:.1f}, server {server_capacity:.1f} proofs/s")

    def stats(rs):
        lat = np.array([r.done - r.arrival for r in rs])
        span = max(r.done for r in rs) - min(r.arrival for r in rs)
        return {"proofs_per_s": round(len(rs) / span, 2), "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 1),
                "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 1)}

    loads = {}
    same = True
    for load in (0.8, 1.05):
        rate = load * capacity
        seq = make_requests(n, 4, rate=rate)
        prove_sequential(be, tw, seq, fri, realtime=True)
        srv_reqs = make_requests(n, 4, rate=rate)
        server.serve(srv_reqs, realtime=True)
        same = same and all(np.array_equal(x.felts, y.felts) for x, y in zip(seq, srv_reqs))
        loads[str(load)] = {"arrival_rate": round(rate, 2), "server": stats(srv_reqs), "plain_sequential": stats(seq)}
        log(f"[bench] load {load}: {loads[str(load)]}")
    server.close()
    top = loads["1.05"]
    out = {"metric": "proof-serving throughput (wide-Fibonacci requests, Poisson arrivals at 105% of plain capacity)",
           "value": top["server"]["proofs_per_s"], "unit": "proofs/s", "n_gpus": 1, "steps": n, "warmup": 4,
           "ms_per_step": round(1e3 / top["server"]["proofs_per_s"], 2), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u32 (M31/QM31)", "data": "synthetic",
           "config": {"workload": "serve_wide_fib", "requests": n, "columns": [32, 64, 128, 256],
                      "log_rows": [14, 20], "seed": 4, "fri": fri, "workers": args.serve_workers,
                      "scheduling": ("shortest job first with 50 ms aging, HBM-budget admission, batched trace commits"
                                     if args.serve_batched else
                                     "HBM-budget admission; whole proofs per worker context, each idle worker taking "
                                     "the oldest waiting request and replaying a captured proof graph per request "
                                     "shape")},
           "loads": loads, "capacity_plain_back_to_back": round(capacity, 2),
           "server_p99_le_plain": {k: v["server"]["p99_ms"] <= v["plain_sequential"]["p99_ms"] for k, v in loads.items()},
           "server_throughput_ge_plain": {k: v["server"]["proofs_per_s"] >= v["plain_sequential"]["proofs_per_s"]
                                          for k, v in loads.items()},
           "capacity_server_all_queued": round(server_capacity, 2), "proofs_identical_to_plain": same}
    print(json.dumps(out), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    # untimed warmup: a GPU idle before the timed region starts at lower clocks, and a 20-step
    # region right after idling measures 8-10% slower than the same steps under sustained load
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", choices=["uneven", "plain", "fri_deep", "prove", "serve"], default="uneven",
                    help="uneven (default, the headline): the 512-column batched commit; plain: 64 x 2^14; "
                         "fri_deep: the FRI + DEEP-quotient phase on 200 x 2^21; prove: a full proof of the "
                         "wide-Fibonacci AIR, 256 x 2^20; serve: a Poisson stream of wide-Fibonacci prove "
                         "requests through the continuous-batching server")
    ap.add_argument("--requests", type=int, default=400,
                    help="serve: number of requests (400: the 99th percentile rests on 4 requests)")
    ap.add_argument("--serve-workers", type=int, default=2, help="serve: worker contexts of the server")
    ap.add_argument("--serve-batched", action="store_true",
                    help="serve: commit the traces of each admitted batch together on a scheduler context "
                         "(shortest job first) instead of the default lanes (whole proofs per worker, each "
                         "idle worker taking the oldest waiting request)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in pass shapes")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the FFT pass shapes now instead of loading proving_amd/fft_tuning_gfx950.json")
    ap.add_argument("--fft-tuning", default=None,
                    help="JSON {log_size: [block_log, stride_layers, tile_log, max_stage]} to use instead "
                         "of autotuning (the 'fft_tuning' field of an earlier run), so that a profiled run "
                         "launches exactly the measured kernels")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain per-column, unfused-FFT and host-to-device variants (their "
                         "FFT launches would mix into a profiler's per-kernel averages)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device-to-device once (counter calibration run) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build runs on a "
              "single MI355X and has no multi-GPU path.", file=sys.stderr)
        sys.exit(2)

    from proving_amd.backend import BaseColumn, ColumnSet, Context, HashLayer, HashLayerSet, HipBackend
    from proving_amd.workloads import PLAIN_COMMIT, UNEVEN_COMMIT, cells as count_cells, splitmix64_columns

    if args.copy_calibration:
        ctx = Context(0)
        n = 256 << 20
        src, dst = ctx.alloc(n), ctx.alloc(n)
        ctx.check(ctx.lib.sh_memset(ctx.h, src.ptr, 1, n), "memset")
        ctx.sync()
        ctx.check(ctx.lib.sh_copy(ctx.h, dst.ptr, src.ptr, n), "copy")
        ctx.sync()
        print(json.dumps({"copy_calibration_bytes": n}), flush=True)
        ctx.close()
        return

    if args.workload == "fri_deep":
        return bench_fri_deep(args)
    if args.workload == "prove":
        return bench_prove(args)
    if args.workload == "serve":
        return bench_serve(args)
    work = UNEVEN_COMMIT if args.workload == "uneven" else PLAIN_COMMIT
    logs, b, seed = work["logs"], work["log_blowup"], work["seed"]
    L = max(logs) + b
    n_cells = count_cells(logs)
    log(f"[bench] workload {work['name']}: {len(logs)} columns, {n_cells} cells, L={L}")

    ctx = Context(0)
    be = HipBackend(ctx)
    tree = be.precompute_canonic_twiddles(L)
    host = splitmix64_columns(seed, logs)
    evals = [BaseColumn.from_cpu(ctx, c) for c in host]
    coeffs = ColumnSet(ctx, logs)                      # one pooled allocation per output kind
    ldes = ColumnSet(ctx, [lg + b for lg in logs])
    layers = HashLayerSet(ctx, L)
    log("[bench] inputs resident in HBM")
    if args.fft_tuning:
        for m, t in json.loads(args.fft_tuning).items():
            be.set_fft_tuning(int(m), *[int(x) for x in t])
        log(f"[bench] FFT pass shapes set from --fft-tuning")
    elif _tune(args, be, logs, b, tree):
        tuned = {m: be.fft_tuning(m) for m in sorted(set(logs) | {lg + b for lg in logs})}
        log(f"[bench] FFT pass shapes: {tuned}")

    _, _, _, root0 = be.commit_trace(evals, b, tree, coeffs, ldes, layers)
    for _ in range(args.warmup):
        be.commit_trace(evals, b, tree, coeffs, ldes, layers, want_root=False)
    ctx.sync()
    log(f"[bench] warmup done ({args.warmup} steps)")

    def timed(step_fn, steps):
        ctx.sync()
        t0 = time.perf_counter()
        ctx.timer_start()
        for _ in range(steps):
            step_fn()
        gpu_ms = ctx.timer_stop()
        ctx.sync()
        wall = time.perf_counter() - t0
        return max(wall, gpu_ms / 1e3)

    # (1) eager batched commit, with per-launch HIP-event spans for the kernel table / roofline
    ctx.profile(True)
    sec_eager = timed(lambda: be.commit_trace(evals, b, tree, coeffs, ldes, layers, want_root=False),
                      args.steps)
    prof = ctx.profile_report()
    ctx.profile(False)
    assert layers[0].to_cpu()[0].tobytes() == root0, "commit is not deterministic across steps"
    log(f"[bench] eager batched: {sec_eager * 1e3 / args.steps:.3f} ms/step")

    # (2) the same commit captured as one HIP graph: the headline number
    graph = be.commit_trace_graph(evals, b, tree, coeffs, ldes, layers)
    for _ in range(args.warmup):
        graph.launch()
    sec = timed(graph.launch, args.steps)
    root = graph.root()
    assert root == root0, "graph replay differs from the eager commit"
    log(f"[bench] graph batched: {sec * 1e3 / args.steps:.3f} ms/step")
    ms_per_step = sec * 1e3 / args.steps
    value = n_cells * args.steps / sec

    variants = {"graph_batched_ms_per_step": round(ms_per_step, 4),
                "eager_batched_ms_per_step": round(sec_eager * 1e3 / args.steps, 4)}

    # (2a) the same graph with the interpolation and the extension as separate transforms (no fused
    # pass): what the fusion is worth on this box
    if not args.no_plain:
        be.set_fft_fusion(0)
        ugraph = be.commit_trace_graph(evals, b, tree, coeffs, ldes, layers)
        for _ in range(args.warmup):
            ugraph.launch()
        sec_u = timed(ugraph.launch, args.steps)
        assert ugraph.root() == root0, "unfused commit differs"
        ugraph.free()
        be.set_fft_fusion(1)
        variants["unfused_graph_ms_per_step"] = round(sec_u * 1e3 / args.steps, 4)
        log(f"[bench] graph batched, interpolation and extension unfused: {sec_u * 1e3 / args.steps:.3f} ms/step")

    # (2b) the same pipeline with the mixed-degree tree that injects the smaller columns at their
    # matching layer (MerkleOps::commit_on_layer), also captured as one graph
    mlayers = HashLayerSet(ctx, L)
    _, _, _, mroot0 = be.commit_trace(evals, b, tree, coeffs, ldes, mlayers, mixed=True)
    mgraph = be.commit_trace_graph(evals, b, tree, coeffs, ldes, mlayers, mixed=True)
    for _ in range(args.warmup):
        mgraph.launch()
    sec_m = timed(mgraph.launch, args.steps)
    assert mgraph.root() == mroot0, "mixed-degree graph replay differs from the eager mixed commit"
    mgraph.free()
    variants["mixed_graph_ms_per_step"] = round(sec_m * 1e3 / args.steps, 4)
    variants["mixed_root"] = mroot0.hex()
    log(f"[bench] graph batched, mixed-degree tree: {sec_m * 1e3 / args.steps:.3f} ms/step")
    if not args.no_plain:
        # (2c) two commits of the same columns in flight, each on a context of its own (its own
        # graph and output buffers): how much of the device one commit leaves idle
        ctx2 = Context(0)
        be2 = HipBackend(ctx2)
        for m in range(32):
            be2.set_fft_tuning(m, *be.fft_tuning(m))
        tree2 = be2.precompute_canonic_twiddles(L)
        out2 = (ColumnSet(ctx2, logs), ColumnSet(ctx2, [lg + b for lg in logs]), HashLayerSet(ctx2, L))
        graph2 = be2.commit_trace_graph(evals, b, tree2, *out2)
        for _ in range(args.warmup):
            graph.launch()
            graph2.launch()
        ctx.sync()
        ctx2.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            graph.launch()
            graph2.launch()
        ctx.sync()
        ctx2.sync()
        sec_2 = time.perf_counter() - t0
        assert graph2.root() == root0 and graph.root() == root0, "concurrent commits differ"
        graph2.free()
        ctx2.close()
        variants["two_in_flight_ms_per_commit"] = round(sec_2 * 1e3 / (2 * args.steps), 4)
        log(f"[bench] two commits in flight (two contexts): {sec_2 * 1e3 / (2 * args.steps):.3f} ms per commit")
        # (3) host-inclusive: the columns start in (pageable) host memory and cross PCIe first
        h2d_steps = max(1, args.steps // 4)
        def h2d_step():
            for e, h in zip(evals, host):
                e.upload(h)
            graph.launch()
        sec_h2d = timed(h2d_step, h2d_steps)
        assert graph.root() == root0
        variants["h2d_inclusive_ms_per_step"] = round(sec_h2d * 1e3 / h2d_steps, 4)
        log(f"[bench] host-to-device inclusive: {sec_h2d * 1e3 / h2d_steps:.3f} ms/step")
    graph.free()

    if not args.no_plain:
        # (4) plain per-column drop-in: one interpolate and one evaluate call per column (as the
        # reference's PolyOps calls are made column by column), then one lifted Merkle commit
        plain_steps = max(1, args.steps // 4)
        for e, c in zip(evals, coeffs):
            c.copy_from(e)
        def plain_step():
            for c in coeffs:
                be.interpolate(c, tree)
            for c, l in zip(coeffs, ldes):
                be.evaluate(c, c.log_size + b, tree, l)
            be.commit_lifted(ldes, L, layers=layers, want_root=False)
        plain_step()
        assert layers[0].to_cpu()[0].tobytes() == root0, "plain per-column path differs from the batched commit"
        sec_plain = timed(plain_step, plain_steps)
        log(f"[bench] plain per-column drop-in: {sec_plain * 1e3 / plain_steps:.3f} ms/step")
        variants["plain_per_column_ms_per_step"] = round(sec_plain * 1e3 / plain_steps, 4)
        variants["graph_vs_plain"] = round(sec_plain / plain_steps / (sec / args.steps), 2)

    # (5) the reference's own Merkle benchmark shape (crates/stwo/benches/merkle.rs:12-26): one
    # commit_on_layer over 256 columns of 2^16 zeros with no children
    zcols = [BaseColumn.from_cpu(ctx, np.zeros(1 << 16, np.uint32)) for _ in range(256)]
    zout = HashLayer(ctx, 16)
    be.commit_on_layer(16, None, zcols, out=zout)
    for _ in range(args.warmup):
        be.commit_on_layer(16, None, zcols, out=zout)
    sec_z = timed(lambda: be.commit_on_layer(16, None, zcols, out=zout), args.steps)
    variants["commit_on_layer_256x2p16_ms"] = round(sec_z * 1e3 / args.steps, 4)
    variants["commit_on_layer_256x2p16_elements_per_s"] = round((256 << 16) * args.steps / sec_z, 1)
    log(f"[bench] commit_on_layer 256 x 2^16 (the reference's Merkle bench shape): "
        f"{sec_z * 1e3 / args.steps:.3f} ms")
    del zcols, zout

    work_per = algorithmic_work(logs, b)
    kernels = {}
    for k, v in prof.items():
        entry = {"launches_per_step": v["n"] / args.steps, "ms_per_step": v["ms"] / args.steps,
                 "avg_launch_us": 1e3 * v["ms"] / v["n"]}
        if k in work_per:
            w = work_per[k]
            entry["gbs"] = w["bytes"] / (v["ms"] / args.steps / 1e3) / 1e9
            if "compressions" in w:
                entry["valu_tops"] = (w["compressions"] * B2_OPS_PER_COMPRESSION
                                      / (v["ms"] / args.steps / 1e3) / 1e12)
        kernels[k] = entry
    # roofline of the dominant kernel class: the FFT passes (the interpolation's, the extension's
    # and the fused pass between them, sh_set_fft_fusion) taken together, since which layers run in
    # which launch depends on the pass plan; algorithmic bytes: the interpolation 8 B per cell, the
    # extension 4 B per coefficient + 4 B per extended word
    fft_classes = [k for k in ("fft_ifft_pass", "fft_ie_pass", "fft_rfft_pass") if k in prof]
    fft_ms = sum(prof[k]["ms"] for k in fft_classes)
    n_launch = sum(prof[k]["n"] for k in fft_classes) / args.steps
    bytes_per_launch = work_per["fft_passes"]["bytes"] / n_launch
    avg_s = fft_ms / (n_launch * args.steps) / 1e3
    achieved = bytes_per_launch / avg_s / 1e9
    # the FFT passes are bound by integer VALU issue, not HBM (the counters in
    # profiles/commit_uneven512_hbm_traffic.md): an M31 butterfly is 11 VALU instructions of
    # arithmetic (multiply by a doubled twiddle 5, add 3, subtract 3), priced against the 32-bit
    # VALU peak; the HBM fraction of the same launches is reported beside it
    bfly = sum(lg << (lg - 1) for lg in logs) + sum(lg << (lg + b - 1) for lg in logs)
    valu_ops = VALU_OPS_PER_BUTTERFLY * bfly / n_launch
    valu_tops = valu_ops / avg_s / 1e12
    tuning_now = {str(m): list(be.fft_tuning(m)) for m in sorted(set(logs) | {lg + b for lg in logs})}
    traffic, same_shapes = profiled_traffic("fft_passes", tuning_now) if args.workload == "uneven" else (None, None)
    roofline = {"kernel": "fft passes (" + ", ".join(fft_classes) + ")", "bound": "valu",
                "achieved": round(valu_tops, 2), "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                "frac": round(valu_tops / VALU_PEAK_TOPS, 4), "traffic": traffic,
                "ops_per_butterfly": VALU_OPS_PER_BUTTERFLY, "butterflies_per_launch": bfly / n_launch,
                "launches_per_step": n_launch, "avg_launch_us": round(avg_s * 1e6, 2)}
    roofline["hbm"] = {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_launch": bytes_per_launch}
    # each launch class's own VALU fraction, from the butterflies its passes run in this plan
    per_cls = fft_butterflies(logs, b, {m: be.fft_tuning(m) for m in set(logs) | {lg + b for lg in logs}})
    roofline["by_class"] = {k: {"ms_per_step": round(prof[k]["ms"] / args.steps, 4),
                                "butterflies_per_step": per_cls[k],
                                "valu_frac": round(VALU_OPS_PER_BUTTERFLY * per_cls[k]
                                                   / (prof[k]["ms"] / args.steps / 1e3) / 1e12 / VALU_PEAK_TOPS, 4)}
                            for k in fft_classes}
    if traffic is not None:
        roofline["traffic_source"] = ("profiles/commit_uneven512_hbm_traffic.md (rocprofv3 FETCH_SIZE x 2.0 + "
                                      "WRITE_SIZE x 1.0 per dispatch)")
        roofline["traffic_pass_shapes_match"] = same_shapes
    biggest = max(prof, key=lambda k: prof[k]["ms"])
    out = {
        "metric": "committed trace cells/sec at log-20 LDE",
        "value": round(value, 1),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (M31)",
        "data": "synthetic",
        "config": {"workload": work["name"], "columns": len(logs), "cells": n_cells,
                   "log_blowup": b, "lifting_log_size": L, "seed": seed,
                   "pipeline": "interpolate + LDE + lifted Blake2s Merkle (sh_commit_trace)"},
        "roofline": roofline,
        "kernels": kernels,
        "variants": variants,
        "largest_kernel_class": biggest,
        "fft_tuning": {str(m): be.fft_tuning(m) for m in sorted(set(logs) | {lg + b for lg in logs})},
        "root": root.hex(),
    }
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (C restatements of SimdBackend and CpuBackend on the same workload) ...")
        out["cpu_baseline"] = cpu_baseline(work, root)
        log("[bench] cpu baseline done")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
