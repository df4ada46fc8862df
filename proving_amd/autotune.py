"""Writes the FFT pass-shape table the benches and provers load by default
(proving_amd/fft_tuning_gfx950.json).

The device autotuner (sh_autotune_fft) times candidate pass shapes per transform size on the size
class alone; in a batched commit the classes overlap on the device, and the shapes that win alone
are not always the ones that win together.  This script starts from the device autotuner's picks
and then runs coordinate descent over the transform sizes of the batched commit of 512 mixed
columns (256 x 2^12, 128 x 2^16, 96 x 2^18, 32 x 2^20, blowup 1), timing the whole captured
commit for every distinct pass plan of a size, until a sweep changes nothing.  Sizes the commit
does not use keep the autotuner's picks (tuned on 256 x 2^20 and 200 x 2^21, the transforms of
the full proof and of the FRI + DEEP benchmark).
Run on the target GPU:  python -m proving_amd.autotune [--sweeps 2] [--out PATH]
"""
import argparse
import json
import os
import time

from .backend import BaseColumn, Context, HashLayer, HipBackend
from .workloads import UNEVEN_COMMIT, splitmix64_columns

HERE = os.path.dirname(os.path.abspath(__file__))


def plan(top, shape):
    """The pass plan sh_fft_* derive from a shape for a transform of `top` layers: the block pass
    and the strided passes (lo, hi), as in plan_passes (fft.hip)."""
    block_log, stride_layers, _, _ = shape
    rem = 0
    if top > block_log:
        rem = (top - block_log) % stride_layers
        if rem and rem <= 2 and block_log + rem <= 13:
            block_log += rem  # a remainder of one or two layers joins the block pass
            rem = 0
    bh = min(top, block_log)
    passes, lo = [(0, bh, 0)], bh
    while lo < top:
        g = min(top - lo, stride_layers)
        if rem and rem <= 2 and stride_layers < top - lo < 2 * stride_layers:
            g = (top - lo + 1) // 2  # otherwise the last two strided passes share the layers
        passes.append((lo, lo + g))
        lo += g
    return tuple(passes)


class CommitTimer:
    def __init__(self, ctx, be, tree, evals, b):
        self.ctx, self.be, self.tree, self.evals, self.b = ctx, be, tree, evals, b
        L = max(c.log_size for c in evals) + b
        self.coeffs = [BaseColumn(ctx, c.log_size) for c in evals]
        self.ldes = [BaseColumn(ctx, c.log_size + b) for c in evals]
        self.layers = [HashLayer(ctx, k) for k in range(L + 1)]

    def __call__(self, steps=20, reps=3):
        graph = self.be.commit_trace_graph(self.evals, self.b, self.tree, self.coeffs, self.ldes, self.layers)
        for _ in range(3):
            graph.launch()
        best = 1e30
        for _ in range(reps):
            self.ctx.sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                graph.launch()
            self.ctx.sync()
            best = min(best, (time.perf_counter() - t0) * 1e3 / steps)
        graph.free()
        return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sweeps", type=int, default=2)
    ap.add_argument("--out", default=os.path.join(HERE, "fft_tuning_gfx950.json"))
    args = ap.parse_args()
    ctx = Context(0)
    be = HipBackend(ctx)
    w = UNEVEN_COMMIT
    logs, b = w["logs"], w["log_blowup"]
    tree = be.precompute_canonic_twiddles(max(logs) + b + 1)
    evals = [BaseColumn.from_cpu(ctx, c) for c in splitmix64_columns(w["seed"], logs)]
    be.autotune_fft([20] * 256, 1, tree)
    be.autotune_fft([21] * 200, 1, tree)
    be.autotune_fft(logs, b, tree)
    timer = CommitTimer(ctx, be, tree, evals, b)
    sizes = sorted(set(logs) | {lg + b for lg in logs})
    shapes = [(blk, st, rb, rs) for blk in (11, 12, 13) for st in (5, 6, 7, 8) for rb in (1, 2, 3)
              for rs in (1, 2, 3)]
    cur = timer()
    print(f"autotuner picks: commit step {cur:.4f} ms", flush=True)
    for sweep in range(args.sweeps):
        changed = False
        for m in sizes:
            keep = be.fft_tuning(m)
            seen, best_t, best_ms = set(), keep, cur
            for t in shapes:
                p = plan(m, t)
                key = (p, t[2], t[3] if len(p) > 1 else 0)  # r_stride only matters with strided passes
                if key in seen:
                    continue
                seen.add(key)
                if tuple(t) == tuple(keep):
                    continue
                be.set_fft_tuning(m, *t)
                ms = timer()
                if ms < best_ms * 0.997:  # a change has to beat the noise
                    best_t, best_ms = t, ms
            be.set_fft_tuning(m, *best_t)
            if tuple(best_t) != tuple(keep):
                changed = True
                cur = timer()
            print(f"sweep {sweep} size {m}: {tuple(keep)} -> {tuple(best_t)}, {len(seen)} plans: {cur:.4f} ms",
                  flush=True)
        if not changed:
            break
    final = timer(steps=40)
    table = {str(m): list(be.fft_tuning(m)) for m in range(3, 23)}
    print(f"final commit step {final:.4f} ms: {table}", flush=True)
    doc = {"device": "gfx950", "workload": "commit of 512 mixed columns, blowup 1, captured graph",
           "commit_step_ms": round(final, 4), "shapes": table,
           "fields": "block_log, stride_layers, r_block, r_stride"}
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(f"wrote {args.out}", flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
