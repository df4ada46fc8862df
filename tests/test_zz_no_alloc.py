"""The allocation-free variant of compute_quotients_and_combine (static device scratch) gives the
same quotient as the allocating entry point, on two calls in a row (the second smaller, so the
shared scratch still holds the first call's larger contents), and rejects a subdomain beyond its
static scratch before launching anything."""
import ctypes

import numpy as np
import pytest

import oracle as O
from proving_amd.backend import Context, HipBackend, SecureColumnByCoords, _ptrs, _u32s

pytestmark = pytest.mark.gpu
P = O.P


def _inputs(ctx, rng, log_part, n_acc):
    """n_acc partial numerators (random QM31 columns of 2^log_part rows) with out-of-domain
    sample points drawn from a channel and random first linear terms."""
    ch = O.Channel()
    ch.mix_u64(log_part * 16 + n_acc)
    pts = [[int(v) for v in ch.random_point()] for _ in range(n_acc)]
    parts_h = [rng.integers(0, P, (4, 1 << log_part), dtype=np.uint32) for _ in range(n_acc)]
    parts = [SecureColumnByCoords.from_cpu(ctx, p) for p in parts_h]
    fls = [[int(v) for v in rng.integers(0, P, 4)] for _ in range(n_acc)]
    return pts, parts_h, parts, fls


def _call(ctx, tree, fn, pts, parts, fls, L, b):
    out = SecureColumnByCoords(ctx, L)
    rc = fn(ctx.h, tree.h, _ptrs([c for p in parts for c in p.columns]), _u32s([p.log_size for p in parts]),
            _u32s([v for pt in pts for v in pt]), _u32s([v for f in fls for v in f]), len(parts), L, b,
            _ptrs(out.columns))
    return rc, out


def test_noalloc_combine_matches_allocating_entry_point():
    ctx = Context(0)
    try:
        be = HipBackend(ctx)
        lib = ctx.lib
        rng = np.random.default_rng(77)
        b = 1
        tree = be.precompute_canonic_twiddles(22)
        # first call: the largest subdomain the static scratch holds (2^21 rows), 3 accumulations;
        # second call: smaller (2^11 rows), 2 accumulations, one of them lifted from 2^9 rows
        for L, logs in [(22, [21, 21, 21]), (12, [11, 9])]:
            print(f"noalloc combine: L={L} partial logs={logs}", flush=True)
            pts, parts_h, parts, fls = [], [], [], []
            for lg in logs:
                p, ph, pc, f = _inputs(ctx, rng, lg, 1)
                pts += p
                parts_h += ph
                parts += pc
                fls += f
            rc0, ref = _call(ctx, tree, lib.sh_compute_quotients_and_combine, pts, parts, fls, L, b)
            assert rc0 == 0, lib.sh_last_error(ctx.h).decode()
            rc1, got = _call(ctx, tree, lib.sh_compute_quotients_and_combine_noalloc, pts, parts, fls, L, b)
            assert rc1 == 0, lib.sh_last_error(ctx.h).decode()
            np.testing.assert_array_equal(got.to_cpu(), ref.to_cpu())
            if L <= 12:
                tw, itw = O.canonic_twiddles(L)
                exp = O.compute_quotients_and_combine(parts_h, pts, fls, L, b, tw, itw)
                np.testing.assert_array_equal(got.to_cpu(), exp)
        # a subdomain of 2^22 rows exceeds the static scratch: refused on the host
        pts, _, parts, fls = _inputs(ctx, rng, 4, 1)
        rc, _ = _call(ctx, tree, lib.sh_compute_quotients_and_combine_noalloc, pts, parts, fls, 23, b)
        assert rc != 0 and "static scratch" in lib.sh_last_error(ctx.h).decode()
        ctx.sync()
    finally:
        ctx.close()
