#!/usr/bin/env python3
"""Autotuner of the batched mixed-precision evaluation: interpolation degree x leaf size x RHS tile
x stored operator.

    python sal-fort_amd/autotune.py [--n N] [--fields R] [--out tuning.json]

For one grid and R fields resident on the device it sweeps
    degree d in {6, 7, 8, 10}, cluster_thresh in {33, 65, 129, 257} (leaves of at most 32..256 points),
    RHS tile in {16, 32} fields per kernel pass (a tile of 8 would leave half of every 16-wide MFMA
    column tile idle, so it is not offered),
    stored operator off / on (on only where it fits the device memory left, with a margin),
times each setting (tree and list built per leaf size, plan per degree, best of a few steps), and
measures its relL2 against the FP64 direct sum at seeded targets for a few fields.  The choice is
the fastest setting whose worst field error is at most `--bar` (default 5e-7: half of the 1e-6
target, margin for the targets the check does not sample).  Where no setting meets it (grids whose
error floor is the FP32 near field, such as the MOM6-shaped one), the choice is the fastest setting
under `--fallback-bar` (default 7e-7), and the result says which bar was used.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import salfort as S  # noqa: E402

THETA = 0.7
DEGREES = (6, 7, 8, 10)
THRESHES = (33, 65, 129, 257)
TILES = (16, 32)
STORES = (0, 1)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _measure(g, da, ds, gx, gy, d, tile, store, m2l, steps, exact, tg, errs):
    """Best of `steps` timed evaluations of one setting, and the per-field error (measured once per
    degree: neither the tile nor the stored operator changes the arithmetic of a field)."""
    import torch
    n, R = da.numel(), ds.shape[0]
    o = S.default_opts(degree=d, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, m2l=m2l, rhs_tile=tile, store=store)
    t = time.perf_counter()
    g.eval(da, ds, gx, gy, o)  # plan (and stored operator) + warm-up
    torch.cuda.synchronize()
    setup = time.perf_counter() - t
    best = float("inf")
    for _ in range(steps):
        t = time.perf_counter()
        g.eval(da, ds, gx, gy, o)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t)
    if errs is None:
        errs = []
        for r, (ex, ey) in enumerate(exact):
            tx, ty = gx[r][tg.long()], gy[r][tg.long()]
            errs.append(float(torch.sqrt(((tx - ex) ** 2 + (ty - ey) ** 2).sum() / (ex ** 2 + ey ** 2).sum())))
    row = dict(degree=d, rhs_tile=tile, store=store, ms=best * 1e3, gradients_per_s=n * R / best,
               relL2_max=max(errs), first_call_s=setup)
    return row, errs


def tune(dx, dy, dz, da, ds, degrees=DEGREES, threshes=THRESHES, tiles=TILES, stores=STORES, steps=2, targets=400,
         check_fields=2, bar=5e-7, fallback_bar=7e-7, m2l=S.M2L_BF16X3, margin=24 << 30):
    """Device tensors: coordinates, areas [n], fields [R, n].  Returns (rows, best row); the best row's
    "bar" is the error bar it was chosen under."""
    import torch
    n, R = dx.numel(), ds.shape[0]
    rng = np.random.default_rng(99)
    tg = torch.as_tensor(np.sort(rng.choice(n, min(targets, n), replace=False)).astype(np.int32), device=dx.device)
    exact = [S.direct_sum(dx, dy, dz, da, ds[r].contiguous(), tg, 2) for r in range(min(check_fields, R))]
    gx = torch.empty_like(ds)
    gy = torch.empty_like(ds)
    rows = []
    for th in threshes:
        t = time.perf_counter()
        g = S.Geometry(device_xyz=(dx, dy, dz), cluster_thresh=th)
        g.dtt(THETA)
        build_s = time.perf_counter() - t
        for d in degrees:
            errs = None
            for store in stores:
                need = 0
                if store:
                    st = g.plan_stats(S.default_opts(degree=d, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED))
                    need = st["stored_near_bytes"] + st["stored_far_bytes"]
                    if need > torch.cuda.mem_get_info()[0] - margin:
                        log(f"  d={d:2d} thresh={th:3d} stored operator: {need / 2**30:.1f} GiB does not fit, skipped")
                        continue
                for tile in tiles:
                    row, errs = _measure(g, da, ds, gx, gy, d, tile, store, m2l, steps, exact, tg, errs)
                    row.update(cluster_thresh=th, tree_dtt_s=build_s, ok=row["relL2_max"] <= bar)
                    if store:
                        row["stored_bytes"] = need
                    rows.append(row)
                    log(f"  d={d:2d} thresh={th:3d} tile={tile:2d} store={store}: {row['ms']:8.2f} ms  "
                        f"{row['gradients_per_s']:.4g} gradients/s  relL2 {row['relL2_max']:.3e}"
                        f"{'' if row['ok'] else '  (over the bar)'}")
            g.release_plans()  # frees this degree's plans and stored operator before the next
        g.close()
    for b in (bar, fallback_bar):
        good = [r for r in rows if r["relL2_max"] <= b]
        if good:
            best = dict(min(good, key=lambda r: r["ms"]), bar=b)
            if b != bar:
                log(f"  no setting under {bar:g}; the fastest under {b:g}")
            return rows, best
    return rows, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="mpas", choices=["mpas", "mom6", "fibonacci"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--fields", type=int, default=32)
    ap.add_argument("--bar", type=float, default=5e-7)
    ap.add_argument("--fallback-bar", type=float, default=7e-7)
    ap.add_argument("--out", default=None, help="write the table and the choice as JSON")
    args = ap.parse_args()
    import torch
    dev = torch.device("cuda:0")
    W = S.workload(args.workload, n=args.n, seed=0, R=args.fields)
    n = len(W["x"])
    tens = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    dx, dy, dz, da = (tens(v) for v in (W["x"], W["y"], W["z"], W["area"]))
    ds = tens(W["ssh"].reshape(args.fields, n))
    rows, best = tune(dx, dy, dz, da, ds, bar=args.bar, fallback_bar=args.fallback_bar)
    out = dict(workload=args.workload, n=n, fields=args.fields, bar=args.bar, best=best, table=rows)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps(dict(best=best)))


if __name__ == "__main__":
    main()
