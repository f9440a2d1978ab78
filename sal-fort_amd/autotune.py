#!/usr/bin/env python3
"""Autotuner of the batched mixed-precision evaluation: interpolation degree x leaf size x precision
variant x stored operator.

    python sal-fort_amd/autotune.py [--workload mpas|mom6|fibonacci] [--fields R] [--out tuning.json]

For one grid and R fields resident on the device it sweeps
    degree d in {6, 7, 8, 10}, cluster_thresh in {33, 65, 129, 257} (leaves of at most 32..256 points),
    precision variants (m2l, near64):
        (SAL_M2L_BF16X3, 0)     split-bf16 far field, every near pair in FP32
        (SAL_M2L_BF16X3, 2)     the same with close near pairs (chord < 2 h) in FP64
        (SAL_M2L_BF16X6, 2)     three-part split-bf16 far field (FP32 precision) with FP64 per-chunk
                                sums, close pairs in FP64
      (grids whose far field carries much of the gradient, such as the land-masked MOM6-shaped one,
      need the last: its two-part split-bf16 far field alone is about 3e-7 of the result; the
      FP32-MFMA far field, SAL_M2L_FP32_MFMA, gives the same error more slowly and is not swept),
    the stored operator where it fits the device memory left (with a margin), else kernel values
      evaluated on every call,
with the RHS tile that covers R (32 above 16 fields: one pass of every kernel).  Each setting is
timed (tree and list built per leaf size, plan per degree and variant, best of a few steps) and its
relL2 against the FP64 direct sum measured at seeded targets for a few fields.  The choice is the
fastest setting whose worst field error is at most `--bar` (default 4e-7: 2.5x below the 1e-6
target, margin for the targets the check does not sample); where none meets it, the fastest under
`--fallback-bar` (default 7e-7), and the result says which bar was used.

The sweep is pruned by cost order: at one leaf size a higher degree only costs more, and the
variants above are listed cheapest first, so for each leaf size the degrees run in ascending order
and the variants in listed order until one setting meets the bar; the rest of that leaf size cannot
be faster and is skipped (`prune=False` times everything).
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
VARIANTS = ((S.M2L_BF16X3, 0), (S.M2L_BF16X3, 2), (S.M2L_BF16X6, 2))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _measure(g, da, ds, gx, gy, d, tile, store, m2l, near64, steps, exact, tg):
    """Best of `steps` timed evaluations of one setting and its per-field errors."""
    import torch
    n, R = da.numel(), ds.shape[0]
    o = S.default_opts(degree=d, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, m2l=m2l, rhs_tile=tile, store=store,
                       near64=near64)
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
    errs = []
    for r, (ex, ey) in enumerate(exact):
        tx, ty = gx[r][tg.long()], gy[r][tg.long()]
        errs.append(float(torch.sqrt(((tx - ex) ** 2 + (ty - ey) ** 2).sum() / (ex ** 2 + ey ** 2).sum())))
    return dict(degree=d, rhs_tile=tile, store=store, m2l=m2l, near64=near64, ms=best * 1e3,
                gradients_per_s=n * R / best, relL2_max=max(errs), first_call_s=setup)


def tune(dx, dy, dz, da, ds, degrees=DEGREES, threshes=THRESHES, variants=VARIANTS, steps=2, targets=2000,
         check_fields=4, bar=4e-7, fallback_bar=7e-7, margin=24 << 30, quiet=False, prune=True):
    """Device tensors: coordinates, areas [n], fields [R, n].  Returns (rows, best row); the best row's
    "bar" is the error bar it was chosen under."""
    import torch
    say = (lambda *a: None) if quiet else log
    n, R = dx.numel(), ds.shape[0]
    tile = 32 if R > 16 else 16
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
        met = False
        for d in sorted(degrees):
            for m2l, near64 in variants:
                st = g.plan_stats(S.default_opts(degree=d, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED,
                                                 near64=near64))
                need = st["stored_near_bytes"] + st["stored_far_bytes"]
                store = 1 if need <= torch.cuda.mem_get_info()[0] - margin else 0
                row = _measure(g, da, ds, gx, gy, d, tile, store, m2l, near64, steps, exact, tg)
                row.update(cluster_thresh=th, tree_dtt_s=build_s, ok=row["relL2_max"] <= bar,
                           stored_bytes=need if store else 0, close_pairs=st["close_pairs"])
                rows.append(row)
                say(f"  d={d:2d} thresh={th:3d} m2l={m2l} near64={near64} store={store}: {row['ms']:8.2f} ms  "
                    f"relL2 {row['relL2_max']:.3e}{'' if row['ok'] else '  (over the bar)'}")
                g.release_plans()  # frees this setting's plans and stored operator before the next
                met = met or row["ok"]
                if met and prune:
                    break
            if met and prune:
                break
        g.close()
    for b in (bar, fallback_bar):
        good = [r for r in rows if r["relL2_max"] <= b]
        if good:
            best = dict(min(good, key=lambda r: r["ms"]), bar=b)
            if b != bar:
                say(f"  no setting under {bar:g}; the fastest under {b:g}")
            return rows, best
    return rows, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="mpas", choices=["mpas", "mom6", "fibonacci"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--fields", type=int, default=32)
    ap.add_argument("--bar", type=float, default=4e-7)
    ap.add_argument("--fallback-bar", type=float, default=7e-7)
    ap.add_argument("--out", default=None, help="write the table and the choice as JSON")
    ap.add_argument("--serving-fields", action="store_true",
                    help="tune on the serving loop's fields (degree <= 64 harmonics plus 0.02 m noise, seed 1) "
                         "instead of the workload's own")
    ap.add_argument("--no-prune", dest="prune", action="store_false", help="time every setting")
    args = ap.parse_args()
    import torch
    dev = torch.device("cuda:0")
    W = S.workload(args.workload, n=args.n, seed=0, R=args.fields)
    n = len(W["x"])
    if args.serving_fields:
        W["ssh"] = S.gen_ssh(W["x"], W["y"], W["z"], lmax=64, amp=0.1, noise=0.02, seed=1, R=args.fields)
    tens = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    dx, dy, dz, da = (tens(v) for v in (W["x"], W["y"], W["z"], W["area"]))
    ds = tens(W["ssh"].reshape(args.fields, n))
    t = time.perf_counter()
    rows, best = tune(dx, dy, dz, da, ds, bar=args.bar, fallback_bar=args.fallback_bar, prune=args.prune)
    out = dict(workload=args.workload, n=n, fields=args.fields, bar=args.bar, best=best, table=rows,
               seconds=round(time.perf_counter() - t, 1))
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps(dict(best=best, settings_tried=len(rows), seconds=out["seconds"])))


if __name__ == "__main__":
    main()
