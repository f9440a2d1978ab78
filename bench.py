#!/usr/bin/env python3
"""Benchmark: SAL gradients/s on the 2,313,486-point MPAS-shaped grid.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode batched|plain] [--fields R]

One step is one evaluation of the treecode over all N points for R fields, with the grid's tree,
interaction list and interpolation data resident in HBM and the fields already on the device
(tree and list build times are reported separately in the JSON line).

  batched  R=32 fields, mixed precision (FP32 far field and near pairs above the chord cutoff,
           split-bf16 MFMA proxy-to-proxy products), accuracy mode          -> the headline value
  plain    R fields evaluated one by one in FP64 with the reference's semantics

Accuracy is measured against the FP64 direct sum at seeded sampled targets, per field.  rank 0
prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sal-fort_amd"))

GRID_N = 2313486
THETA = 0.7
PEAKS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}  # TFLOP/s, MI355X dense


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def prof_read(S):
    import ctypes as C
    L = S.lib()
    n = C.c_int()
    S.check(L.sal_prof_count(C.byref(n)))
    out = {}
    for i in range(n.value):
        buf = C.create_string_buffer(128)
        ms, cnt = C.c_double(), C.c_int64()
        S.check(L.sal_prof_get(i, buf, 128, C.byref(ms), C.byref(cnt)))
        out[buf.value.decode()] = (ms.value, cnt.value)
    return out


def work_counts(S, g, degree):
    return g.work_counts(degree)


def cpu_baseline(W, degree, stride):
    """The compiled reference's fast_sum on every stride-th interaction of its own list (its
    round-robin split, fast_sum_module.f90:304), one core; scaled to gradients/s for one field."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_py import Reference, REF_SO
    if not os.path.exists(REF_SO):
        return None
    import numpy as np
    R = Reference()
    ssh = W["ssh"][0] if W["ssh"].ndim == 2 else W["ssh"]
    # the reference prints a line per call on stdout: keep stdout for the JSON line
    fd = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        R.tree_build(W["x"], W["y"], W["z"], W["thresh"])
        R.dtt(THETA, W["thresh"])
        t = time.perf_counter()
        R.fast_sum(W["x"], W["y"], W["z"], W["area"], ssh, degree, rank=0, nproc=stride)
        dt = time.perf_counter() - t
    finally:
        os.dup2(fd, 1)
        os.close(fd)
        os.close(devnull)
    n = len(W["x"])
    return {"value": n / (dt * stride), "unit": "gradients/s", "cores": 1, "kind": "reference",
            "sample": f"reference fast_sum (amdflang -O2) over every {stride}th interaction of its DTT list "
                      f"(its own round-robin split), {dt:.1f} s, one field, d={degree}, scaled by {stride}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=["batched", "plain"], default="batched")
    ap.add_argument("--fields", type=int, default=None)
    ap.add_argument("--n", type=int, default=GRID_N)
    ap.add_argument("--degree", type=int, default=8)
    ap.add_argument("--targets", type=int, default=2000, help="sampled targets for the direct-sum check")
    ap.add_argument("--cpu-stride", type=int, default=60, help="reference CPU sample: every k-th interaction")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--m2l", type=int, default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import salfort as S

    dev = torch.device("cuda:0")
    R = args.fields or (32 if args.mode == "batched" else 1)
    t = time.perf_counter()
    W = S.workload("mpas", n=args.n, seed=0, R=R)
    n = len(W["x"])
    ssh = W["ssh"].reshape(R, n)
    log(f"inputs: n={n} R={R} ({time.perf_counter() - t:.1f} s)")
    tens = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    dx, dy, dz, da, ds = (tens(v) for v in (W["x"], W["y"], W["z"], W["area"], ssh))
    torch.cuda.synchronize()

    t = time.perf_counter()
    g = S.Geometry(device_xyz=(dx, dy, dz), cluster_thresh=W["thresh"])
    tree_s = time.perf_counter() - t
    t = time.perf_counter()
    g.dtt(THETA)
    dtt_s = time.perf_counter() - t
    P = g.counts()[0]
    log(f"tree: {P} panels {tree_s:.3f} s; dtt: {g.n_interactions} interactions {dtt_s:.3f} s")

    deg = args.degree
    if args.mode == "batched":
        opts = S.default_opts(degree=deg, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED)
        if args.m2l is not None:
            opts.m2l = args.m2l
        direct_mode, dtype = 2, "fp32"
    else:
        opts = S.default_opts(degree=deg, mode=S.MODE_REFERENCE, precision=S.PREC_FP64)
        direct_mode, dtype = 0, "fp64"
    gx = torch.zeros((R, n), dtype=torch.float64, device=dev)
    gy = torch.zeros_like(gx)
    t = time.perf_counter()
    g.eval(da, ds, gx, gy, opts)  # first call builds the evaluation plan
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t
    for _ in range(args.warmup):
        g.eval(da, ds, gx, gy, opts)
    torch.cuda.synchronize()
    S.check(S.lib().sal_prof_reset())
    S.check(S.lib().sal_prof_enable(1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.eval(da, ds, gx, gy, opts)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    S.check(S.lib().sal_prof_enable(0))
    prof = prof_read(S)
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = n * R / (ms_step * 1e-3)
    log(f"eval: {ms_step:.2f} ms/step  {value:.4g} gradients/s  (plan {plan_s:.2f} s)")
    for k, (ms, c) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
        log(f"  kernel {k:12s} {ms / max(c, 1):9.3f} ms/launch  x{c}")

    # accuracy against the FP64 direct sum at seeded targets
    rng = np.random.default_rng(1234)
    tg = torch.as_tensor(np.sort(rng.choice(n, args.targets, replace=False)).astype(np.int32), device=dev)
    nchk = min(R, 4)
    rel = []
    for r in range(nchk):
        ex, ey = S.direct_sum(dx, dy, dz, da, ds[r].contiguous(), tg, direct_mode)
        tx, ty = gx[r][tg.long()], gy[r][tg.long()]
        rel.append(float(torch.sqrt(((tx - ex) ** 2 + (ty - ey) ** 2).sum() / (ex ** 2 + ey ** 2).sum())))
    log(f"relL2 vs FP64 direct sum (mode {direct_mode}) per field: {['%.3e' % v for v in rel]}")

    # roofline of the dominant kernel: algorithmic pair-evaluation flops per launch / its duration
    K = work_counts(S, g, deg)
    dom, (dom_ms, dom_n) = max(prof.items(), key=lambda kv: kv[1][0])
    avg_ms = dom_ms / max(dom_n, 1)
    per_launch_fields = R if args.mode == "batched" else 1
    fl = 34 + 4 * per_launch_fields
    flops = {"leaf64": (K[0] + K[1]) * 38, "proxy64": (K[2] + K[3]) * 38,
             "near32": K[0] * fl, "far32": (K[1]) * fl}.get(dom)
    roof = None
    if flops is not None:
        peak = PEAKS["fp64" if dtype == "fp64" else "fp32"]
        ach = flops / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "engine": "vector ALU (" + ("FP64" if dtype == "fp64" else "FP32") + ")",
                "kernel": dom, "achieved": round(ach, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 4), "traffic": None,
                "flops_per_launch": int(flops), "avg_launch_ms": round(avg_ms, 4)}

    cpu = None
    if not args.no_cpu:
        log("cpu baseline: reference fast_sum sample ...")
        cpu = cpu_baseline(W, deg, args.cpu_stride)
        if cpu:
            log(f"cpu baseline: {cpu['value']:.4g} gradients/s")

    out = {
        "metric": "2.3M-pt SAL gradients/s @relL2<=1e-6, batched vs plain GPU; p99 latency",
        "value": value, "unit": "gradients/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic (seeded MPAS-shaped grid and harmonic fields)",
        "config": {"workload": f"MPAS-shaped N={n}, {R} field(s), {args.mode}", "n_points": n, "fields": R,
                   "degree": deg, "theta": THETA, "cluster_thresh": W["thresh"], "panels": P,
                   "interactions": g.n_interactions, "pair_evals": [int(k) for k in K],
                   "tree_build_s": round(tree_s, 4), "dtt_s": round(dtt_s, 4), "plan_s": round(plan_s, 3),
                   "relL2_vs_direct": rel, "direct_mode": direct_mode, "targets": args.targets},
        "roofline": roof, "cpu_baseline": cpu,
        "kernels_ms_per_launch": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
