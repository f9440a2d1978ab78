#!/usr/bin/env python3
"""Benchmark: SAL gradients/s on the 2,313,486-point MPAS-shaped grid.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode batched|plain] [--fields R]
                    [--workload mpas|mom6|fibonacci] [--graph] [--calib-copy-mb MB]

One step is one evaluation of the treecode over all N points for R fields, with the grid's tree,
interaction list and interpolation data resident in HBM and the fields already on the device
(tree, list and plan build times are reported separately in the JSON line).

  batched  R=32 fields, mixed precision (FP32 kernel values, FP32 MFMA near-field products with
           FP64 per-panel accumulation, split-bf16 MFMA proxy products), accuracy mode -> headline
  plain    R fields evaluated one by one in FP64 with the reference's semantics

Accuracy is measured against the FP64 direct sum at seeded sampled targets, per field.  The one
JSON line goes to stdout; progress goes to stderr.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sal-fort_amd"))

THETA = 0.7
# MI355X published dense peaks, TFLOP/s; FP32 vector (packed) and FP32 MFMA share the 157.3 figure
PEAKS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
HBM_PEAK_GBS = 8000.0


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


def kernel_flops(kernel, K, fields):
    """Algorithmic flops of one launch of `kernel` over `fields` fields: 34 + 4 R per pair evaluation
    (38 at R = 1), pairs K = [PP c_t c_s, PC c_t m, CP m c_s, CC m^2] summed over the list."""
    per_pair = 34 + 4 * fields
    if kernel in ("leaf64", "near_mix"):   # target points: PP and PC entries
        return (K[0] + K[1]) * per_pair
    if kernel in ("proxy64", "far_mix"):   # target proxies: CP and CC entries
        return (K[2] + K[3]) * per_pair
    return None


def read_traffic(kernel, tag):
    """HBM bytes per launch of `kernel` measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by
    the same counters on a copy of known size) and kept as a text summary under profiles/."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{tag}_traffic.txt")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        m = re.search(r"^hbm_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read(), re.M)
    return (float(m.group(1)), os.path.relpath(path, ROOT)) if m else (None, None)


def cpu_cores():
    """Host cores this process may use, at most 16 (the share of one GPU's host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(W, degree, stride, cores):
    """The compiled reference's fast_sum as `cores` MPI ranks would run it, one worker process per
    host core (oracle/ref_sample.py), over every stride-th interaction of its own list: worker r
    takes entries i with i mod (stride cores) == r stride, so together they take i mod stride == 0
    (the round-robin split of fast_sum_module.f90:304).  Wall time from the common start to the last
    worker's end, times stride: gradients/s for one field."""
    import subprocess
    import tempfile
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_py import REF_SO
    if not os.path.exists(REF_SO):
        return None
    ssh = W["ssh"][0] if W["ssh"].ndim == 2 else W["ssh"]
    worker = os.path.join(ROOT, "oracle", "ref_sample.py")
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "grid.npz")
        np.savez(inp, x=W["x"], y=W["y"], z=W["z"], area=W["area"], ssh=ssh, thresh=W["thresh"])
        procs = [subprocess.Popen([sys.executable, worker, inp, str(r * stride), str(stride * cores), str(degree),
                                   str(THETA)], stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)
                 for r in range(cores)]
        try:
            for p in procs:
                if p.stdout.readline().strip() != "ready":
                    raise RuntimeError("reference worker failed before its sample")
            t = time.perf_counter()
            for p in procs:
                p.stdin.write("go\n")
                p.stdin.flush()
            own = []
            for p in procs:
                line = p.stdout.readline().split()
                if not line or line[0] != "done":
                    raise RuntimeError("reference worker failed during its sample")
                own.append(float(line[1]))
            dt = time.perf_counter() - t
        finally:
            for p in procs:
                if p.poll() is None:
                    p.kill()
                p.wait()
    n = len(W["x"])
    return {"value": n / (dt * stride), "unit": "gradients/s", "cores": cores, "kind": "reference",
            "sample": f"reference fast_sum (amdflang -O2) as {cores} MPI-rank-style worker processes, one per "
                      f"host core, over every {stride}th entry of its own interaction list (its round-robin "
                      f"split), {dt:.2f} s wall ({sum(own):.1f} s of CPU work, slowest rank {max(own):.2f} s), "
                      f"one field, d={degree}, time x {stride}"}


def run_stt(args, S, W, g, deg, thresh, tree_s):
    """The MOM6 driver's FP64 path: SINGLE_TREE_TRAVERSAL for every point, then
    proxy_source_compute + pc_compute + pp_compute (mom6_fast_sum_module.f90:33-192), with the
    reference's transposed weight read, on host arrays as the driver passes them.  A step is one
    field; its time includes the field upload and the sums' download."""
    import numpy as np
    x, y, z, area = W["x"], W["y"], W["z"], W["area"]
    n = len(x)
    ssh = np.ascontiguousarray(W["ssh"].reshape(-1, n)[0])
    t = time.perf_counter()
    npc, npp = g.stt(THETA, x, y, z)
    stt_s = time.perf_counter() - t
    _, _, _, pps = g.stt_export()
    count = g.export()["count"]
    m = (deg + 1) ** 2
    K = [int(count[pps].astype(np.int64).sum()), int(npc) * m]   # pp pairs, pc pairs
    for _ in range(args.warmup):
        g.mom6_sums(area, ssh, deg)
    S.check(S.lib().sal_prof_reset())
    S.check(S.lib().sal_prof_enable(1))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gx, gy = g.mom6_sums(area, ssh, deg)
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    S.check(S.lib().sal_prof_enable(0))
    prof = prof_read(S)
    kms = {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items()}
    flops = 38 * (K[0] + K[1])
    ach = flops / (kms["mom6_sums"] * 1e-3) / 1e12
    log(f"stt: {npc} pc + {npp} pp entries ({stt_s:.3f} s); sums {ms:.2f} ms per field; kernels {kms}")
    out = {"metric": "MOM6 FP64 gradients/s (single tree traversal path, reference semantics)",
           "value": n / (ms * 1e-3), "unit": "gradients/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp64", "data": "synthetic (seeded grid and harmonic fields)",
           "config": {"workload": f"{args.workload} N={n}, 1 field, stt", "n_points": n, "degree": deg,
                      "theta": THETA, "cluster_thresh": thresh, "pc_entries": int(npc), "pp_entries": int(npp),
                      "pair_evals": K, "tree_build_s": round(tree_s, 4), "stt_s": round(stt_s, 4),
                      "timing": "host arrays in and out (the driver's call), PCIe included"},
           "roofline": {"bound": "mfma", "kernel": "mom6_sums", "engine": "FP64 VALU", "achieved": round(ach, 3),
                        "peak": PEAKS["fp64"], "unit": "TFLOP/s", "frac": round(ach / PEAKS["fp64"], 4),
                        "flops_per_launch": flops, "avg_launch_ms": kms["mom6_sums"]},
           "kernels_ms_per_launch": kms}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=["batched", "plain", "stt"], default="batched",
                    help="stt: the MOM6 driver's FP64 path (single tree traversal + per-point sums)")
    ap.add_argument("--workload", choices=["mpas", "mom6", "fibonacci"], default="mpas")
    ap.add_argument("--fields", type=int, default=None)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--degree", type=int, default=None)
    ap.add_argument("--thresh", type=int, default=None, help="cluster_thresh (leaf size < thresh)")
    ap.add_argument("--targets", type=int, default=2000, help="sampled targets for the direct-sum check")
    ap.add_argument("--cpu-stride", type=int, default=15, help="reference CPU sample: every k-th interaction")
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="reference CPU sample: worker processes, one per core (0: the host cores, at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--m2l", type=int, default=None)
    ap.add_argument("--rhs-tile", type=int, default=0)
    ap.add_argument("--store", type=int, default=None,
                    help="batched: stream a stored operator (1) or evaluate kernel values on every call (0)")
    ap.add_argument("--no-tune", action="store_true",
                    help="batched: skip the autotuner and use the workload's degree and leaf size")
    ap.add_argument("--graph", action="store_true", help="replay the evaluation as a captured HIP graph")
    ap.add_argument("--serve-requests", type=int, default=60,
                    help="requests of the serving leg (batched mpas only; 0 skips it)")
    ap.add_argument("--serve-rate", default="15,40,120",
                    help="mean arrivals per second of the serving leg (comma list: a sweep; the first is reported)")
    ap.add_argument("--no-cache", dest="cache", action="store_false",
                    help="skip the geometry cache leg (save and reload the tree and lists)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="run one device copy of this size first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import salfort as S

    dev = torch.device("cuda:0")
    if args.calib_copy_mb > 0:
        nb = args.calib_copy_mb << 20
        a = torch.ones(nb // 4, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration copy: {nb} bytes read, {nb} bytes written")
        del a, b

    R = args.fields or (32 if args.mode == "batched" else 1)
    t = time.perf_counter()
    W = S.workload(args.workload, n=args.n, seed=0, R=R)
    thresh = args.thresh or W["thresh"]
    deg = args.degree or W["degree"]
    n = len(W["x"])
    ssh = W["ssh"].reshape(R, n)
    log(f"inputs: {args.workload} n={n} R={R} ({time.perf_counter() - t:.1f} s)")
    tens = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    dx, dy, dz, da, ds = (tens(v) for v in (W["x"], W["y"], W["z"], W["area"], ssh))
    torch.cuda.synchronize()

    tuned = None
    if args.mode == "batched" and not args.no_tune and args.degree is None and args.thresh is None:
        import autotune
        log("autotune: degree x leaf size x RHS tile ...")
        t = time.perf_counter()
        rows, best = autotune.tune(dx, dy, dz, da, ds)
        tuned = {"best": best, "table": rows, "seconds": round(time.perf_counter() - t, 1),
                 "bar": best["bar"] if best else None}
        if best is not None:
            deg, thresh = best["degree"], best["cluster_thresh"]
            if not args.rhs_tile:
                args.rhs_tile = best["rhs_tile"]
            if args.store is None:
                args.store = best["store"]
            log(f"autotune: d={deg} cluster_thresh={thresh} rhs_tile={args.rhs_tile} store={args.store} "
                f"({tuned['seconds']} s)")

    t = time.perf_counter()
    g = S.Geometry(device_xyz=(dx, dy, dz), cluster_thresh=thresh)
    tree_s = time.perf_counter() - t
    if args.mode == "stt":
        run_stt(args, S, W, g, deg, thresh, tree_s)
        return
    t = time.perf_counter()
    g.dtt(THETA)
    dtt_s = time.perf_counter() - t
    P = g.counts()[0]
    log(f"tree: {P} panels {tree_s:.3f} s; dtt: {g.n_interactions} interactions {dtt_s:.3f} s")

    if args.mode == "batched":
        opts = S.default_opts(degree=deg, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED)
        if args.m2l is not None:
            opts.m2l = args.m2l
        direct_mode, dtype = 2, "fp32"
    else:
        opts = S.default_opts(degree=deg, mode=S.MODE_REFERENCE, precision=S.PREC_FP64)
        direct_mode, dtype = 0, "fp64"
    opts.rhs_tile = args.rhs_tile
    opts.graph = 1 if args.graph else 0
    if args.mode == "batched":
        opts.store = 1 if args.store else 0
    gx = torch.zeros((R, n), dtype=torch.float64, device=dev)
    gy = torch.zeros_like(gx)
    t = time.perf_counter()
    g.eval(da, ds, gx, gy, opts)  # first call builds the evaluation plan
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t
    for _ in range(args.warmup):
        g.eval(da, ds, gx, gy, opts)
    torch.cuda.synchronize()

    # timed region: per-kernel HIP events on the evaluation stream, per-step events for the latency
    # spread, wall clock around all K steps
    S.check(S.lib().sal_prof_reset())
    S.check(S.lib().sal_prof_enable(0 if args.graph else 1))
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record()
    for i in range(args.steps):
        g.eval(da, ds, gx, gy, opts)
        ev[i + 1].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    S.check(S.lib().sal_prof_enable(0))
    prof = prof_read(S)
    ms_step = (t1 - t0) * 1e3 / args.steps
    step_ms = sorted(ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps))
    value = n * R / (ms_step * 1e-3)
    log(f"eval: {ms_step:.2f} ms/step  {value:.4g} gradients/s  (plan {plan_s:.2f} s)")
    for k, (ms, c) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
        log(f"  kernel {k:12s} {ms / max(c, 1):9.3f} ms/launch  x{c}")

    # accuracy against the FP64 direct sum at seeded targets
    rng = np.random.default_rng(1234)
    tg = torch.as_tensor(np.sort(rng.choice(n, min(args.targets, n), replace=False)).astype(np.int32), device=dev)
    nchk = min(R, 4)
    rel = []
    for r in range(nchk):
        ex, ey = S.direct_sum(dx, dy, dz, da, ds[r].contiguous(), tg, direct_mode)
        tx, ty = gx[r][tg.long()], gy[r][tg.long()]
        rel.append(float(torch.sqrt(((tx - ex) ** 2 + (ty - ey) ** 2).sum() / (ex ** 2 + ey ** 2).sum())))
    log(f"relL2 vs FP64 direct sum (mode {direct_mode}) per field: {['%.3e' % v for v in rel]}")

    # roofline of the dominant kernel: algorithmic pair-evaluation flops per launch / its duration
    K = g.work_counts(deg)
    plan_info = g.plan_stats(opts) if args.mode == "batched" else None
    dev_bytes = g.device_bytes()
    roof = None
    if prof:
        dom, (dom_ms, dom_n) = max(prof.items(), key=lambda kv: kv[1][0])
        avg_ms = dom_ms / max(dom_n, 1)
        per_launch = 1 if args.mode == "plain" else min(R, opts.rhs_tile or (32 if R > 16 else 16))
        flops = kernel_flops(dom, K, per_launch)
        tag = f"{args.workload}{n}_r{R}_d{deg}_t{thresh}_{args.mode}" + ("_stored" if opts.store else "")
        traffic, tsrc = read_traffic(dom, tag)
        if flops is not None and opts.store and dom in ("near_mix", "far_mix"):
            # streaming the stored operator: HBM-bound.  Algorithmic bytes per launch: the stored
            # values once, plus (near) the FP32 field rows read and the FP64 east/north rows written
            ps = plan_info
            byts = (ps["stored_near_bytes"] + n * per_launch * (4 + 16)) if dom == "near_mix" else ps["stored_far_bytes"]
            ach = byts / (avg_ms * 1e-3) / 1e9
            roof = {"bound": "hbm", "kernel": dom, "engine": "stored-operator stream, FP32/bf16 MFMA products",
                    "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "traffic_source": tsrc, "bytes_per_launch": int(byts),
                    "avg_launch_ms": round(avg_ms, 4), "flops_per_launch": int(flops),
                    "TFLOPs": round(flops / (avg_ms * 1e-3) / 1e12, 2)}
        elif flops is not None:
            peak = PEAKS[dtype]
            ach = flops / (avg_ms * 1e-3) / 1e12
            roof = {"bound": "mfma", "kernel": dom,
                    "engine": "FP64 VALU" if dtype == "fp64" else "FP32 VALU kernel values + FP32 MFMA products",
                    "achieved": round(ach, 3), "peak": peak, "unit": "TFLOP/s", "frac": round(ach / peak, 4),
                    "traffic": traffic, "traffic_source": tsrc,
                    "flops_per_launch": int(flops), "avg_launch_ms": round(avg_ms, 4)}
            if traffic:
                roof["hbm_GBps"] = round(traffic / (avg_ms * 1e-3) / 1e9, 1)
                roof["hbm_frac"] = round(roof["hbm_GBps"] / HBM_PEAK_GBS, 4)

    # on-disk geometry cache: save and reload the tree and lists, against building them
    cache = None
    if args.cache:
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "grid.salgeom")
            t = time.perf_counter()
            g.save(path)
            save_s = time.perf_counter() - t
            t = time.perf_counter()
            h = S.Geometry.load(path)
            load_s = time.perf_counter() - t
            same = all(np.array_equal(a, b) for a, b in zip(g.dtt_export(), h.dtt_export()))
            cache = {"file_bytes": os.path.getsize(path), "save_s": round(save_s, 4), "load_s": round(load_s, 4),
                     "build_tree_dtt_s": round(tree_s + dtt_s, 4), "lists_identical": bool(same)}
            h.close()
        log(f"geometry cache: {cache}")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline: reference fast_sum sample ...")
        cpu = cpu_baseline(W, deg, args.cpu_stride, args.cpu_cores or cpu_cores())
        if cpu:
            log(f"cpu baseline: {cpu['value']:.4g} gradients/s")

    serving = None
    if args.serve_requests > 0 and args.mode == "batched" and args.workload == "mpas" and args.n is None:
        import serve
        del gx, gy
        g.close()  # the serving leg keeps its own geometries (and stored operators) resident
        rates = [float(r) for r in args.serve_rate.split(",")]
        log(f"serving: {args.serve_requests} requests at {rates}/s on fibonacci, mom6, mpas ...")
        geos = serve.build(["fibonacci", "mom6", "mpas"], degree=deg, thresh=thresh, store=opts.store)
        sweep = []
        for k, rate in enumerate(rates):
            for G in geos:
                G["free"] = list(range(len(G["outs"])))
            srv = S.Server(slots=3, batching=True)
            serve.register(srv, geos)
            _, summ = serve.run_load(srv, geos, rate, args.serve_requests, seed=k, check_every=6)
            srv.close()
            s = {key: summ[key] for key in ("requests", "rate_per_s", "p50_ms", "p99_ms", "mean_ms",
                                           "gradients_per_s", "relL2_max", "relL2_max_by_geometry", "checked")}
            s.update(batches=summ["stats"]["batches"], merged_requests=summ["stats"]["merged_requests"])
            sweep.append(s)
            log(f"serving at {rate}/s: p50 {s['p50_ms']:.1f} ms  p99 {s['p99_ms']:.1f} ms  "
                f"{s['gradients_per_s']:.4g} gradients/s  batches {s['batches']} merged {s['merged_requests']}  "
                f"relL2 max {s['relL2_max']}")
        serving = dict(sweep[0])
        serving.update(slots=3, fields_per_request="1..32", geometries={G["name"]: G["n"] for G in geos},
                       settings={G["name"]: [G["degree"], G["thresh"]] for G in geos},
                       store=int(opts.store),
                       latency="scheduled arrival to results in pinned host memory (PCIe both ways)",
                       sweep=sweep)

    out = {
        "metric": "2.3M-pt SAL gradients/s @relL2<=1e-6, batched vs plain GPU; p99 latency",
        "value": value, "unit": "gradients/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic (seeded grid and harmonic fields)",
        "config": {"workload": f"{args.workload} N={n}, {R} field(s), {args.mode}", "n_points": n, "fields": R,
                   "degree": deg, "theta": THETA, "cluster_thresh": thresh, "panels": P,
                   "interactions": g.n_interactions, "pair_evals": [int(k) for k in K],
                   "plan": plan_info,
                   "tree_build_s": round(tree_s, 4), "dtt_s": round(dtt_s, 4), "plan_s": round(plan_s, 3),
                   "step_ms_p50": round(step_ms[len(step_ms) // 2], 3), "step_ms_max": round(step_ms[-1], 3),
                   "graph": bool(args.graph), "m2l": int(opts.m2l), "rhs_tile": int(opts.rhs_tile),
                   "store": int(opts.store), "device_bytes_resident": dev_bytes,
                   "relL2_vs_direct": rel, "direct_mode": direct_mode, "targets": int(len(tg))},
        "roofline": roof, "cpu_baseline": cpu, "serving": serving, "autotune": tuned, "geometry_cache": cache,
        "kernels_ms_per_launch": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
