#!/usr/bin/env python3
"""Benchmark: SAL gradients/s on the 2,313,486-point MPAS-shaped grid, batched against plain FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode batched|plain|stt] [--fields R]
                    [--workload mpas|mom6|fibonacci] [--no-graph] [--calib-copy-mb MB]

One step is one evaluation of the treecode over all N points for R fields, with the grid's tree,
interaction list and interpolation data resident in HBM and the fields already on the device
(tree, list and plan build times are reported separately in the JSON line).

  batched  (default) R = 32 fields, mixed precision (FP32 kernel values, FP32 MFMA near-field products
           with FP64 per-panel sums, close near pairs in FP64, two- or three-part split-bf16 far-field
           products), accuracy mode, at the autotuner's setting, replayed as one captured HIP graph
           per step -> the headline.  The same run also times, on the same grid:
             plain_fp64    the reference's semantics in FP64 at d = 8, leaves <= 128, one field (the
                           "plain GPU" side of the metric), with a full-size parity sample against
                           the compiled reference (its list bit for bit, its sums on every 100th entry)
             cpu_baseline  the compiled reference's fast_sum on the host cores at the headline's
                           degree and leaf size
             serving       a request loop over three resident geometries at swept arrival rates
  plain    R fields evaluated one by one in FP64 with the reference's semantics
  stt      the MOM6 driver's FP64 path (single tree traversal + per-point sums)

Accuracy is measured against the FP64 direct sum at seeded sampled targets, per field.  The one JSON
line goes to stdout, last; a few progress lines go to stderr; the autotune table, the serving sweep
and the per-request records go to gpurun_out/bench_detail.json.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sal-fort_amd"))

THETA = 0.7
# MI355X published dense peaks, TFLOP/s; FP32 vector and FP32 MFMA share the 157.3 figure
PEAKS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
HBM_PEAK_GBS = 8000.0
PLAIN_DEGREE, PLAIN_THRESH = 8, 129    # the plain FP64 GPU treecode's configuration (d = 8, leaves <= 128)
PARITY_STRIDE = 100                    # reference parity sample: every 100th interaction


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


def kernel_flops(kernel, K, fields, close_pairs=0, prof=None):
    """Algorithmic flops of one launch of `kernel` over `fields` fields: 34 + 4 R per pair evaluation
    (38 at R = 1), pairs K = [PP c_t c_s, PC c_t m, CP m c_s, CC m^2] summed over the list; the close
    pairs of the mixed path are summed by close_mix, not near_mix.  In the FP64 reference path pair64
    sums the PP entries (each mirrored pair of entries evaluates a kernel factor once for both, the
    list's PP pairs still count once each); farpair64 sums mirrored CC entries and CP entries with
    their PC mirrors, leaving leaf64 and proxy64 a mix (no per-kernel figure: fp64_pairs_frac covers
    the four together)."""
    per_pair = 34 + 4 * fields
    if kernel == "pair64":
        return K[0] * per_pair
    if prof is not None and "farpair64" in prof and kernel in ("leaf64", "proxy64", "farpair64"):
        return None
    if kernel == "leaf64" and prof is not None and "pair64" in prof:
        return K[1] * per_pair
    if kernel in ("leaf64", "near_mix"):   # target points: PP and PC entries
        return (K[0] + K[1] - close_pairs) * per_pair
    if kernel in ("proxy64", "far_mix"):   # target proxies: CP and CC entries
        return (K[2] + K[3]) * per_pair
    if kernel == "close_mix":
        return close_pairs * per_pair
    return None


def dominant(prof):
    """The pair-evaluation kernel with the most time (the roofline's kernel)."""
    pair = [k for k in prof if kernel_flops(k, [1, 1, 1, 1], 1, prof=prof) is not None]
    return max(pair or prof, key=lambda k: prof[k][0])


def read_traffic(kernel, tag):
    """HBM bytes per launch of `kernel` measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by
    the same counters on a copy of known size), kept as a text summary under profiles/."""
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


def grid_npz(td, W, ssh, thresh):
    import numpy as np
    path = os.path.join(td, f"grid_t{thresh}.npz")
    np.savez(path, x=W["x"], y=W["y"], z=W["z"], area=W["area"], ssh=ssh, thresh=thresh)
    return path


def cpu_baseline(inp, n, degree, thresh, stride, cores):
    """The compiled reference's fast_sum as `cores` MPI ranks would run it, one worker process per
    host core (oracle/ref_sample.py), over every stride-th interaction of its own list: worker r
    takes entries i with i mod (stride cores) == r stride, so together they take i mod stride == 0
    (the round-robin split of fast_sum_module.f90:304).  Wall time from the common start to the last
    worker's end, times stride: gradients/s for one field."""
    worker = os.path.join(ROOT, "oracle", "ref_sample.py")
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
    return {"value": n / (dt * stride), "unit": "gradients/s", "cores": cores, "kind": "reference",
            "sample": f"compiled reference fast_sum, d={degree}, cluster_thresh={thresh}, {cores} rank processes "
                      f"over every {stride}th list entry, one field, wall x {stride}",
            "stride": stride, "wall_s": round(dt, 3), "cpu_s": round(sum(own), 2)}


def relerr(S, torch, dx, dy, dz, da, ds, gx, gy, tg, mode):
    """Per-field relL2 of (gx, gy) [R, n] against the FP64 direct sum at targets tg."""
    ex, ey = S.direct_sum(dx, dy, dz, da, ds, tg, mode)
    if ex.dim() == 1:
        ex, ey = ex[None], ey[None]
    tx, ty = gx[:, tg.long()], gy[:, tg.long()]
    return torch.sqrt(((tx - ex) ** 2 + (ty - ey) ** 2).sum(dim=1) / (ex ** 2 + ey ** 2).sum(dim=1)).tolist()


def timed_steps(S, torch, g, da, ds, gx, gy, opts, steps, prof):
    """K evaluations bracketed by a synchronize on both sides: (ms per step, sorted per-step ms from
    events on the evaluation stream, per-kernel {name: (total ms, launches)} when prof)."""
    S.check(S.lib().sal_prof_reset())
    S.check(S.lib().sal_prof_enable(1 if prof else 0))
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record()
    for i in range(steps):
        g.eval(da, ds, gx, gy, opts)
        ev[i + 1].record()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    S.check(S.lib().sal_prof_enable(0))
    return ms, sorted(ev[i].elapsed_time(ev[i + 1]) for i in range(steps)), prof_read(S) if prof else {}


def kms(prof):
    return {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items()}


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
        g.mom6_sums(area, ssh, deg)
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    S.check(S.lib().sal_prof_enable(0))
    km = kms(prof_read(S))
    flops = 38 * (K[0] + K[1])
    ach = flops / (km["mom6_sums"] * 1e-3) / 1e12
    traffic, tsrc = read_traffic("mom6_sums", f"{args.workload}_{n}_r1_d{deg}_t{thresh}_stt")
    log(f"stt: {npc} pc + {npp} pp entries ({stt_s:.3f} s); sums {ms:.2f} ms per field")
    out = {"metric": "MOM6 FP64 gradients/s (single tree traversal path, reference semantics)",
           "value": n / (ms * 1e-3), "unit": "gradients/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded grid and harmonic fields)",
           "config": {"workload": f"{args.workload} N={n}, 1 field, stt", "n_points": n, "degree": deg,
                      "theta": THETA, "cluster_thresh": thresh, "pc_entries": int(npc), "pp_entries": int(npp),
                      "pair_evals": K, "tree_build_s": round(tree_s, 4), "stt_s": round(stt_s, 4),
                      "timing": "host arrays in and out (the driver's call), PCIe included"},
           "roofline": {"bound": "mfma", "kernel": "mom6_sums", "engine": "FP64 VALU", "achieved": round(ach, 3),
                        "peak": PEAKS["fp64"], "unit": "TFLOP/s", "frac": round(ach / PEAKS["fp64"], 4),
                        "traffic": traffic, "traffic_source": tsrc, "flops_per_launch": flops,
                        "avg_launch_ms": km["mom6_sums"]},
           "kernels_ms_per_launch": km}
    print(json.dumps(out), flush=True)


def plain_leg(S, torch, W, dx, dy, dz, da, ssh0, steps, warmup, parity):
    """The plain FP64 GPU treecode on the same grid at d = 8, leaves <= 128, one field, reference
    semantics: gradients/s, the dominant kernel's roofline, and the parity sample against the
    compiled reference (`parity`: its list and its sums over every PARITY_STRIDE-th entry)."""
    import numpy as np
    n = dx.numel()
    t = time.perf_counter()
    g = S.Geometry(device_xyz=(dx, dy, dz), cluster_thresh=PLAIN_THRESH)
    g.dtt(THETA)
    build_s = time.perf_counter() - t
    o = S.default_opts(degree=PLAIN_DEGREE, mode=S.MODE_REFERENCE, precision=S.PREC_FP64)
    gx = torch.zeros(n, dtype=torch.float64, device=dx.device)
    gy = torch.zeros_like(gx)
    for _ in range(max(1, warmup)):
        g.eval(da, ssh0, gx, gy, o)
    ms, step_ms, prof = timed_steps(S, torch, g, da, ssh0, gx, gy, o, steps, True)
    K = g.work_counts(PLAIN_DEGREE)
    km = kms(prof)
    dom = dominant(prof)
    fl = kernel_flops(dom, K, 1, prof=prof)
    ach = fl / (km[dom] * 1e-3) / 1e12
    # every pair of the list (PP, PC, CP, CC) over the four pair kernels' time
    pt_ms = sum(km.get(k, 0.0) for k in ("pair64", "farpair64", "leaf64", "proxy64"))
    pt_frac = sum(K) * 38 / (pt_ms * 1e-3) / 1e12 / PEAKS["fp64"] if pt_ms > 0 else None
    traffic, tsrc = read_traffic(dom, f"mpas{n}_r1_d{PLAIN_DEGREE}_t{PLAIN_THRESH}_plain")
    out = {"value": n / (ms * 1e-3), "unit": "gradients/s", "ms_per_step": round(ms, 3),
           "step_ms_p50": round(step_ms[len(step_ms) // 2], 3), "degree": PLAIN_DEGREE,
           "cluster_thresh": PLAIN_THRESH, "fields": 1, "dtype": "f64", "panels": g.counts()[0],
           "interactions": g.n_interactions, "tree_dtt_s": round(build_s, 3),
           "roofline": {"bound": "mfma", "kernel": dom, "engine": "FP64 VALU", "achieved": round(ach, 2),
                        "peak": PEAKS["fp64"], "unit": "TFLOP/s", "frac": round(ach / PEAKS["fp64"], 4),
                        "traffic": traffic, "traffic_source": tsrc, "flops_per_launch": int(fl),
                        "avg_launch_ms": km[dom]},
           "fp64_pairs_frac": round(pt_frac, 4) if pt_frac else None,
           "kernels_ms_per_launch": km}
    if parity is not None:
        proc, path = parity
        line = proc.stdout.readline().split()
        proc.wait()
        if line and line[0] == "done" and os.path.exists(path):
            ref = np.load(path)
            it, is_, ty = g.dtt_export()
            same = bool(np.array_equal(it, ref["it"]) and np.array_equal(is_, ref["is_"]) and np.array_equal(ty, ref["ty"]))
            sel = (np.arange(len(it)) + 1) % PARITY_STRIDE == 0    # the reference's rank 0 of PARITY_STRIDE
            h = S.Geometry(device_xyz=(dx, dy, dz), cluster_thresh=PLAIN_THRESH)
            h.dtt_import(it[sel], is_[sel], ty[sel])
            hx = torch.zeros_like(gx)
            hy = torch.zeros_like(gx)
            h.eval(da, ssh0, hx, hy, o)
            torch.cuda.synchronize()
            hx, hy = hx.cpu().numpy(), hy.cpu().numpy()
            rel = float(np.sqrt(((hx - ref["gx"]) ** 2 + (hy - ref["gy"]) ** 2).sum() / (ref["gx"] ** 2 + ref["gy"] ** 2).sum()))
            out["parity"] = {"list_bit_exact": same, "entries": int(sel.sum()), "stride": PARITY_STRIDE,
                             "types_sampled": [int((ty[sel] == k).sum()) for k in range(4)],
                             "relL2_vs_reference": rel, "reference_s": float(line[1])}
            h.close()
        else:
            out["parity"] = {"error": "reference worker failed"}
    g.close()
    return out


def serving_leg(S, serve, deg, thresh, m2l, near64, store, n_requests, rates, detail):
    """Three resident geometries, seeded Poisson arrivals at each rate: p50 / p99 latency, gradients/s and
    the worst sampled error per geometry (all rates)."""
    geos = serve.build(["fibonacci", "mom6", "mpas"], degree=deg, thresh=thresh, store=store, m2l=m2l, near64=near64,
                       quiet=True)
    sweep = []
    for k, rate in enumerate(rates):
        for G in geos:
            G["free"] = list(range(len(G["outs"])))
        srv = S.Server(slots=3, batching=True)
        serve.register(srv, geos)
        recs, summ = serve.run_load(srv, geos, rate, n_requests, seed=k, check_every=3)
        srv.close()
        s = {key: summ[key] for key in ("requests", "rate_per_s", "p50_ms", "p99_ms", "mean_ms", "gradients_per_s",
                                       "relL2_max", "relL2_max_by_geometry", "checked", "overlapped_requests",
                                       "running_mean")}
        s.update(batches=summ["stats"]["batches"], merged_requests=summ["stats"]["merged_requests"])
        sweep.append(s)
        detail.setdefault("serving_requests", []).append(
            [{k2: v for k2, v in r.items() if k2 in ("geo", "R", "batch", "batch_fields", "relL2", "t_submit",
                                                     "t_done", "late")} for r in recs])
        log(f"serving {rate}/s: p50 {s['p50_ms']:.1f} ms  p99 {s['p99_ms']:.1f} ms  {s['gradients_per_s']:.3g} "
            f"gradients/s  merged {s['merged_requests']}  overlapped {s['overlapped_requests']}  relL2 max "
            f"{s['relL2_max']:.2e}")
    detail["serving_sweep"] = sweep
    worst = {}
    for s in sweep:
        for name, v in s["relL2_max_by_geometry"].items():
            worst[name] = max(worst.get(name, 0.0), v)
    out = {"rates_per_s": rates, "requests_per_rate": n_requests,
           "p50_ms": [round(s["p50_ms"], 1) for s in sweep], "p99_ms": [round(s["p99_ms"], 1) for s in sweep],
           "gradients_per_s": [float("%.4g" % s["gradients_per_s"]) for s in sweep],
           "merged_requests": [s["merged_requests"] for s in sweep],
           "overlapped_requests": [s["overlapped_requests"] for s in sweep],
           "running_mean": [round(s["running_mean"], 2) for s in sweep],
           "relL2_max": max(worst.values()) if worst else None,
           "relL2_max_by_geometry": {k: float("%.3g" % v) for k, v in worst.items()},
           "checked": sum(s["checked"] for s in sweep), "slots": 3, "fields_per_request": "1..32",
           "settings": {G["name"]: [G["n"], G["degree"], G["thresh"], G["m2l"], G["near64"]] for G in geos}}
    for G in geos:
        G["g"].close()
    return out


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
    ap.add_argument("--m2l", type=int, default=None)
    ap.add_argument("--near64", type=int, default=None)
    ap.add_argument("--rhs-tile", type=int, default=0)
    ap.add_argument("--store", type=int, default=None,
                    help="batched: stream a stored operator (1) or evaluate kernel values on every call (0)")
    ap.add_argument("--no-tune", action="store_true",
                    help="batched: skip the autotuner and use the workload's degree and leaf size")
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="batched: launch the kernels directly instead of replaying a captured HIP graph")
    ap.add_argument("--targets", type=int, default=2000, help="sampled targets for the direct-sum check")
    ap.add_argument("--cpu-stride", type=int, default=15, help="reference CPU sample: every k-th interaction")
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="reference CPU sample: worker processes, one per core (0: the host cores, at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="batched: skip the plain FP64 leg")
    ap.add_argument("--serve-requests", type=int, default=45,
                    help="requests per rate of the serving leg (batched mpas only; 0 skips it)")
    ap.add_argument("--serve-rate", default="15,40,120", help="mean arrivals per second of the serving leg (a sweep)")
    ap.add_argument("--no-cache", dest="cache", action="store_false",
                    help="skip the geometry cache leg (save and reload the tree and lists)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="run one device copy of this size first (HBM counter calibration)")
    ap.add_argument("--detail", default=os.path.join(ROOT, "gpurun_out", "bench_detail.json"))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import salfort as S

    dev = torch.device("cuda:0")
    detail = {}
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
    W = S.workload(args.workload, n=args.n, seed=0, R=R)
    thresh = args.thresh or W["thresh"]
    deg = args.degree or W["degree"]
    n = len(W["x"])
    ssh = W["ssh"].reshape(R, n)
    full = args.mode == "batched" and args.workload == "mpas" and args.n is None
    tens = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    dx, dy, dz, da, ds = (tens(v) for v in (W["x"], W["y"], W["z"], W["area"], ssh))
    torch.cuda.synchronize()
    td = tempfile.TemporaryDirectory()

    # the reference side of the plain leg's parity sample runs on the host meanwhile
    parity = None
    if full and not args.no_plain:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle_py import REF_SO
        if os.path.exists(REF_SO):
            inp = grid_npz(td.name, W, ssh[0], PLAIN_THRESH)
            out_path = os.path.join(td.name, "ref_parity.npz")
            proc = subprocess.Popen([sys.executable, os.path.join(ROOT, "oracle", "ref_sample.py"), "--parity", inp,
                                     str(PARITY_STRIDE), str(PLAIN_DEGREE), str(THETA), out_path],
                                    stdout=subprocess.PIPE, text=True)
            parity = (proc, out_path)

    m2l = args.m2l if args.m2l is not None else S.M2L_BF16X3
    near64 = args.near64
    tuned = None
    if args.mode == "batched" and not args.no_tune and args.degree is None and args.thresh is None:
        import autotune
        t = time.perf_counter()
        rows, best = autotune.tune(dx, dy, dz, da, ds, quiet=True)
        tuned = {"best": best, "seconds": round(time.perf_counter() - t, 1), "settings_tried": len(rows)}
        detail["autotune_table"] = rows
        if best is not None:
            deg, thresh, m2l, near64 = best["degree"], best["cluster_thresh"], best["m2l"], best["near64"]
            args.rhs_tile = args.rhs_tile or best["rhs_tile"]
            if args.store is None:
                args.store = best["store"]
        log(f"autotune: {len(rows)} settings in {tuned['seconds']} s -> d={deg} cluster_thresh={thresh} m2l={m2l} "
            f"near64={near64} store={args.store}")

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

    if args.mode == "batched":
        opts = S.default_opts(degree=deg, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, m2l=m2l,
                              store=1 if args.store is None else int(args.store))
        if near64 is not None:
            opts.near64 = int(near64)
        direct_mode, dtype = 2, "fp32"
    else:
        opts = S.default_opts(degree=deg, mode=S.MODE_REFERENCE, precision=S.PREC_FP64)
        direct_mode, dtype = 0, "f64"
    opts.rhs_tile = args.rhs_tile
    gx = torch.zeros((R, n), dtype=torch.float64, device=dev)
    gy = torch.zeros_like(gx)
    t = time.perf_counter()
    g.eval(da, ds, gx, gy, opts)  # first call builds the evaluation plan (and the stored operator)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t

    # timed region: K steps (captured HIP graph replays in the batched mode), then K direct-launch steps
    # with per-kernel HIP events on the evaluation stream for the kernel times and the roofline
    graph = args.mode == "batched" and args.graph
    og = S.default_opts(**{f: getattr(opts, f) for f, _ in S.EvalOpts._fields_})
    og.graph = 1 if graph else 0
    for _ in range(max(args.warmup, 1 if graph else 0)):  # the first graph call captures
        g.eval(da, ds, gx, gy, og)
    ms_step, step_ms, prof = timed_steps(S, torch, g, da, ds, gx, gy, og, args.steps, not graph)
    ms_direct = ms_step
    if graph:
        ms_direct, _, prof = timed_steps(S, torch, g, da, ds, gx, gy, opts, args.steps, True)
    value = n * R / (ms_step * 1e-3)
    log(f"eval: {ms_step:.2f} ms/step ({'graph' if graph else 'direct'}), {value:.4g} gradients/s; direct "
        f"{ms_direct:.2f} ms; plan {plan_s:.2f} s")

    # accuracy against the FP64 direct sum at seeded targets, every field
    rng = np.random.default_rng(1234)
    tg = torch.as_tensor(np.sort(rng.choice(n, min(args.targets, n), replace=False)).astype(np.int32), device=dev)
    rel = relerr(S, torch, dx, dy, dz, da, ds, gx, gy, tg, direct_mode)

    # roofline of the dominant kernel: algorithmic pair-evaluation flops per launch / its launch time
    K = g.work_counts(deg)
    plan_info = g.plan_stats(opts) if args.mode == "batched" else None
    close = plan_info["close_pairs"] if plan_info else 0
    dev_bytes = g.device_bytes()
    km = kms(prof)
    dom = dominant(prof)
    per_launch = 1 if args.mode == "plain" else min(R, opts.rhs_tile or (32 if R > 16 else 16))
    flops = kernel_flops(dom, K, per_launch, close, prof=prof)
    tag = f"{args.workload}{n}_r{R}_d{deg}_t{thresh}_{args.mode}" + ("_stored" if opts.store else "")
    traffic, tsrc = read_traffic(dom, tag)
    peak = PEAKS["fp64" if dtype == "f64" else "fp32"]
    ach = flops / (km[dom] * 1e-3) / 1e12
    roof = {"bound": "mfma", "kernel": dom,
            "engine": "FP64 VALU" if dtype == "f64" else "FP32 MFMA products (stored FP32 kernel values)" if opts.store
            else "FP32 VALU kernel values + FP32 MFMA products",
            "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s", "frac": round(ach / peak, 4),
            "traffic": traffic, "traffic_source": tsrc, "flops_per_launch": int(flops), "avg_launch_ms": km[dom],
            "timing": "HIP events, K direct steps after the graph replays" if graph else "HIP events, timed steps"}
    if traffic:
        roof["hbm_GBps"] = round(traffic / (km[dom] * 1e-3) / 1e9, 1)
    if plan_info and opts.store and dom == "near_mix":
        roof["stored_operator_bytes"] = plan_info["stored_near_bytes"]

    config = {"workload": f"{args.workload} N={n}, {R} field(s), {args.mode}", "n_points": n, "fields": R,
              "degree": deg, "theta": THETA, "cluster_thresh": thresh, "panels": P, "interactions": g.n_interactions,
              "pair_evals": [int(k) for k in K], "tree_build_s": round(tree_s, 4), "dtt_s": round(dtt_s, 4),
              "plan_s": round(plan_s, 3), "graph": graph, "ms_per_step_direct": round(ms_direct, 3),
              "step_ms_p50": round(step_ms[len(step_ms) // 2], 3), "step_ms_max": round(step_ms[-1], 3),
              "device_bytes_resident": dev_bytes,
              "relL2_vs_direct": {"fields": len(rel), "targets": int(len(tg)), "max": float("%.3g" % max(rel)),
                                  "median": float("%.3g" % float(np.median(rel))), "direct_mode": direct_mode}}
    if args.mode == "batched":
        config.update(m2l=int(opts.m2l), near64=int(opts.near64), rhs_tile=int(opts.rhs_tile), store=int(opts.store),
                      close_pairs=close, stored_operator_bytes=(plan_info["stored_near_bytes"] +
                                                                plan_info["stored_far_bytes"])
                      if opts.store else 0)
        if opts.store:  # the same setting with kernel values evaluated on every call
            of = S.default_opts(**{f: getattr(opts, f) for f, _ in S.EvalOpts._fields_})
            of.store = 0
            g.eval(da, ds, gx, gy, of)
            ms_fly, _, _ = timed_steps(S, torch, g, da, ds, gx, gy, of, max(2, args.steps // 4), False)
            config["on_the_fly_ms_per_step"] = round(ms_fly, 3)
    detail["headline"] = {"config": config, "relL2_per_field": rel, "plan": plan_info, "kernels_ms_per_launch": km}
    g.close()
    del gx, gy
    torch.cuda.empty_cache()

    plain = None
    if full and not args.no_plain:
        plain = plain_leg(S, torch, W, dx, dy, dz, da, ds[0].contiguous(), args.steps, args.warmup, parity)
        plain["batched_over_plain"] = round(value / plain["value"], 2)
        detail["plain_fp64"] = plain
        pr = plain.get("parity", {})
        log(f"plain FP64: {plain['ms_per_step']:.2f} ms/step, {plain['value']:.4g} gradients/s "
            f"(batched {plain['batched_over_plain']}x); parity lists {pr.get('list_bit_exact')} relL2 "
            f"{pr.get('relL2_vs_reference')}")
        plain.pop("kernels_ms_per_launch", None)

    # on-disk geometry cache: save and reload the tree and lists, against building them
    cache = None
    if args.cache and args.mode == "batched":
        g2 = S.Geometry(device_xyz=(dx, dy, dz), cluster_thresh=thresh)
        g2.dtt(THETA)
        path = os.path.join(td.name, "grid.salgeom")
        t = time.perf_counter()
        g2.save(path)
        save_s = time.perf_counter() - t
        t = time.perf_counter()
        h = S.Geometry.load(path)
        load_s = time.perf_counter() - t
        same = all(np.array_equal(a, b) for a, b in zip(g2.dtt_export(), h.dtt_export()))
        cache = {"file_bytes": os.path.getsize(path), "save_s": round(save_s, 3), "load_s": round(load_s, 3),
                 "build_tree_dtt_s": round(tree_s + dtt_s, 3), "lists_identical": bool(same)}
        h.close()
        g2.close()
        os.remove(path)

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle_py import REF_SO
        if os.path.exists(REF_SO):
            inp = grid_npz(td.name, W, ssh[0], thresh)
            cpu = cpu_baseline(inp, n, deg, thresh, args.cpu_stride, args.cpu_cores or cpu_cores())
            log(f"cpu baseline: {cpu['value']:.4g} gradients/s on {cpu['cores']} cores")

    serving = None
    if args.serve_requests > 0 and full:
        import serve
        rates = [float(r) for r in args.serve_rate.split(",")]
        serving = serving_leg(S, serve, deg, thresh, m2l, opts.near64, int(opts.store), args.serve_requests, rates,
                              detail)
    td.cleanup()

    try:
        os.makedirs(os.path.dirname(args.detail), exist_ok=True)
        with open(args.detail, "w") as f:
            json.dump(detail, f)
    except OSError:
        pass

    out = {
        "metric": "2.3M-pt SAL gradients/s @relL2<=1e-6, batched vs plain GPU; p99 latency",
        "value": value, "unit": "gradients/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic (seeded grid and harmonic fields)", "config": config,
        "roofline": roof, "cpu_baseline": cpu, "plain_fp64": plain, "serving": serving,
        "autotune": ({"best": {k: (float("%.4g" % v) if isinstance(v, float) else v) for k, v in tuned["best"].items()
                               if k in ("degree", "cluster_thresh", "m2l", "near64", "rhs_tile", "store", "ms",
                                        "relL2_max", "bar")},
                      "settings_tried": tuned["settings_tried"], "seconds": tuned["seconds"]}
                     if tuned and tuned["best"] else None),
        "geometry_cache": cache, "kernels_ms_per_launch": km,
        "detail_file": os.path.relpath(args.detail, ROOT),
    }
    sys.stderr.flush()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
