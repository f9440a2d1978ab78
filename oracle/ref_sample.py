"""Test infrastructure: one worker of the reference CPU baseline (the bench's cpu_baseline leg), and
the reference side of the bench's full-size parity sample.

    python oracle/ref_sample.py INPUTS.npz RANK NPROC DEGREE THETA
    python oracle/ref_sample.py --parity INPUTS.npz STRIDE DEGREE THETA OUT.npz

Loads the grid from INPUTS.npz (x, y, z, area, ssh, thresh), builds the compiled reference's tree
and interaction list (tree_traverse, DUAL_TREE_TRAVERSAL), prints "ready", waits for one line on
stdin, then runs the reference's fast_sum as MPI rank RANK of NPROC would (entries i with
i mod NPROC == RANK, fast_sum_module.f90:304) and prints "done <seconds>".  The parent runs several
workers at once, one per host core, each with its own rank, as the reference's MPI ranks run.

With --parity the worker runs at once: the reference's fast_sum as rank 0 of STRIDE ranks (every
STRIDE-th entry, i = STRIDE, 2 STRIDE, ... in the reference's 1-based numbering), and writes its
sums and the reference's whole interaction list (0-based) to OUT.npz, then prints "done <seconds>".
"""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from oracle_py import Reference  # noqa: E402


def parity(path, stride, degree, theta, out_path):
    d = np.load(path, allow_pickle=False)
    x, y, z, area, ssh, thresh = d["x"], d["y"], d["z"], d["area"], d["ssh"], int(d["thresh"])
    R = Reference()
    out = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        t = time.perf_counter()
        R.tree_build(x, y, z, thresh)
        it, is_, ty = R.dtt(theta, thresh)
        gx, gy = R.fast_sum(x, y, z, area, ssh, degree, rank=0, nproc=stride)
        dt = time.perf_counter() - t
        R.flush()
    finally:
        os.dup2(out, 1)
        os.close(devnull)
    np.savez(out_path, gx=gx, gy=gy, it=it, is_=is_, ty=ty)
    print(f"done {dt:.6f}", flush=True)


def main():
    if sys.argv[1] == "--parity":
        parity(sys.argv[2], int(sys.argv[3]), int(sys.argv[4]), float(sys.argv[5]), sys.argv[6])
        return
    path, rank, nproc, degree, theta = sys.argv[1], int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4]), float(sys.argv[5])
    d = np.load(path, allow_pickle=False)
    x, y, z, area, ssh, thresh = d["x"], d["y"], d["z"], d["area"], d["ssh"], int(d["thresh"])
    R = Reference()
    # the reference prints on stdout; keep stdout for this worker's own two lines
    out = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        R.tree_build(x, y, z, thresh)
        R.dtt(theta, thresh)
        R.flush()
        os.write(out, b"ready\n")
        sys.stdin.readline()
        t = time.perf_counter()
        R.fast_sum(x, y, z, area, ssh, degree, rank=rank, nproc=nproc)
        dt = time.perf_counter() - t
        R.flush()
    finally:
        os.dup2(out, 1)
        os.close(devnull)
    print(f"done {dt:.6f}", flush=True)


if __name__ == "__main__":
    main()
