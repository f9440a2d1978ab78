"""Write the golden fixtures of the parity tests from the compiled reference.

Runs in the build container only (it needs the reference tree, built into oracle/_ref by
`make -C oracle ref`).  Each fixture stores its inputs next to the reference's outputs, so the tests
never need the reference or the input generator to reproduce them:

    python tests/golden/make_golden.py

Cases
  fib642     Fibonacci N=642, cluster_thresh=10, theta=0.7, degree 3: tree, DTT list, fast_sum,
             assign_points_to_panels, STT lists, proxy_source_compute, pc_compute+pp_compute
  mpas4000   variable-resolution grid N=4000 (the MPAS-shaped generator, seed 7), cluster_thresh=33,
             theta=0.7, degree 4: tree, DTT list, fast_sum
  kernels    sal_grad_gfunc on seeded pairs (including self pairs, near pairs and poles),
             bli_coeffs / bli_interp_points_shift / interp_vals_bli, cube-sphere maps
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "sal-fort_amd"))

from oracle_py import Reference  # noqa: E402


def fib(n):
    i = np.arange(n)
    z = 1 - (2 * i + 1) / n
    ph = i * np.pi * (3 - np.sqrt(5))
    r = np.sqrt(1 - z * z)
    return r * np.cos(ph), r * np.sin(ph), z


def tree_case(R, x, y, z, area, ssh, thresh, theta, degree, mom6):
    out = dict(x=x, y=y, z=z, area=area, ssh=ssh, thresh=thresh, theta=theta, degree=degree)
    R.tree_build(x, y, z, thresh)
    t = R.tree_export()
    for k, v in t.items():
        out["tree_" + k] = v
    it, is_, ty = R.dtt(theta, thresh)
    out.update(dtt_t=it, dtt_s=is_, dtt_ty=ty)
    gx, gy = R.fast_sum(x, y, z, area, ssh, degree)
    out.update(fs_gx=gx, fs_gy=gy)
    if mom6:
        nlev = int(t["level"].max()) + 1
        pp = R.assign(x, y, z, nlev)
        out["assign"] = pp
        pct, pcs, ppt, pps = R.stt(theta, thresh, x, y, z)
        out.update(stt_pc_t=pct, stt_pc_s=pcs, stt_pp_t=ppt, stt_pp_s=pps)
        w = R.proxy_source(pp, x, y, z, ssh, area, degree)
        out["proxy_w"] = w
        mx, my = R.mom6_sums(x, y, z, area, ssh, degree, w)
        out.update(m6_gx=mx, m6_gy=my)
    return out


def kernel_case(R):
    rng = np.random.default_rng(12345)
    n = 400
    t = rng.normal(size=(n, 3))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    s = rng.normal(size=(n, 3))
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    s[:40] = t[:40]                                   # self pairs
    s[40:80] = t[40:80] + 1e-4 * rng.normal(size=(40, 3))  # near pairs
    s[40:80] /= np.linalg.norm(s[40:80], axis=1, keepdims=True)
    t[80] = (0.0, 0.0, 1.0)                           # poles
    t[81] = (0.0, 0.0, -1.0)
    g = np.array([R.gfunc(tuple(t[i]), tuple(s[i])) for i in range(n)])
    g6 = np.array([R.gfunc(tuple(t[i]), tuple(s[i]), which=1) for i in range(n)])
    out = dict(k_t=t, k_s=s, k_g=g, k_g6=g6)
    # BLI
    degs = [0, 1, 2, 3, 4, 6, 8, 10]
    pts, coefs, vals, args = [], [], [], []
    for d in degs:
        a, b = -0.3 + 0.01 * d, 0.45 - 0.02 * d
        coefs.append(R.bli_coeffs(d))
        p = R.bli_points(a, b, d)
        pts.append(p)
        for xi, eta in [(0.1, -0.2), (p[0], 0.05), (a, b), (0.123456789, 0.2)]:
            args.append((d, xi, eta, a, b, a - 0.1, b + 0.1))
            vals.append(R.interp_vals(xi, eta, a, b, a - 0.1, b + 0.1, d).ravel())
    out.update(bli_degs=np.array(degs), bli_coefs=np.concatenate(coefs), bli_pts=np.concatenate(pts),
               bli_args=np.array(args), bli_vals=np.concatenate(vals))
    # cube-sphere maps
    q = rng.normal(size=(300, 3))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[0] = (1, 1, 0) / np.sqrt(2)                     # face ties
    q[1] = (1, 1, 1) / np.sqrt(3)
    faces = np.array([R.face(*p) for p in q], np.int32)
    xe = np.array([R.xieta(*p) for p in q])
    back = np.array([R.xyz(xe[i, 0], xe[i, 1], int(faces[i])) for i in range(len(q))])
    out.update(cs_p=q, cs_face=faces, cs_xieta=xe, cs_back=back)
    return out


def main():
    R = Reference()
    x, y, z = fib(642)
    area = np.full(642, 4 * np.pi * 6.371e6 ** 2 / 642)
    ssh = 0.1 * np.sin(3 * x) * np.cos(2 * y) + 0.05 * z ** 3 + 0.02 * x * y
    np.savez_compressed(os.path.join(HERE, "fib642.npz"),
                        **tree_case(R, x, y, z, area, ssh, 10, 0.7, 3, mom6=True))

    import salfort as S
    x, y, z, area = S.gen_mpas(4000, seed=7)
    ssh = S.gen_ssh(x, y, z, lmax=16, amp=0.1, noise=0.02, seed=8)
    np.savez_compressed(os.path.join(HERE, "mpas4000.npz"),
                        **tree_case(R, x, y, z, area, ssh, 33, 0.7, 4, mom6=False))

    np.savez_compressed(os.path.join(HERE, "kernels.npz"), **kernel_case(R))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
