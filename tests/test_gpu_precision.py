"""Precision controls of the mixed-precision path on the GPU.

* close pairs (near64): the plan's close-pair list is exactly the set of point-to-point pairs of the
  near field whose FP32 chord squared is below the target leaf's cutoff (near64 h_L)^2, rebuilt here
  on the host from the tree and list exports (pairs within 1e-6 of a cutoff, where the host's
  emulation of the FP32 chord could round the other way, are left out of the comparison);
* close pairs summed in FP64 bring the mixed result closer to the FP64 accuracy mode on the same
  list: at least 2x on an MPAS-shaped grid, whose error is near-field rounding;
* the FP32-MFMA far field with FP64 per-chunk sums, and the three-part split-bf16 one with the same
  sums, are closer to FP64 than the two-part split-bf16 one on a land-masked MOM6-shaped grid, whose
  far field carries much of the gradient, and the three-part one is as close as FP32 MFMA;
* stored and on-the-fly evaluations stay bit-identical with close pairs and with the FP32 far field.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _dev(a, cuda):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=cuda)


def _solid_angle(b):
    """Solid angle of a cube-face panel from its bounds (min_xi, mid, max_xi, min_eta, mid, max_eta)."""
    def F(xi, eta):
        a, c = np.tan(xi), np.tan(eta)
        return np.arctan(a * c / np.sqrt(1 + a * a + c * c))
    return F(b[2], b[5]) - F(b[0], b[5]) - F(b[2], b[3]) + F(b[0], b[3])


def _chord2_f32(sx, sy, sz, tx, ty, tz):
    """fmaf(dx, dx, fmaf(dy, dy, dz * dz)) with d = float32(s - t): each fused step as the exactly
    computed FP64 value rounded once to FP32 (exact products of FP32 values fit in FP64)."""
    f32 = np.float32
    dx, dy, dz = f32(sx - tx), f32(sy - ty), f32(sz - tz)
    p = (dz * dz).astype(f32)
    p = f32(dy.astype(np.float64) * dy + p)
    return f32(dx.astype(np.float64) * dx + p)


@pytest.mark.parametrize("near64", [2, 3])
def test_close_pair_list_is_the_brute_force_set(cuda, near64):
    import salfort as S
    x, y, z, area = S.gen_mpas(20000, seed=6)
    g = S.Geometry(x, y, z, 65)
    g.dtt(0.7)
    o = S.default_opts(degree=7, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, near64=near64)
    ct, cs, cut2 = g.close_pairs(o)
    T = g.export()
    it, is_, ty = g.dtt_export()
    pp = {}
    for a, b in zip(it[ty == 0], is_[ty == 0]):
        pp.setdefault(int(a), []).append(int(b))
    pts = lambda P: T["points"][T["offsets"][P]:T["offsets"][P] + T["count"][P]]
    leaves = np.where((T["is_leaf"] == 1) & (T["count"] > 0))[0]
    n = len(x)
    want, ambiguous = [], []
    for L in leaves:
        c_host = np.float32(near64 * near64 * _solid_angle(T["bounds"][L]) / T["count"][L])
        assert abs(float(cut2[L]) - float(c_host)) <= 2 * np.spacing(c_host), (L, cut2[L], c_host)
        src = []
        A = L
        while A >= 0:
            src += [pts(S_) for S_ in pp.get(int(A), [])]
            A = T["parent"][A]
        src = np.concatenate(src)
        for t in pts(L):
            c2 = _chord2_f32(x[src], y[src], z[src], x[t], y[t], z[t])
            close = (c2 < cut2[L]) & (src != t)
            near_cut = np.abs(c2.astype(np.float64) - cut2[L]) <= 1e-6 * cut2[L]
            want.append(t * np.int64(n) + src[close & ~near_cut])
            ambiguous.append(t * np.int64(n) + src[near_cut])
    want = np.concatenate(want)
    amb = np.concatenate(ambiguous)
    got = ct.astype(np.int64) * n + cs
    got = got[~np.isin(got, amb)]
    assert len(np.unique(got)) == len(got)   # no pair twice
    assert np.array_equal(np.sort(got), np.sort(want))
    assert len(want) > 5 * n                  # a few pairs per target at this cutoff
    # non-leaf panels have no cutoff
    assert np.all(cut2[T["is_leaf"] == 0] < 0)


def test_close_pairs_bring_mixed_closer_to_fp64(cuda):
    import torch
    import salfort as S
    x, y, z, area = S.gen_mpas(200000, seed=2)
    R = 4
    ssh = S.gen_ssh(x, y, z, lmax=64, amp=0.1, noise=0.02, seed=9, R=R)
    dx, dy, dz, da, ds = (_dev(v, cuda) for v in (x, y, z, area, ssh))
    g = S.Geometry(device_xyz=(dx, dy, dz), cluster_thresh=65)
    g.dtt(0.7)
    res = {}
    for name, o in [("f64", S.default_opts(degree=7, mode=S.MODE_ACCURATE, precision=S.PREC_FP64)),
                    ("mix0", S.default_opts(degree=7, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, near64=0)),
                    ("mix2", S.default_opts(degree=7, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, near64=2))]:
        gx = torch.zeros((R, len(x)), dtype=torch.float64, device=cuda)
        gy = torch.zeros_like(gx)
        g.eval(da, ds, gx, gy, o)
        res[name] = (gx.cpu().numpy(), gy.cpu().numpy())
    for r in range(R):
        e0 = rel_l2(res["mix0"][0][r], res["mix0"][1][r], res["f64"][0][r], res["f64"][1][r])
        e2 = rel_l2(res["mix2"][0][r], res["mix2"][1][r], res["f64"][0][r], res["f64"][1][r])
        assert e2 <= 0.5 * e0 and e2 < 1e-7, (r, e0, e2)


def test_fp32_far_field_with_fp64_sums_on_a_masked_grid(cuda):
    import torch
    import salfort as S
    x, y, z, area = S.gen_mom6(nlon=720, nlat=540)
    R = 4
    ssh = S.gen_ssh(x, y, z, lmax=64, amp=0.1, noise=0.02, seed=1, R=R)
    dx, dy, dz, da, ds = (_dev(v, cuda) for v in (x, y, z, area, ssh))
    g = S.Geometry(device_xyz=(dx, dy, dz), cluster_thresh=129)
    g.dtt(0.7)
    res = {}
    for name, o in [("f64", S.default_opts(degree=8, mode=S.MODE_ACCURATE, precision=S.PREC_FP64)),
                    ("bf16", S.default_opts(degree=8, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, m2l=S.M2L_BF16X3)),
                    ("fp32", S.default_opts(degree=8, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED,
                                            m2l=S.M2L_FP32_MFMA)),
                    ("fp32s", S.default_opts(degree=8, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED,
                                             m2l=S.M2L_FP32_MFMA, store=1)),
                    ("x6", S.default_opts(degree=8, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, m2l=S.M2L_BF16X6)),
                    ("x6s", S.default_opts(degree=8, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, m2l=S.M2L_BF16X6,
                                           store=1))]:
        gx = torch.zeros((R, len(x)), dtype=torch.float64, device=cuda)
        gy = torch.zeros_like(gx)
        g.eval(da, ds, gx, gy, o)
        res[name] = (gx.cpu().numpy(), gy.cpu().numpy())
    # the stored near operator with the FP32 far field: the same bits as on the fly
    assert np.array_equal(res["fp32"][0], res["fp32s"][0]) and np.array_equal(res["fp32"][1], res["fp32s"][1])
    assert np.array_equal(res["x6"][0], res["x6s"][0]) and np.array_equal(res["x6"][1], res["x6s"][1])
    eb = [rel_l2(res["bf16"][0][r], res["bf16"][1][r], res["f64"][0][r], res["f64"][1][r]) for r in range(R)]
    ef = [rel_l2(res["fp32"][0][r], res["fp32"][1][r], res["f64"][0][r], res["f64"][1][r]) for r in range(R)]
    e6 = [rel_l2(res["x6"][0][r], res["x6"][1][r], res["f64"][0][r], res["f64"][1][r]) for r in range(R)]
    print("far-field variants, max relL2 vs FP64: bf16x3 %.3e  fp32 %.3e  bf16x6 %.3e" % (max(eb), max(ef), max(e6)))
    assert max(ef) < max(eb) and max(ef) < 2.5e-7, (eb, ef)
    assert max(e6) < max(eb) and max(e6) <= 1.1 * max(ef), (eb, ef, e6)


@pytest.mark.parametrize("m2l,near64", [(1, 2), (2, 2), (2, 0), (3, 2)])
def test_stored_equals_on_the_fly_with_precision_controls(cuda, m2l, near64):
    import torch
    import salfort as S
    x, y, z, area = S.gen_mpas(60000, seed=4)
    R = 20
    ssh = S.gen_ssh(x, y, z, lmax=32, amp=0.1, noise=0.02, seed=3, R=R)
    dx, dy, dz, da, ds = (_dev(v, cuda) for v in (x, y, z, area, ssh))
    g = S.Geometry(device_xyz=(dx, dy, dz), cluster_thresh=65)
    g.dtt(0.7)
    out = []
    for store in (0, 1):
        o = S.default_opts(degree=7, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, m2l=m2l, near64=near64,
                           store=store, rhs_tile=32)
        gx = torch.zeros((R, len(x)), dtype=torch.float64, device=cuda)
        gy = torch.zeros_like(gx)
        g.eval(da, ds, gx, gy, o)
        out.append((gx.cpu().numpy(), gy.cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert np.all(np.isfinite(out[0][0]))
