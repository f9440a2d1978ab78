"""Full-size parity on the GPU, at the benchmark's grids.

* MPAS-shaped grid, N = 2,313,486, d = 8, theta = 0.7, cluster_thresh = 129 (the plain FP64 GPU
  treecode's configuration): the tree export and the DTT list are bit-exact against the oracle's
  (tree_module.f90:59-398), and the FP64 sums of every 100th list entry (fast_sum as rank 0 of 100
  ranks, fast_sum_module.f90:291-332) are within 1e-12 relative L2 of the oracle's.
* MPAS-shaped grid, 32 fields, the batched mixed-precision setting the autotuner picks for it (d = 7,
  leaves <= 64, split-bf16 far field, stored operator): every field within 5e-7 relative L2 of the
  FP64 direct sum at 2,000 seeded targets, and the stored operator's result bit-identical to the
  same evaluation with kernel values computed on the fly.
* MOM6-shaped grid, 1440 x 1080 cells (1,104,192 ocean cells), the MOM6 driver's FP64 path at d = 8,
  leaves <= 128: the single-tree-traversal lists bit-exact against the oracle's
  (tree_module.f90:401-487), the proxy weights within 1e-13 (mom6_fast_sum_module.f90:33-74) and the
  per-point sums of 2,000 seeded targets within 1e-12 (:76-192).
* MOM6-shaped grid, 1440 x 1080 cells (about 1.1M ocean cells), 32 fields at the serving loop's
  setting for it (d = 8, leaves <= 128, three-part split-bf16 far field with FP64 sums, close pairs in FP64):
  every field within 5e-7 relative L2 of the FP64 direct sum at 20,000 seeded targets (the metric's
  bar is 1e-6).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TREE_KEYS = ["level", "is_leaf", "parent", "child", "face", "bounds", "radius", "count", "offsets", "points"]
STRIDE = 100


def test_mpas_full_size_tree_list_and_sums(cuda, oracle):
    import salfort as S
    W = S.workload("mpas", R=1)
    x, y, z, area, ssh = W["x"], W["y"], W["z"], W["area"], W["ssh"]
    assert len(x) == 2313486
    g = S.Geometry(x, y, z, 129)
    oracle.tree_build(x, y, z, 129)
    t, o = g.export(), oracle.tree_export()
    for k in TREE_KEYS:
        assert np.array_equal(t[k], o[k]), k
    g.dtt(0.7)
    a = g.dtt_export()
    b = oracle.dtt(0.7, 129)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    it, is_, ty = a
    sel = (np.arange(len(it)) + 1) % STRIDE == 0
    assert all((ty[sel] == k).sum() > 100 for k in range(4))  # every interaction type sampled
    h = S.Geometry(x, y, z, 129)
    h.dtt_import(it[sel], is_[sel], ty[sel])
    gx, gy = h.fast_sum(area, ssh, 8)
    ox, oy = oracle.fast_sum(x, y, z, area, ssh, 8, rank=0, nproc=STRIDE)
    assert rel_l2(gx, gy, ox, oy) < 1e-12


def test_mom6_full_grid_stt_weights_and_sums(cuda, oracle):
    """The MOM6 driver's FP64 path on the 1440 x 1080 MOM6-shaped grid (1,104,192 ocean cells), d = 8,
    leaves <= 128: the single-tree-traversal lists bit-exact against the oracle (reference order),
    the proxy weights within 1e-13 of the oracle's, and the per-point sums (the reference's transposed
    weight read) of 2,000 seeded targets within 1e-12 of the oracle's."""
    import salfort as S
    W = S.workload("mom6", R=1)
    x, y, z, area = W["x"], W["y"], W["z"], W["area"]
    ssh = np.ascontiguousarray(W["ssh"]).reshape(-1)
    n = len(x)
    assert 1.0e6 < n < 1.2e6
    g = S.Geometry(x, y, z, 129)
    g.stt(0.7, x, y, z)
    oracle.tree_build(x, y, z, 129)
    o = oracle.stt(0.7, 129, x, y, z)
    for a, b in zip(g.stt_export(), o):
        assert np.array_equal(a, b)
    pp = oracle.assign(x, y, z, g.counts()[1] + 1)
    w = oracle.proxy_source(pp, x, y, z, ssh, area, 8)
    gw = g.mom6_proxy_source(area, ssh, 8)
    assert np.max(np.abs(gw - w)) <= 1e-13 * np.max(np.abs(w))
    tg = np.sort(np.random.default_rng(5).choice(n, 2000, replace=False))
    ox, oy = oracle.mom6_sums(x, y, z, area, ssh, 8, w, targets=tg)
    gx, gy = g.mom6_sums(area, ssh, 8, transposed=True)
    assert rel_l2(gx[tg], gy[tg], ox[tg], oy[tg]) < 1e-12


def test_mom6_full_grid_accuracy_at_the_serving_setting(cuda):
    import torch
    import salfort as S
    import serve
    st = serve.GEOMETRY_SETTINGS["mom6"]
    x, y, z, area = S.gen_mom6()
    n, R = len(x), 32
    assert 1.0e6 < n < 1.2e6
    ssh = S.gen_ssh(x, y, z, lmax=64, amp=0.1, noise=0.02, seed=1, R=R).reshape(R, n)
    tens = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=cuda)
    dx, dy, dz, da, ds = (tens(v) for v in (x, y, z, area, ssh))
    g = S.Geometry(device_xyz=(dx, dy, dz), cluster_thresh=st["thresh"])
    g.dtt(0.7)
    o = S.default_opts(degree=st["degree"], mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, m2l=st["m2l"],
                       near64=st["near64"], store=1)
    gx = torch.zeros((R, n), dtype=torch.float64, device=cuda)
    gy = torch.zeros_like(gx)
    g.eval(da, ds, gx, gy, o)
    rng = np.random.default_rng(1234)
    tg = torch.as_tensor(np.sort(rng.choice(n, 20000, replace=False)).astype(np.int32), device=cuda)
    ex, ey = S.direct_sum(dx, dy, dz, da, ds, tg, 2)
    tx, ty = gx[:, tg.long()], gy[:, tg.long()]
    rel = torch.sqrt(((tx - ex) ** 2 + (ty - ey) ** 2).sum(dim=1) / (ex ** 2 + ey ** 2).sum(dim=1)).cpu().numpy()
    print("MOM6-shaped relL2 per field: max %.3e median %.3e" % (rel.max(), np.median(rel)))
    assert np.all(rel <= 5e-7), rel


def test_mpas_full_size_batched_accuracy(cuda):
    import torch
    import salfort as S
    W = S.workload("mpas", R=32)
    x, y, z, area = W["x"], W["y"], W["z"], W["area"]
    n, R = len(x), 32
    ssh = W["ssh"].reshape(R, n)
    tens = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=cuda)
    dx, dy, dz, da, ds = (tens(v) for v in (x, y, z, area, ssh))
    g = S.Geometry(device_xyz=(dx, dy, dz), cluster_thresh=65)
    g.dtt(0.7)
    out = []
    for store in (1, 0):
        o = S.default_opts(degree=7, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, m2l=S.M2L_BF16X3, near64=0,
                           store=store, rhs_tile=32)
        gx = torch.zeros((R, n), dtype=torch.float64, device=cuda)
        gy = torch.zeros_like(gx)
        g.eval(da, ds, gx, gy, o)
        out.append((gx, gy))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
    rng = np.random.default_rng(1234)
    tg = torch.as_tensor(np.sort(rng.choice(n, 2000, replace=False)).astype(np.int32), device=cuda)
    ex, ey = S.direct_sum(dx, dy, dz, da, ds, tg, 2)
    gx, gy = out[0]
    tx, ty = gx[:, tg.long()], gy[:, tg.long()]
    rel = torch.sqrt(((tx - ex) ** 2 + (ty - ey) ** 2).sum(dim=1) / (ex ** 2 + ey ** 2).sum(dim=1)).cpu().numpy()
    print("MPAS-shaped relL2 per field: max %.3e median %.3e" % (rel.max(), np.median(rel)))
    assert np.all(rel <= 5e-7), rel
    g.close()
