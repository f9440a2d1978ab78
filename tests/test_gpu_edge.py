"""Degenerate grids on the GPU against the oracle: a single point, fewer points than cube faces, a
tree that never refines, and a dense cap that leaves most panels empty.  Tree and list bit-exact,
FP64 sums within 1e-12 of the oracle, the mixed path within 1e-6 of FP64 accurate mode."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _cap(n, seed, radius_deg=2.0):
    """n points in a spherical cap of the given radius around (lat 40, lon -60)."""
    rng = np.random.default_rng(seed)
    c = np.array([np.cos(np.radians(40)) * np.cos(np.radians(-60)), np.cos(np.radians(40)) * np.sin(np.radians(-60)),
                  np.sin(np.radians(40))])
    u = np.cross(c, [0.0, 0.0, 1.0])
    u /= np.linalg.norm(u)
    v = np.cross(c, u)
    r = np.radians(radius_deg) * np.sqrt(rng.random(n))
    a = 2 * np.pi * rng.random(n)
    p = np.cos(r)[:, None] * c + np.sin(r)[:, None] * (np.cos(a)[:, None] * u + np.sin(a)[:, None] * v)
    p /= np.linalg.norm(p, axis=1)[:, None]
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()


def _grid(kind):
    import salfort as S
    if kind == "one":
        x, y, z = np.array([0.3]), np.array([-0.5]), np.array([np.sqrt(1 - 0.34)])
        return x, y, z, 65
    if kind == "seven":
        x, y, z, _ = S.gen_mpas(7, seed=3)
        return x, y, z, 65
    if kind == "no_refine":
        x, y, z, _ = S.gen_mpas(200, seed=4)
        return x, y, z, 1000
    x, y, z = _cap(6000, seed=5)
    return x, y, z, 33


@pytest.mark.parametrize("kind", ["one", "seven", "no_refine", "cap"])
def test_degenerate_grids_match_oracle(cuda, oracle, kind):
    import salfort as S
    x, y, z, thresh = _grid(kind)
    n = len(x)
    area = np.full(n, 1.0e10)
    ssh = 0.1 + 0.05 * x + 0.02 * np.sin(5 * y)
    g = S.Geometry(x, y, z, thresh)
    g.dtt(0.7)
    oracle.tree_build(x, y, z, thresh)
    t = g.export()
    o = oracle.tree_export()
    for k in ("level", "is_leaf", "parent", "child", "face", "bounds", "radius", "count", "offsets", "points"):
        assert np.array_equal(t[k], o[k]), k
    it, is_, ty = g.dtt_export()
    oi, os_, oty = oracle.dtt(0.7, thresh)
    assert np.array_equal(it, oi) and np.array_equal(is_, os_) and np.array_equal(ty, oty)
    gx, gy = g.fast_sum(area, ssh, 4)
    ox, oy = oracle.fast_sum(x, y, z, area, ssh, 4)
    if not (np.any(ox) or np.any(oy)):  # a lone point: its self pair's numerator cancels exactly
        assert not (np.any(gx) or np.any(gy))
    else:
        assert rel_l2(gx, gy, ox, oy) < 1e-12


@pytest.mark.parametrize("kind", ["one", "seven", "no_refine", "cap"])
def test_degenerate_grids_mixed(cuda, kind):
    import salfort as S
    import torch
    x, y, z, thresh = _grid(kind)
    n = len(x)
    area = np.full(n, 1.0e10)
    R = 3
    ssh = np.stack([0.1 + 0.05 * x, 0.02 * np.sin(5 * y), -0.03 + 0.01 * z])
    g = S.Geometry(x, y, z, thresh)
    g.dtt(0.7)
    dev = lambda v: torch.as_tensor(np.ascontiguousarray(v), dtype=torch.float64, device=cuda)
    a, s = dev(area), dev(ssh)
    fx, fy = torch.zeros((R, n), dtype=torch.float64, device=cuda), torch.zeros((R, n), dtype=torch.float64, device=cuda)
    mx, my = torch.zeros_like(fx), torch.zeros_like(fy)
    g.eval(a, s, fx, fy, S.default_opts(degree=6, mode=S.MODE_ACCURATE, precision=S.PREC_FP64))
    for store in (0, 1):
        g.eval(a, s, mx, my, S.default_opts(degree=6, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, store=store))
        torch.cuda.synchronize()
        if n == 1:  # a lone point has no other source: every sum is exactly zero
            assert float(fx.abs().max()) == 0.0 and float(mx.abs().max()) == 0.0
            continue
        for r in range(R):
            assert rel_l2(mx[r].cpu().numpy(), my[r].cpu().numpy(), fx[r].cpu().numpy(), fy[r].cpu().numpy()) < 1e-6
