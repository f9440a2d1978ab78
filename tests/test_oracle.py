"""The oracle (C restatement of the reference) against the reference's golden outputs and against
what the mathematics fixes.  CPU only."""
import numpy as np
import pytest

from conftest import fib, load_golden, rel_l2

TREE_KEYS = ["level", "is_leaf", "parent", "child", "face", "bounds", "radius", "count", "offsets", "points"]


@pytest.mark.parametrize("case", ["fib642", "mpas4000"])
def test_tree_and_list_bit_exact(oracle, case):
    G = load_golden(case)
    oracle.tree_build(G["x"], G["y"], G["z"], int(G["thresh"]))
    t = oracle.tree_export()
    for k in TREE_KEYS:
        assert np.array_equal(t[k], G["tree_" + k]), k
    it, is_, ty = oracle.dtt(float(G["theta"]), int(G["thresh"]))
    assert np.array_equal(it, G["dtt_t"]) and np.array_equal(is_, G["dtt_s"]) and np.array_equal(ty, G["dtt_ty"])


@pytest.mark.parametrize("case", ["fib642", "mpas4000"])
def test_fast_sum_matches_reference(oracle, case):
    G = load_golden(case)
    oracle.tree_build(G["x"], G["y"], G["z"], int(G["thresh"]))
    oracle.dtt(float(G["theta"]), int(G["thresh"]))
    gx, gy = oracle.fast_sum(G["x"], G["y"], G["z"], G["area"], G["ssh"], int(G["degree"]))
    assert np.array_equal(gx, G["fs_gx"]) and np.array_equal(gy, G["fs_gy"])  # same order, same libm


def test_mom6_path_matches_reference(oracle):
    G = load_golden("fib642")
    x, y, z = G["x"], G["y"], G["z"]
    oracle.tree_build(x, y, z, int(G["thresh"]))
    pp = oracle.assign(x, y, z, G["assign"].shape[1])
    assert np.array_equal(pp, G["assign"])
    pct, pcs, ppt, pps = oracle.stt(float(G["theta"]), int(G["thresh"]), x, y, z)
    for a, k in [(pct, "stt_pc_t"), (pcs, "stt_pc_s"), (ppt, "stt_pp_t"), (pps, "stt_pp_s")]:
        assert np.array_equal(a, G[k]), k
    w = oracle.proxy_source(pp, x, y, z, G["ssh"], G["area"], int(G["degree"]))
    assert np.array_equal(w, G["proxy_w"])
    gx, gy = oracle.mom6_sums(x, y, z, G["area"], G["ssh"], int(G["degree"]), w)
    assert np.array_equal(gx, G["m6_gx"]) and np.array_equal(gy, G["m6_gy"])
    # the sums of selected targets only: the same bits there, zeros elsewhere
    tg = np.arange(3, len(x), 7)
    sx, sy = oracle.mom6_sums(x, y, z, G["area"], G["ssh"], int(G["degree"]), w, targets=tg)
    assert np.array_equal(sx[tg], gx[tg]) and np.array_equal(sy[tg], gy[tg])
    rest = np.setdiff1d(np.arange(len(x)), tg)
    assert not sx[rest].any() and not sy[rest].any()


def test_kernel_bit_exact(oracle):
    G = load_golden("kernels")
    g = np.array([oracle.gfunc(tuple(a), tuple(b)) for a, b in zip(G["k_t"], G["k_s"])])
    assert np.array_equal(g, G["k_g"])
    assert np.array_equal(G["k_g"], G["k_g6"])  # the MOM6 copy is the same function


def test_kernel_closed_form(oracle):
    """p1 + p2 equals -5.21196/c^3 - 17.4 (1+c)/(c^2 (2+c)) (up to the 1e-16 regularisers) and the
    east/north components are the projection of val * s on the target's basis."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        t = rng.normal(size=3); t /= np.linalg.norm(t)
        s = rng.normal(size=3); s /= np.linalg.norm(s)
        c = np.linalg.norm(s - t)
        val = (-5.21196 / c ** 3 - 17.4 * (1 + c) / (c ** 2 * (2 + c))) * -2.343256857908601e-9
        cp = np.sqrt(1 - t[2] ** 2)
        e = np.array([-t[1], t[0], 0]) / cp
        nvec = np.array([-t[2] * t[0], -t[2] * t[1], 1 - t[2] ** 2]) / cp
        gx, gy = oracle.gfunc(tuple(t), tuple(s))
        assert gx == pytest.approx(val * (s @ e), rel=1e-9, abs=1e-20)
        assert gy == pytest.approx(val * (s @ nvec), rel=1e-9, abs=1e-20)
    assert oracle.gfunc((0.0, 0.0, 1.0), (0.5, 0.5, np.sqrt(0.5))) == (0.0, 0.0)  # pole


def test_bli_matches_reference_and_closed_forms(oracle):
    G = load_golden("kernels")
    off = 0
    for d in G["bli_degs"]:
        c = oracle.bli_coeffs(int(d))
        assert np.array_equal(c, G["bli_coefs"][off:off + d + 1])
        if d > 0:  # [1/2, -1, 1, ..., (-1)^d / 2]
            ref = np.array([(-1.0) ** i for i in range(d + 1)])
            ref[0] *= 0.5
            ref[-1] *= 0.5
            assert np.array_equal(c, ref)
        off += d + 1
    off = 0
    for k, (d, xi, eta, a, b, c_, e_) in enumerate(G["bli_args"]):
        d = int(d)
        v = oracle.interp_vals(xi, eta, a, b, c_, e_, d).ravel()
        assert np.array_equal(v, G["bli_vals"][off:off + (d + 1) ** 2])
        off += (d + 1) ** 2


@pytest.mark.parametrize("d", [1, 2, 4, 8, 12])
def test_bli_polynomial_exactness(oracle, d):
    """degree-d polynomials are reproduced; the basis sums to one; Kronecker at the nodes;
    nodes are cos(k pi / d) mapped to [a, b], descending."""
    a, b = -0.37, 0.52
    p = oracle.bli_points(a, b, d)
    assert np.allclose(p, 0.5 * (b - a) * np.cos(np.pi * np.arange(d + 1) / d) + 0.5 * (a + b), rtol=0, atol=1e-15)
    rng = np.random.default_rng(d)
    cx = rng.normal(size=d + 1)
    cy = rng.normal(size=d + 1)
    for xi, eta in rng.uniform(a, b, size=(5, 2)):
        v = oracle.interp_vals(xi, eta, a, b, a, b, d)
        assert abs(v.sum() - 1) < 1e-12
        f = np.polyval(cx, p)[:, None] * np.polyval(cy, p)[None, :]
        assert (v * f).sum() == pytest.approx(np.polyval(cx, xi) * np.polyval(cy, eta), rel=1e-10, abs=1e-10)
    v = oracle.interp_vals(p[2 % (d + 1)], p[0], a, b, a, b, d)
    assert v[2 % (d + 1), 0] == 1.0 and v.sum() == 1.0


def test_cube_sphere_maps(oracle):
    G = load_golden("kernels")
    for p, f, xe, back in zip(G["cs_p"], G["cs_face"], G["cs_xieta"], G["cs_back"]):
        assert oracle.face(*p) == f
        assert oracle.xieta(*p) == tuple(xe)
        assert oracle.xyz(xe[0], xe[1], int(f)) == tuple(back)
        assert np.allclose(back, p, atol=1e-15)  # round trip
        assert abs(np.linalg.norm(back) - 1) < 1e-15


def test_tree_invariants(oracle):
    x, y, z = fib(10242)
    thresh = 40
    oracle.tree_build(x, y, z, thresh)
    t = oracle.tree_export()
    P = len(t["level"])
    cnt, off, pts = t["count"], t["offsets"], t["points"]
    for i in range(P):
        seg = pts[off[i]:off[i + 1]]
        assert np.all(np.diff(seg) > 0)  # ascending
        if t["is_leaf"][i]:
            assert cnt[i] < thresh
        else:
            ch = t["child"][i]
            assert np.array_equal(ch, ch[0] + np.arange(4)) and ch[0] > i  # consecutive, BFS
            assert np.array_equal(np.sort(np.concatenate([pts[off[c]:off[c + 1]] for c in ch])), seg)
    roots = np.concatenate([pts[off[i]:off[i + 1]] for i in range(6)])
    assert np.array_equal(np.sort(roots), np.arange(len(x)))
    assert np.all(np.diff(t["level"]) >= 0)  # BFS: levels non-decreasing with id


def test_dtt_covers_every_pair_once(oracle):
    x, y, z = fib(2562)
    thresh = 12
    oracle.tree_build(x, y, z, thresh)
    t = oracle.tree_export()
    it, is_, ty = oracle.dtt(0.7, thresh)
    off, pts = t["offsets"], t["points"]
    cover = np.zeros((len(x), len(x)), np.int32)
    for a, b in zip(it, is_):
        ta, sb = pts[off[a]:off[a + 1]], pts[off[b]:off[b + 1]]
        cover[np.ix_(ta, sb)] += 1
    assert np.all(cover == 1)


def test_all_pp_limit_equals_direct_sum(oracle):
    """theta -> 0 makes every entry a PP entry: the treecode is the direct sum (same order per
    target up to the panel grouping)."""
    x, y, z = fib(1500)
    area = np.full(1500, 1e10)
    ssh = np.cos(3 * x) * y
    oracle.tree_build(x, y, z, 10)
    it, is_, ty = oracle.dtt(1e-6, 10)
    assert np.all(ty == 0)
    gx, gy = oracle.fast_sum(x, y, z, area, ssh, 3)
    dx, dy = oracle.direct_sum(x, y, z, area, ssh, np.arange(1500), 0)
    assert rel_l2(gx, gy, dx, dy) < 1e-13


def test_direct_sum_brute_force(oracle):
    """the direct sum against a numpy evaluation of the kernel on a tiny grid"""
    x, y, z = fib(200)
    area = np.linspace(1, 2, 200) * 1e10
    ssh = np.sin(2 * y)
    P = np.stack([x, y, z], 1)
    tg = np.array([0, 17, 99, 150])
    for mode in (1, 2):
        gx, gy = oracle.direct_sum(x, y, z, area, ssh, tg, mode)
        for k, t in enumerate(tg):
            s = np.delete(P, t, 0)
            q = np.delete(area * ssh, t)
            c = np.linalg.norm(s - P[t], axis=1)
            val = (-5.21196 / (c ** 3 + 1e-16) + 8.7 * (2 - c * c + c) / (-0.5 * c * c * (4 - c * c) + 1e-16))
            val *= -2.343256857908601e-9
            cp = np.sqrt(1 - P[t, 2] ** 2)
            e = np.array([-P[t, 1], P[t, 0], 0]) / cp
            nv = np.array([-P[t, 2] * P[t, 0], -P[t, 2] * P[t, 1], 1 - P[t, 2] ** 2]) / cp
            assert gx[k] == pytest.approx(np.sum(val * q * (s @ e)), rel=1e-11)
            assert gy[k] == pytest.approx(np.sum(val * q * (s @ nv)), rel=1e-11)


def test_accuracy_mode_converges(oracle):
    """accuracy mode: error against the direct sum falls with the degree (closed-form check of the
    tangential interpolation); the reference mode stalls on the poles (F5)."""
    x, y, z = fib(10242)
    area = np.full(len(x), 4 * np.pi * 6.371e6 ** 2 / len(x))
    ssh = 0.1 * np.sin(3 * x) * np.cos(2 * y) + 0.05 * z ** 3
    oracle.tree_build(x, y, z, 33)
    oracle.dtt(0.7, 33)
    tg = np.arange(0, len(x), 37)
    dx, dy = oracle.direct_sum(x, y, z, area, ssh, tg, 2)
    errs = []
    for d in (3, 5, 7):
        gx, gy = oracle.fast_sum_cart(x, y, z, area, ssh, d)
        errs.append(rel_l2(gx[tg], gy[tg], dx, dy))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 3e-6
