"""GPU evaluation against the oracle and the reference's golden outputs.

Bars: FP64 reference mode within 1e-12 relative L2 of the reference treecode (the reference's own
fast_sum output, and the oracle's bit-identical restatement); FP64 accuracy mode within 1e-12 of the
oracle's accuracy mode; the FP64 direct sum within 1e-13 of the oracle's direct sum."""
import numpy as np
import pytest

from conftest import fib, load_golden, rel_l2

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _dev(a, cuda):
    return _torch().as_tensor(np.ascontiguousarray(a), dtype=_torch().float64, device=cuda)


@pytest.mark.parametrize("case", ["fib642", "mpas4000"])
def test_golden_fast_sum(cuda, case):
    import salfort as S
    G = load_golden(case)
    g = S.Geometry(G["x"], G["y"], G["z"], int(G["thresh"]))
    g.dtt(float(G["theta"]))
    gx, gy = g.fast_sum(G["area"], G["ssh"], int(G["degree"]))
    assert rel_l2(gx, gy, G["fs_gx"], G["fs_gy"]) < 1e-12


@pytest.mark.parametrize("kind,n,thresh,theta,deg", [
    ("fib", 40962, 65, 0.7, 4), ("mpas", 50000, 129, 0.7, 8), ("mpas", 20000, 33, 0.5, 6), ("fib", 3000, 10, 0.9, 2),
    # 256 proxies (one pass of the proxy-target block) and 289 (two passes)
    ("fib", 12000, 65, 0.7, 15), ("fib", 12000, 65, 0.7, 16),
])
def test_fast_sum_matches_oracle(cuda, oracle, kind, n, thresh, theta, deg):
    import salfort as S
    if kind == "fib":
        x, y, z = fib(n)
        area = np.full(n, 4 * np.pi * 6.371e6 ** 2 / n)
    else:
        x, y, z, area = S.gen_mpas(n, seed=11)
    ssh = S.gen_ssh(x, y, z, lmax=16, amp=0.1, noise=0.02, seed=4)
    g = S.Geometry(x, y, z, thresh)
    g.dtt(theta)
    gx, gy = g.fast_sum(area, ssh, deg)
    oracle.tree_build(x, y, z, thresh)
    oracle.dtt(theta, thresh)
    ox, oy = oracle.fast_sum(x, y, z, area, ssh, deg)
    assert rel_l2(gx, gy, ox, oy) < 1e-12
    # accumulate semantics (fast_sum_module.f90:69-70): adds into the caller's arrays
    hx, hy = g.fast_sum(area, ssh, deg, gx.copy(), gy.copy())
    assert rel_l2(hx, hy, 2 * ox, 2 * oy) < 1e-12


@pytest.mark.parametrize("stride", [1, 2, 3])
def test_pair_tasks_every_shape(cuda, oracle, stride):
    """The reference path sums PP entries once per mirrored pair of entries.  Leaves of up to 128
    points give a panel's entry with itself above 64 points and tasks with more lane points than
    staged points; a list of every stride-th entry leaves many entries without their mirror
    (one-directional tasks).  Against the oracle's sums over the same entries (rank 0 of `stride`)."""
    import salfort as S
    x, y, z, area = S.gen_mpas(40000, seed=21)
    ssh = S.gen_ssh(x, y, z, lmax=16, amp=0.1, noise=0.02, seed=6)
    g = S.Geometry(x, y, z, 129)
    g.dtt(0.7)
    it, is_, ty = g.dtt_export()
    cnt = g.export()["count"]
    assert cnt[is_[ty == 0]].max() > 64 and cnt[is_[ty == 0]].min() < 32  # both task shapes occur
    sel = (np.arange(len(it)) + 1) % stride == 0
    h = S.Geometry(x, y, z, 129)
    h.dtt_import(it[sel], is_[sel], ty[sel])
    gx, gy = h.fast_sum(area, ssh, 6)
    oracle.tree_build(x, y, z, 129)
    oracle.dtt(0.7, 129)
    ox, oy = oracle.fast_sum(x, y, z, area, ssh, 6, rank=0, nproc=stride)
    assert rel_l2(gx, gy, ox, oy) < 1e-12
    # the LDS sums of the pair tasks take their adds in a fixed order: the same bits on every call
    for _ in range(2):
        hx, hy = h.fast_sum(area, ssh, 6)
        assert np.array_equal(hx, gx) and np.array_equal(hy, gy)


def test_pp_entries_of_large_panels(cuda, oracle):
    """A list typed with a larger cluster threshold than the tree's makes PP entries between panels
    of more than 128 points, which the leaf kernel sums itself (beside the pair tasks of the small
    ones); against the oracle on the same list."""
    import salfort as S
    x, y, z, area = S.gen_mpas(20000, seed=8)
    ssh = S.gen_ssh(x, y, z, lmax=16, amp=0.1, noise=0.02, seed=9)
    g = S.Geometry(x, y, z, 33)
    g.dtt(0.7, 300)
    it, is_, ty = g.dtt_export()
    cnt = g.export()["count"]
    assert cnt[it[ty == 0]].max() > 128 and cnt[it[ty == 0]].min() <= 128
    gx, gy = g.fast_sum(area, ssh, 4)
    oracle.tree_build(x, y, z, 33)
    oracle.dtt(0.7, 300)
    ox, oy = oracle.fast_sum(x, y, z, area, ssh, 4)
    assert rel_l2(gx, gy, ox, oy) < 1e-12


def test_accurate_fp64_matches_oracle(cuda, oracle):
    import salfort as S
    x, y, z, area = S.gen_mpas(30000, seed=2)
    ssh = S.gen_ssh(x, y, z, lmax=32, amp=0.1, noise=0.02, seed=5)
    g = S.Geometry(x, y, z, 65)
    g.dtt(0.7)
    oracle.tree_build(x, y, z, 65)
    oracle.dtt(0.7, 65)
    ox, oy = oracle.fast_sum_cart(x, y, z, area, ssh, 6)
    a, s = _dev(area, cuda), _dev(ssh, cuda)
    gx, gy = _torch().zeros_like(s), _torch().zeros_like(s)
    g.eval(a, s, gx, gy, S.default_opts(degree=6, mode=S.MODE_ACCURATE, precision=S.PREC_FP64))
    _torch().cuda.synchronize()
    assert rel_l2(gx.cpu().numpy(), gy.cpu().numpy(), ox, oy) < 1e-12


def test_eval_many_fields_and_accumulate(cuda, oracle):
    import salfort as S
    x, y, z, area = S.gen_mpas(20000, seed=3)
    ssh = S.gen_ssh(x, y, z, lmax=16, amp=0.1, noise=0.02, seed=9, R=3)
    g = S.Geometry(x, y, z, 33)
    g.dtt(0.7)
    a, s = _dev(area, cuda), _dev(ssh, cuda)
    gx = _torch().ones_like(s)
    gy = _torch().ones_like(s)
    g.eval(a, s, gx, gy, S.default_opts(degree=5, mode=S.MODE_REFERENCE, precision=S.PREC_FP64), accumulate=True)
    _torch().cuda.synchronize()
    oracle.tree_build(x, y, z, 33)
    oracle.dtt(0.7, 33)
    for r in range(3):
        ox, oy = oracle.fast_sum(x, y, z, area, ssh[r], 5)
        assert rel_l2(gx[r].cpu().numpy() - 1, gy[r].cpu().numpy() - 1, ox, oy) < 1e-12


def test_all_pp_limit_equals_direct_sum(cuda):
    """theta -> 0: every entry is PP and the GPU treecode is the GPU direct sum (self included)."""
    import salfort as S
    x, y, z = fib(4000)
    area = np.full(4000, 1e10)
    ssh = np.cos(3 * x) * y
    g = S.Geometry(x, y, z, 20)
    g.dtt(1e-7)
    assert g.dtt_counts()[1:].sum() == 0
    gx, gy = g.fast_sum(area, ssh, 3)
    X = [_dev(v, cuda) for v in (x, y, z, area, ssh)]
    tg = _torch().arange(4000, dtype=_torch().int32, device=cuda)
    dx, dy = S.direct_sum(*X, tg, 0)
    assert rel_l2(gx, gy, dx.cpu().numpy(), dy.cpu().numpy()) < 1e-13


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_direct_sum_matches_oracle(cuda, oracle, mode):
    import salfort as S
    x, y, z, area = S.gen_mpas(40000, seed=1)
    ssh = S.gen_ssh(x, y, z, lmax=16, amp=0.1, noise=0.02, seed=2)
    tg = np.random.default_rng(0).choice(40000, 300, replace=False).astype(np.int32)
    ox, oy = oracle.direct_sum(x, y, z, area, ssh, tg, mode)
    X = [_dev(v, cuda) for v in (x, y, z, area, ssh)]
    dx, dy = S.direct_sum(*X, _torch().as_tensor(tg, device=cuda), mode)
    assert rel_l2(dx.cpu().numpy(), dy.cpu().numpy(), ox, oy) < 1e-13


def test_device_kernel_expression_is_the_reference(cuda):
    """Self pairs are rounding noise of order 1e-8 * q (F2 in the reference); the treecode's
    reference mode reproduces them: on a grid where every target's self pair is its largest term,
    the GPU and the reference agree to 1e-12 on the self-pair-only sum."""
    import salfort as S
    G = load_golden("fib642")
    g = S.Geometry(G["x"], G["y"], G["z"], 10)
    g.dtt(1e-7)  # all PP
    gx, gy = g.fast_sum(G["area"], G["ssh"], 3)
    from oracle_py import Oracle
    o = Oracle()
    o.tree_build(G["x"], G["y"], G["z"], 10)
    o.dtt(1e-7, 10)
    ox, oy = o.fast_sum(G["x"], G["y"], G["z"], G["area"], G["ssh"], 3)
    assert rel_l2(gx, gy, ox, oy) < 1e-12


@pytest.mark.parametrize("R,tile,deg,m2l", [(1, 16, 8, 1), (5, 16, 6, 1), (20, 16, 8, 1), (32, 32, 8, 1),
                                           (32, 32, 8, 2), (17, 32, 10, 1), (3, 16, 4, 2), (32, 32, 7, 1),
                                           (9, 16, 5, 1), (13, 32, 9, 1), (32, 32, 8, 3), (3, 16, 4, 3)])
def test_mixed_matches_fp64_accurate(cuda, R, tile, deg, m2l):
    """Mixed precision (FP32 kernel, FP64 differences and running sums) against the FP64 accuracy
    mode on the same list: the difference is rounding only, well under the 1e-6 accuracy bar."""
    import salfort as S
    torch = _torch()
    x, y, z, area = S.gen_mpas(40000, seed=6)
    ssh = S.gen_ssh(x, y, z, lmax=32, amp=0.1, noise=0.02, seed=7, R=R).reshape(R, -1)
    g = S.Geometry(x, y, z, 65)
    g.dtt(0.7)
    a, s = _dev(area, cuda), _dev(ssh, cuda)
    fx, fy = torch.zeros_like(s), torch.zeros_like(s)
    g.eval(a, s, fx, fy, S.default_opts(degree=deg, mode=S.MODE_ACCURATE, precision=S.PREC_FP64))
    mx, my = torch.zeros_like(s), torch.zeros_like(s)
    g.eval(a, s, mx, my, S.default_opts(degree=deg, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED,
                                         m2l=m2l, rhs_tile=tile))
    torch.cuda.synchronize()
    for r in range(R):
        e = rel_l2(mx[r].cpu().numpy(), my[r].cpu().numpy(), fx[r].cpu().numpy(), fy[r].cpu().numpy())
        assert e < 1e-6, (r, e)  # FP32 kernel and sums, split-bf16 cluster products


@pytest.mark.parametrize("deg", [7, 8])
def test_mixed_accuracy_vs_direct_sum(cuda, deg):
    """The headline accuracy bar at reduced size: mixed precision, degree 7 or 8, theta 0.7, within
    1e-6 relative L2 of the FP64 direct sum (self pair excluded) at sampled targets."""
    import salfort as S
    torch = _torch()
    x, y, z, area = S.gen_mpas(200000, seed=0, total_area=3.61e14)
    ssh = S.gen_ssh(x, y, z, lmax=64, amp=0.1, noise=0.02, seed=1, R=4).reshape(4, -1)
    g = S.Geometry(x, y, z, 129)
    g.dtt(0.7)
    a, s = _dev(area, cuda), _dev(ssh, cuda)
    X = [_dev(v, cuda) for v in (x, y, z)]
    mx, my = torch.zeros_like(s), torch.zeros_like(s)
    g.eval(a, s, mx, my, S.default_opts(degree=deg, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED))
    tg = torch.as_tensor(np.sort(np.random.default_rng(5).choice(len(x), 2000, replace=False)).astype(np.int32),
                         device=cuda)
    for r in range(4):
        dx, dy = S.direct_sum(*X, a, s[r].contiguous(), tg, 2)
        e = rel_l2(mx[r][tg.long()].cpu().numpy(), my[r][tg.long()].cpu().numpy(), dx.cpu().numpy(), dy.cpu().numpy())
        assert e < 1e-6, (r, e)


@pytest.mark.parametrize("R,tile,deg,m2l", [(7, 16, 8, 1), (20, 32, 8, 1), (33, 32, 6, 1), (5, 16, 10, 1),
                                           (32, 32, 7, 1), (7, 16, 7, 1), (9, 16, 5, 1), (13, 32, 9, 1),
                                           (20, 32, 4, 1),
                                           (7, 16, 8, 2), (33, 32, 6, 2), (5, 16, 10, 2), (9, 16, 5, 2),
                                           (13, 32, 9, 2), (20, 32, 4, 2), (32, 32, 7, 2),
                                           (7, 16, 8, 3), (33, 32, 6, 3), (5, 16, 10, 3), (9, 16, 5, 3),
                                           (13, 32, 9, 3), (20, 32, 4, 3), (32, 32, 7, 3)])
def test_stored_operator_matches_on_the_fly(cuda, R, tile, deg, m2l):
    """The stored operator (kernel values evaluated once by the kernels' own traversal, streamed on
    every call) reproduces the on-the-fly evaluation bit for bit, also replayed as a HIP graph, with
    the split-bf16 far field (m2l 1), the FP32 one (m2l 2) and the three-part bf16 one (m2l 3); the
    stored forms of the last two take a whole RHS tile per pass where their on-the-fly forms take 16
    fields."""
    import salfort as S
    torch = _torch()
    x, y, z, area = S.gen_mpas(60000, seed=8)
    ssh = S.gen_ssh(x, y, z, lmax=32, amp=0.1, noise=0.02, seed=9, R=R).reshape(R, -1)
    g = S.Geometry(x, y, z, 65)
    g.dtt(0.7)
    a, s = _dev(area, cuda), _dev(ssh, cuda)
    o0 = S.default_opts(degree=deg, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, rhs_tile=tile, m2l=m2l)
    o1 = S.default_opts(degree=deg, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED, rhs_tile=tile, store=1, m2l=m2l)
    fx, fy = torch.zeros_like(s), torch.zeros_like(s)
    g.eval(a, s, fx, fy, o0)
    st = g.plan_stats(o1)
    assert st["stored_near_bytes"] > 0 and st["stored_far_bytes"] > 0
    for graph in (0, 1, 1):
        o1.graph = graph
        sx, sy = torch.full_like(s, float("nan")), torch.full_like(s, float("nan"))
        g.eval(a, s, sx, sy, o1)
        torch.cuda.synchronize()
        assert torch.equal(sx, fx) and torch.equal(sy, fy), graph


@pytest.mark.parametrize("tree_thresh,list_thresh", [(65, 100), (65, 20), (129, 129)])
def test_mixed_upward_pass_any_list_threshold(cuda, tree_thresh, list_thresh):
    """The mixed upward pass (P2M into the lowest cluster ancestor, then M2M) against the FP64
    accuracy mode, with the list's cluster threshold above, below and equal to the tree's."""
    import salfort as S
    torch = _torch()
    x, y, z, area = S.gen_mpas(50000, seed=11)
    ssh = S.gen_ssh(x, y, z, lmax=32, amp=0.1, noise=0.02, seed=12, R=3).reshape(3, -1)
    g = S.Geometry(x, y, z, tree_thresh)
    g.dtt(0.7, cluster_thresh=list_thresh)
    a, s = _dev(area, cuda), _dev(ssh, cuda)
    fx, fy = torch.zeros_like(s), torch.zeros_like(s)
    g.eval(a, s, fx, fy, S.default_opts(degree=7, mode=S.MODE_ACCURATE, precision=S.PREC_FP64))
    mx, my = torch.zeros_like(s), torch.zeros_like(s)
    g.eval(a, s, mx, my, S.default_opts(degree=7, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED))
    torch.cuda.synchronize()
    for r in range(3):
        e = rel_l2(mx[r].cpu().numpy(), my[r].cpu().numpy(), fx[r].cpu().numpy(), fy[r].cpu().numpy())
        assert e < 1e-6, (r, e)


def test_gradient_xyz_is_the_tangential_sum(cuda):
    """sal_gradient_xyz_dev turns the direct sum's (east, north) into a 3-vector; that vector must be
    the tangential form of the same sum, sum_s val(t, s) q_s P_t s with P_t s = s - (t.s) t, evaluated
    here in numpy (a different expression from the east/north projections: a wrong basis vector,
    sign or component order fails it), and zero at the poles, where the reference returns zero.
    Both sides use the cancellation-free form (direct-sum mode 2): with d = s - t, c^2 = |d|^2,
    t.s = 1 - c^2/2 in the reference's val (fast_sum_module.f90:20-24) and P_t s = d + (c^2/2) t.
    (With the reference's t.s the close pairs' val is ill-conditioned: two FP64 evaluations of it
    differ by ~1e-12 from the dot product's rounding alone.)"""
    import salfort as S
    torch = _torch()
    x, y, z, _ = S.gen_mpas(6000, seed=11)
    x = np.concatenate([x, [0.0, 0.0]])
    y = np.concatenate([y, [0.0, 0.0]])
    z = np.concatenate([z, [1.0, -1.0]])
    n = len(x)
    rng = np.random.default_rng(5)
    area = rng.uniform(0.5e10, 2e10, n)
    ssh = 0.1 * np.sin(3 * x) + 0.05 * y * z
    tg = np.concatenate([rng.choice(n - 2, 40, replace=False), [n - 2, n - 1]]).astype(np.int32)
    dx, dy, dz = _dev(x, cuda), _dev(y, cuda), _dev(z, cuda)
    gx, gy = S.direct_sum(dx, dy, dz, _dev(area, cuda), _dev(ssh, cuda),
                          torch.as_tensor(tg, device=cuda), mode=2)
    T = S.gradient_xyz(dx[tg.tolist()], dy[tg.tolist()], dz[tg.tolist()], gx, gy).cpu().numpy()
    P = np.stack([x, y, z], axis=1)
    q = area * ssh
    cons = -2.343256857908601e-9
    E = np.zeros((len(tg), 3))
    for a, t in enumerate(tg[:-2]):
        d = np.delete(P, t, axis=0) - P[t]
        c2 = (d * d).sum(axis=1)
        c = np.sqrt(c2)
        val = ((1.0 - 6.21196) / (c * c2 + 1e-16) + 8.7 * ((2.0 - c2) + c) / (-0.5 * c2 * (4.0 - c2) + 1e-16)) * cons
        w = val * np.delete(q, t)
        E[a] = (w[:, None] * (d + 0.5 * c2[:, None] * P[t][None, :])).sum(axis=0)
    assert np.array_equal(T[-2:], np.zeros((2, 3)))
    err = np.sqrt(((T[:-2] - E[:-2]) ** 2).sum() / (E[:-2] ** 2).sum())
    assert err < 1e-12, err
    assert np.abs((T[:-2] * P[tg[:-2]]).sum(axis=1)).max() < 1e-12 * np.abs(T).max()  # tangent
    # [R, n] input: each field's vectors are those of the single-field call
    T2 = S.gradient_xyz(dx[tg.tolist()], dy[tg.tolist()], dz[tg.tolist()], torch.stack([gx, 2 * gx]),
                        torch.stack([gy, 2 * gy])).cpu().numpy()
    assert np.array_equal(T2[0], T) and np.array_equal(T2[1], 2 * T)
