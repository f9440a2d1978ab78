"""GPU tree build, assign and dual-tree traversal against the oracle and the reference's golden
outputs: bit-exact (ids, bounds, radii, membership, list order)."""
import numpy as np
import pytest

from conftest import fib, load_golden

pytestmark = pytest.mark.gpu

TREE_KEYS = ["level", "is_leaf", "parent", "child", "face", "bounds", "radius", "count", "offsets", "points"]


def _grid(kind, n, seed=0):
    import salfort as S
    if kind == "fib":
        return fib(n)
    x, y, z, _ = S.gen_mpas(n, seed=seed)
    return x, y, z


@pytest.mark.parametrize("case", ["fib642", "mpas4000"])
def test_golden_tree_and_list(cuda, case):
    import salfort as S
    G = load_golden(case)
    g = S.Geometry(G["x"], G["y"], G["z"], int(G["thresh"]))
    t = g.export()
    for k in TREE_KEYS:
        assert np.array_equal(t[k], G["tree_" + k]), k
    g.dtt(float(G["theta"]))
    it, is_, ty = g.dtt_export()
    assert np.array_equal(it, G["dtt_t"]) and np.array_equal(is_, G["dtt_s"]) and np.array_equal(ty, G["dtt_ty"])


@pytest.mark.parametrize("kind,n,thresh,theta", [
    ("fib", 40962, 65, 0.7), ("fib", 2562, 10, 0.5), ("mpas", 60000, 33, 0.7), ("mpas", 30000, 129, 0.9),
    ("fib", 1000, 2, 0.7),  # every panel with two points refines: deep tree
])
def test_tree_and_dtt_match_oracle(cuda, oracle, kind, n, thresh, theta):
    import salfort as S
    x, y, z = _grid(kind, n)
    g = S.Geometry(x, y, z, thresh)
    oracle.tree_build(x, y, z, thresh)
    t, o = g.export(), oracle.tree_export()
    for k in TREE_KEYS:
        assert np.array_equal(t[k], o[k]), k
    g.dtt(theta)
    a = g.dtt_export()
    b = oracle.dtt(theta, thresh)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert np.array_equal(g.dtt_counts(), np.bincount(b[2], minlength=4))


def test_forced_ties_take_the_host_path(cuda, oracle):
    """With the tie window forced wide, every point and MAC decision is re-decided on the host;
    the tree and list must still be the reference's."""
    import salfort as S
    x, y, z = _grid("mpas", 20000, seed=3)
    oracle.tree_build(x, y, z, 40)
    S.set_tie_ulps(1e14)
    try:
        g = S.Geometry(x, y, z, 40)
        g.dtt(0.7)
        pt, pr = g.ties()
    finally:
        S.set_tie_ulps(16)
    assert pt > 100 and pr > 0
    t, o = g.export(), oracle.tree_export()
    for k in TREE_KEYS:
        assert np.array_equal(t[k], o[k]), k
    for u, v in zip(g.dtt_export(), oracle.dtt(0.7, 40)):
        assert np.array_equal(u, v)


def test_assign_points_matches_reference(cuda, oracle):
    import salfort as S
    G = load_golden("fib642")
    g = S.Geometry(G["x"], G["y"], G["z"], int(G["thresh"]))
    pp = g.assign_points(G["x"], G["y"], G["z"], G["assign"].shape[1])
    assert np.array_equal(pp, G["assign"])
    x, y, z = _grid("mpas", 50000, seed=5)
    g = S.Geometry(x, y, z, 33)
    oracle.tree_build(x, y, z, 33)
    nlev = g.counts()[1] + 1
    xq, yq, zq, _ = S.gen_mpas(5000, seed=99)  # points that are not the tree's own
    assert np.array_equal(g.assign_points(xq, yq, zq, nlev), oracle.assign(xq, yq, zq, nlev))


def test_face_edge_point_is_not_contained(cuda):
    """A point on a face's max edge (xi = pi/4) lies in no panel under contains_point's half-open
    test; the reference walks past its panel array there, the GPU walk stops and reports -1."""
    import salfort as S
    x, y, z = fib(3000)
    g = S.Geometry(x, y, z, 20)
    e = np.array([1.0, 1.0, 0.3]) / np.linalg.norm([1.0, 1.0, 0.3])
    pp = g.assign_points([e[0]], [e[1]], [e[2]], g.counts()[1] + 1)
    assert np.all(pp == -1)


def test_bad_arguments_raise(cuda):
    import salfort as S
    with pytest.raises(S.SalError):
        S.Geometry([1.0], [0.0], [0.0], 0)
    g = S.Geometry(*fib(100), 10)
    with pytest.raises(S.SalError):
        g.fast_sum(np.ones(100), np.ones(100), 4)  # no list yet
    with pytest.raises(S.SalError):
        g.dtt(-1.0)
