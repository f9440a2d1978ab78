"""On-disk geometry cache: a handle read back from a file written by sal_geom_save has the saved
handle's tree, lists and sums bit for bit, and a damaged file creates no handle."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ERR_IO = 6


def _grid(n=20000, seed=41):
    import salfort as S
    x, y, z, area = S.gen_mpas(n, seed=seed)
    ssh = S.gen_ssh(x, y, z, lmax=16, amp=0.1, noise=0.02, seed=seed + 1, R=2)
    return x, y, z, area, ssh


def _same(a, b):
    return all(np.array_equal(u, v) for u, v in zip(a, b))


def test_loaded_geometry_is_identical(cuda, tmp_path):
    import salfort as S
    import torch
    x, y, z, area, ssh = _grid()
    g = S.Geometry(x, y, z, 33)
    g.dtt(0.7)
    tg = np.arange(0, len(x), 7)
    g.stt(0.7, x[tg], y[tg], z[tg])
    path = str(tmp_path / "mpas.salgeom")
    g.save(path)
    assert not os.path.exists(path + ".tmp")
    h = S.Geometry.load(path)
    assert (h.n, h.thresh, h.n_interactions) == (g.n, g.thresh, g.n_interactions)
    assert h.counts() == g.counts() and h.ties() == g.ties()
    tg_, th = g.export(), h.export()
    for k in tg_:
        assert np.array_equal(tg_[k], th[k]), k
    assert _same(g.dtt_export(), h.dtt_export())
    assert np.array_equal(g.dtt_counts(), h.dtt_counts())
    assert _same(g.stt_export(), h.stt_export())
    nlev = g.counts()[1] + 1
    assert np.array_equal(g.assign_points(x, y, z, nlev), h.assign_points(x, y, z, nlev))
    # sums: FP64 reference semantics, the MOM6 per-point sums, and the mixed batched path
    a = g.fast_sum(area, ssh[0], 6)
    b = h.fast_sum(area, ssh[0], 6)
    assert _same(a, b)
    # the MOM6 sums run over the STT targets; both handles hold the same ones
    assert _same(g.mom6_sums(area, ssh[0], 5), h.mom6_sums(area, ssh[0], 5))
    X = [torch.as_tensor(np.ascontiguousarray(v), device=cuda) for v in (area, ssh)]
    outs = []
    for G in (g, h):
        gx = torch.zeros((2, len(x)), dtype=torch.float64, device=cuda)
        gy = torch.zeros_like(gx)
        o = S.default_opts(degree=7, mode=S.MODE_ACCURATE, precision=S.PREC_MIXED)
        G.eval(X[0], X[1], gx, gy, o)
        torch.cuda.synchronize()
        outs.append((gx.cpu().numpy(), gy.cpu().numpy()))
    assert _same(outs[0], outs[1])
    # a loaded handle saves the same bytes again
    path2 = str(tmp_path / "again.salgeom")
    h.save(path2)
    assert open(path, "rb").read() == open(path2, "rb").read()


def test_geometry_without_lists_round_trips(cuda, tmp_path):
    import salfort as S
    x, y, z, _, _ = _grid(5000, seed=7)
    g = S.Geometry(x, y, z, 65)
    path = str(tmp_path / "tree.salgeom")
    g.save(path)
    h = S.Geometry.load(path)
    tg_, th = g.export(), h.export()
    for k in tg_:
        assert np.array_equal(tg_[k], th[k]), k
    assert h.n_interactions == 0
    h.dtt(0.7)
    g.dtt(0.7)
    assert _same(g.dtt_export(), h.dtt_export())


def test_damaged_files_are_rejected(cuda, tmp_path):
    import salfort as S
    x, y, z, _, _ = _grid(5000, seed=9)
    g = S.Geometry(x, y, z, 33)
    g.dtt(0.7)
    path = str(tmp_path / "g.salgeom")
    g.save(path)
    data = open(path, "rb").read()
    cases = {
        "flipped byte": data[:len(data) // 2] + bytes([data[len(data) // 2] ^ 1]) + data[len(data) // 2 + 1:],
        "truncated": data[:len(data) - 100],
        "trailing": data + b"\0",
        "bad magic": b"X" + data[1:],
    }
    for name, blob in cases.items():
        p = str(tmp_path / "bad.salgeom")
        with open(p, "wb") as f:
            f.write(blob)
        with pytest.raises(S.SalError) as e:
            S.Geometry.load(p)
        assert e.value.code == ERR_IO, name
