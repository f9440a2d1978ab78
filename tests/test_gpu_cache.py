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


# tags of the cache format's sections (sal-fort_amd/csrc/cache.cpp)
T_HEADER, T_LEVEL, T_LEAF, T_PARENT, T_CHILD1, T_COUNT, T_LEVEL_LO, T_START = 1, 9, 10, 11, 12, 14, 15, 19


def _sections(data):
    """[(tag, element size, count, offset of the elements)] of a cache file, END last."""
    out, off = [], 16
    while True:
        tag, es, cnt = np.frombuffer(data, dtype=np.uint32, count=2, offset=off).tolist() + \
            [int(np.frombuffer(data, dtype=np.uint64, count=1, offset=off + 8)[0])]
        out.append((tag, es, cnt, off + 16))
        if tag == 0xE0D:
            return out
        off += 16 + es * cnt


def _rehash(data):
    """The file with its FNV-1a 64 hash recomputed over every byte before the hash."""
    h = 1469598103934665603
    for b in data[:-8]:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return data[:-8] + int(h).to_bytes(8, "little")


def _edit(data, tag, fn):
    """Applies fn to the int32 / int64 elements of section `tag` and returns the rehashed file."""
    for t, es, cnt, off in _sections(data):
        if t == tag:
            dt = np.int64 if es == 8 else np.int32     # a record (the header) as int32 words
            a = np.frombuffer(data, dtype=dt, count=cnt * es // np.dtype(dt).itemsize, offset=off).copy()
            fn(a)
            return _rehash(data[:off] + a.tobytes() + data[off + a.nbytes:])
    raise KeyError(tag)


def test_consistently_wrong_trees_are_rejected(cuda, tmp_path):
    """Files that pass the checksum but whose panel arrays contradict one another (a leaf with
    children, a non-leaf without, children of another parent, counts that do not add up, a panel
    outside its level's id range, level ranges out of order, a leaf with too many points) are each
    rejected with SAL_ERR_IO; the unedited file, rehashed, still loads."""
    import salfort as S
    x, y, z, _, _ = _grid(4000, seed=10)
    g = S.Geometry(x, y, z, 33)
    g.dtt(0.7)
    path = str(tmp_path / "g.salgeom")
    g.save(path)
    data = open(path, "rb").read()
    T = g.export()
    leaf = int(np.where((T["is_leaf"] == 1) & (np.arange(len(T["level"])) >= 6))[0][0])
    inner = int(np.where(T["is_leaf"] == 0)[0][0])
    kid = int(T["child"][inner][0])
    lvl = int(T["level"][leaf])

    def header_thresh(a):   # the header as int32 words: thresh is the first int32 after 11 int64
        a[22] = 2
    cases = {
        "leaf with children": (T_CHILD1, lambda a: a.__setitem__(leaf, kid)),
        "non-leaf without children": (T_CHILD1, lambda a: a.__setitem__(inner, -1)),
        "child of another parent": (T_PARENT, lambda a: a.__setitem__(kid, leaf)),
        "counts do not add up": (T_COUNT, lambda a: a.__setitem__(kid, a[kid] + 1)),
        "panel outside its level": (T_LEVEL, lambda a: a.__setitem__(leaf, lvl + 1)),
        "level ranges out of order": (T_LEVEL_LO, lambda a: a.__setitem__(1, a[2] + 1)),
        "leaf over the threshold": (T_HEADER, header_thresh),
    }
    ok = _rehash(data)
    p = str(tmp_path / "ok.salgeom")
    with open(p, "wb") as f:
        f.write(ok)
    S.Geometry.load(p).close()
    for name, (tag, fn) in cases.items():
        p = str(tmp_path / "bad.salgeom")
        with open(p, "wb") as f:
            f.write(_edit(data, tag, fn))
        with pytest.raises(S.SalError) as e:
            S.Geometry.load(p)
        assert e.value.code == ERR_IO, name
