"""Serving loop and workspaces on the GPU.

* each request's result is the same alone (its own evaluation in a fresh workspace) and in the mix
  (merged with other requests of its geometry, overlapping with other geometries' batches on other
  slots); stated tolerance: relative max difference <= 1e-12 (the kernels compute every field's
  column with the same operation order, so the results are expected to be bit-identical);
* admission by device memory: a budget that holds one batch at a time still completes every
  request without exceeding it, and a request that can never fit fails with SAL_ERR_LIMIT;
* a captured HIP graph replayed several times, with eager work in between, gives the eager result.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = {"fibonacci": 20000, "mpas": 60000}


@pytest.fixture(scope="module")
def geos(cuda):
    import serve
    return serve.build(["fibonacci", "mom6", "mpas"], degree=8, thresh=65, sizes=SIZES, mom6_shape=(360, 270),
                       out_bufs={"fibonacci": 18, "mom6": 18, "mpas": 18})


def _alone(G, row, R, cuda):
    import torch
    import salfort as S
    ws = S.Workspace()
    dx, dy, dz, da = G["dev"]
    ssh = G["pool"][row:row + R].to(cuda)
    gx = torch.empty_like(ssh)
    gy = torch.empty_like(ssh)
    G["g"].eval_work(ws, da, ssh, gx, gy, G["opts"])
    torch.cuda.synchronize()
    ws.close()
    return gx.cpu().numpy(), gy.cpu().numpy()


def test_same_alone_and_in_mix(cuda, geos):
    import salfort as S
    srv = S.Server(slots=3, max_fields=32, batching=True)
    import serve
    serve.register(srv, geos)
    rng = np.random.default_rng(5)
    reqs = []
    # a burst: every request is queued while the first ones run
    for i in range(18):
        gi = int(rng.integers(0, len(geos)))
        G = geos[gi]
        R = int(rng.integers(1, 13))
        row = int(rng.integers(0, serve.POOL_FIELDS - R + 1))
        buf = G["free"].pop()
        gx, gy = G["outs"][buf]
        sid = srv.submit(G["gid"], G["pool"][row:row + R], gx[:R], gy[:R])
        reqs.append((gi, R, row, buf, sid))
    identical = 0
    for gi, R, row, buf, sid in reqs:
        assert srv.wait(sid, 120.0)
        G = geos[gi]
        gx, gy = G["outs"][buf]
        mx, my = gx[:R].numpy().copy(), gy[:R].numpy().copy()
        ax, ay = _alone(G, row, R, cuda)
        scale = max(np.abs(ax).max(), np.abs(ay).max())
        err = max(np.abs(mx - ax).max(), np.abs(my - ay).max()) / scale
        assert err <= 1e-12, (G["name"], R, row, err)
        identical += int(np.array_equal(mx, ax) and np.array_equal(my, ay))
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["rejected"] == 0
    assert st["merged_requests"] > 0, st        # requests were batched together
    assert st["peak_busy_slots"] >= 2, st       # and batches overlapped on the device
    print(f"{identical}/{len(reqs)} bit-identical; stats {st}")


def test_admission_by_memory(cuda, geos):
    import salfort as S
    import serve
    G = geos[2]  # mpas-shaped
    R = 8
    need = G["g"].prepare(G["opts"], R) + 3 * 8 * G["n"] * R
    srv = S.Server(slots=3, max_fields=32, batching=False, hbm_budget=int(need * 1.5))
    serve.register(srv, [G])
    ids = []
    for k in range(4):
        gx, gy = G["outs"][k]
        ids.append(srv.submit(G["gid"], G["pool"][k:k + R], gx[:R], gy[:R]))
    for sid in ids:
        assert srv.wait(sid, 120.0)
    st = srv.stats()
    assert st["completed"] == 4 and st["peak_held_bytes"] <= st["budget_bytes"], st
    assert st["peak_busy_slots"] == 1, st  # one batch's memory at a time
    # more fields than the budget can ever hold
    gx, gy = G["outs"][4]
    big = srv.submit(G["gid"], G["pool"][0:32], gx, gy)
    with pytest.raises(S.SalError) as e:
        srv.wait(big, 120.0)
    assert e.value.code == 5  # SAL_ERR_LIMIT
    srv.close()
    # the server's results equal the workspace path
    ax, ay = _alone(G, 0, R, cuda)
    assert np.array_equal(G["outs"][0][0][:R].numpy(), ax)


def test_graph_replay_matches_eager(cuda, geos):
    import torch
    import salfort as S
    G = geos[1]
    dx, dy, dz, da = G["dev"]
    ssh = G["pool"][0:5].to(cuda)
    ref_x = torch.empty_like(ssh)
    ref_y = torch.empty_like(ssh)
    o = G["opts"]
    ws = S.Workspace()
    G["g"].eval_work(ws, da, ssh, ref_x, ref_y, o)
    torch.cuda.synchronize()
    og = S.default_opts(**{f: getattr(o, f) for f, _ in S.EvalOpts._fields_})  # the geometry's setting
    og.graph = 1
    gx = torch.empty_like(ssh)
    gy = torch.empty_like(ssh)
    for k in range(4):
        gx.fill_(float("nan"))
        gy.fill_(float("nan"))
        G["g"].eval_work(ws, da, ssh, gx, gy, og)  # first call captures, then replays
        torch.cuda.synchronize()
        assert torch.equal(gx, ref_x) and torch.equal(gy, ref_y), k
        # eager work between replays: an evaluation of other fields in the same workspace
        G["g"].eval_work(ws, da, G["pool"][5:7].to(cuda), torch.empty((2, G["n"]), dtype=torch.float64,
                         device=cuda), torch.empty((2, G["n"]), dtype=torch.float64, device=cuda), o)
        torch.cuda.synchronize()
    ws.close()


def test_serving_load_accuracy(cuda, geos):
    """A short Poisson load: every request completes, sampled results meet relL2 <= 1e-6 against the
    FP64 direct sum, and the latency percentiles are reported."""
    import salfort as S
    import serve
    for G in geos:
        G["free"] = list(range(len(G["outs"])))
    srv = S.Server(slots=3, batching=True)
    serve.register(srv, geos)
    recs, summary = serve.run_load(srv, geos, rate=40.0, n_requests=30, seed=3, check_every=3)
    srv.close()
    assert summary["stats"]["completed"] == 30
    assert summary["checked"] >= 10 and summary["relL2_max"] <= 1e-6, summary
    assert summary["p99_ms"] >= summary["p50_ms"] > 0
    print(summary)
