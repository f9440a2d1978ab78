"""Argument checks of the Python host side, on the CPU: every entry point that hands raw pointers to
the library raises ValueError for a wrong dtype, device, layout or size before any library call (a
float32 field, a host tensor where a device one is expected or a short output would otherwise make
the kernels read or write out of bounds)."""
import numpy as np
import pytest
import torch


def _geom(n):
    """A handle-less Geometry: the checks run before any library call, so no GPU is needed."""
    import salfort as S
    g = S.Geometry.__new__(S.Geometry)
    g.h, g.n, g.thresh, g.n_stt = None, n, 65, n
    return g


def test_eval_checks_dtype_device_and_shape():
    g = _geom(100)
    a = torch.zeros(100, dtype=torch.float64)
    s = torch.zeros((3, 100), dtype=torch.float64)
    o = torch.zeros((3, 100), dtype=torch.float64)
    with pytest.raises(ValueError, match="CUDA"):
        g.eval(a, s, o, o.clone())                      # host tensors where device ones are expected
    with pytest.raises(ValueError, match="float64"):
        g.eval(a, s.float(), o, o)
    with pytest.raises(ValueError, match="shape"):
        g.eval(a, torch.zeros((3, 99), dtype=torch.float64), o, o)
    with pytest.raises(ValueError, match="ssh"):
        g.eval(a, torch.zeros((2, 3, 100), dtype=torch.float64), o, o)
    with pytest.raises(ValueError):
        g.eval_work(None, a, s, o, o, None)


def test_host_sums_check_lengths_and_outputs():
    g = _geom(50)
    with pytest.raises(ValueError, match="50"):
        g.fast_sum(np.zeros(49), np.zeros(50), 4)
    with pytest.raises(ValueError, match="gx"):
        g.fast_sum(np.zeros(50), np.zeros(50), 4, np.zeros(40), np.zeros(50))
    with pytest.raises(ValueError, match="gx"):
        g.fast_sum(np.zeros(50), np.zeros(50), 4, np.zeros(50, np.float32), np.zeros(50))
    with pytest.raises(ValueError):
        g.mom6_sums(np.zeros(10), np.zeros(50), 4)
    with pytest.raises(ValueError):
        g.mom6_pp_compute(np.zeros(50), np.zeros(50), np.zeros(3), np.zeros(50))


def test_direct_sum_and_gradient_checks():
    import salfort as S
    x = torch.zeros(10, dtype=torch.float64)
    t = torch.zeros(3, dtype=torch.int32)
    with pytest.raises(ValueError, match="CUDA"):
        S.direct_sum(x, x, x, x, x, t, 2)
    with pytest.raises(ValueError, match="CUDA"):
        S.gradient_xyz(x, x, x, x, x)
    with pytest.raises(ValueError):
        S.Geometry(device_xyz=(x, x, x[:5]))


def test_server_submit_checks_shapes():
    import salfort as S
    srv = S.Server.__new__(S.Server)
    srv.h, srv.geoms = None, [_geom(20)]
    with pytest.raises(ValueError, match="no geometry"):
        srv.submit(1, np.zeros((2, 20)), np.zeros((2, 20)), np.zeros((2, 20)))
    with pytest.raises(ValueError, match="shape"):
        srv.submit(0, np.zeros((2, 19)), np.zeros((2, 19)), np.zeros((2, 19)))
    with pytest.raises(ValueError, match="gx"):
        srv.submit(0, np.zeros((2, 20)), np.zeros((1, 20)), np.zeros((2, 20)))
    with pytest.raises(ValueError, match="host"):
        S._host_ptr(np.zeros((2, 20), np.float32))
