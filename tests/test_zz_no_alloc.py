"""The entry point that allocates nothing (sal_direct_sum_dev_noalloc), through the C ABI on the GPU:
two calls in a row, the second on a smaller grid with fewer targets, each bit-identical to the
allocating sal_direct_sum_dev on the same inputs (the same kernels, the same summation order); a
call beyond the static buffer is refused with SAL_ERR_LIMIT before any launch and leaves the
outputs untouched; a call after the refusal still gives the allocating entry point's result."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ERR_LIMIT = 5


def _inputs(n, nt, R, seed, cuda):
    import torch
    import salfort as S
    x, y, z, area = S.gen_mpas(n, seed=seed)
    ssh = S.gen_ssh(x, y, z, lmax=16, amp=0.1, noise=0.02, seed=seed + 1, R=R).reshape(R, n)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=cuda)
    tg = np.sort(np.random.default_rng(seed).choice(n, nt, replace=False)).astype(np.int32)
    return [t(v) for v in (x, y, z, area, ssh)] + [torch.as_tensor(tg, device=cuda)]


def _noalloc(X, mode, gx, gy):
    import torch
    import salfort as S
    x, y, z, area, ssh, tg = X
    p = lambda a: C.c_void_p(a.data_ptr())
    rc = S.lib().sal_direct_sum_dev_noalloc(p(x), p(y), p(z), p(area), p(ssh), x.numel(), ssh.shape[0], p(tg),
                                            tg.numel(), mode, p(gx), p(gy), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return rc


@pytest.mark.parametrize("mode", [0, 2])
def test_noalloc_direct_sum_equals_the_allocating_one(cuda, mode):
    import torch
    import salfort as S
    for n, nt, R, seed in [(60000, 3000, 3, 1), (9000, 700, 2, 5)]:   # the second call is smaller
        X = _inputs(n, nt, R, seed, cuda)
        ex, ey = S.direct_sum(*X, mode)
        gx = torch.full((R, nt), float("nan"), dtype=torch.float64, device=cuda)
        gy = torch.full_like(gx, float("nan"))
        assert _noalloc(X, mode, gx, gy) == 0, S.lib().sal_last_error()
        assert torch.equal(gx, ex) and torch.equal(gy, ey), (n, nt)


def test_noalloc_direct_sum_refuses_sizes_beyond_its_buffer(cuda):
    import torch
    import salfort as S
    X = _inputs(1000000, 70000, 1, 3, cuda)    # 62 chunks x 70000 targets x 2 > 8388608 doubles
    gx = torch.zeros((1, 70000), dtype=torch.float64, device=cuda)
    gy = torch.zeros_like(gx)
    assert _noalloc(X, 2, gx, gy) == ERR_LIMIT
    assert b"scratch" in S.lib().sal_last_error()
    assert not gx.any() and not gy.any()
    Y = _inputs(20000, 1000, 1, 4, cuda)
    hx = torch.zeros((1, 1000), dtype=torch.float64, device=cuda)
    hy = torch.zeros_like(hx)
    assert _noalloc(Y, 2, hx, hy) == 0
    ex, ey = S.direct_sum(*Y, 2)
    assert torch.equal(hx, ex) and torch.equal(hy, ey)
