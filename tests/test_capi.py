"""The C ABI library loads and exports every symbol include/salfort.h declares; the input
generators are deterministic.  CPU only (no compute call touches a GPU)."""
import ctypes
import os
import re

import numpy as np

from conftest import ROOT


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "salfort.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sal_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import salfort as S
    L = S.lib()
    names = declared_symbols()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_error_path_without_gpu_calls():
    import salfort as S
    L = S.lib()
    assert L.sal_eval_opts_default(None) == 1
    assert b"null" in L.sal_last_error()
    o = S.default_opts()
    assert o.degree == 8 and o.mode == S.MODE_ACCURATE and o.precision == S.PREC_MIXED


def test_generators_deterministic():
    import salfort as S
    a = S.gen_mpas(5000, seed=3)
    b = S.gen_mpas(5000, seed=3)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    x, y, z, area = a
    assert np.allclose(x * x + y * y + z * z, 1, atol=1e-15)
    assert abs(area.sum() / 3.61e14 - 1) < 1e-12
    s1 = S.gen_ssh(x, y, z, lmax=16, seed=4, R=2)
    s2 = S.gen_ssh(x, y, z, lmax=16, seed=5, R=1)
    assert np.array_equal(s1[1], s2)  # field r uses seed0 + r
    f = S.gen_fibonacci(1000)
    assert np.allclose(f[2], 1 - (2 * np.arange(1000) + 1) / 1000, atol=0)


def test_spherical_harmonics_orthonormal():
    """Y_lm orthonormal under a fine Fibonacci quadrature: a field with one unit coefficient has
    mean square 1/(4 pi)."""
    import salfort as S
    x, y, z = S.gen_fibonacci(200000)
    s = S.gen_ssh(x, y, z, lmax=6, amp=1.0, seed=0, R=1)
    # Parseval: mean(f^2) * 4 pi = sum of squared coefficients
    f2 = np.mean(s ** 2) * 4 * np.pi
    assert 0.1 < f2 < 10
    for (L, expect) in [(0, 1.0 / (4 * np.pi))]:
        t = S.gen_ssh(x, y, z, lmax=0, amp=1.0, seed=0)
        c = t[0] * np.sqrt(4 * np.pi)  # constant field = a00 * Y00
        assert np.allclose(t, c / np.sqrt(4 * np.pi))


def test_mom6_grid_shape():
    import salfort as S
    x, y, z, a = S.gen_mom6(nlon=144, nlat=108, seed=1)
    n = len(x)
    assert 0.69 * 144 * 108 < n < 0.73 * 144 * 108
    lat = np.degrees(np.arcsin(z))
    assert lat.min() > -80 and lat.max() < 90


def test_geometry_cache_rejects_foreign_files_without_gpu_calls(tmp_path):
    """sal_geom_load checks the file's magic, format version and header before it touches the GPU:
    these failures return SAL_ERR_IO (6) and create no handle."""
    import struct
    import salfort as S
    L = S.lib()
    h = ctypes.c_void_p()
    assert L.sal_geom_load(os.fsencode(str(tmp_path / "absent.salgeom")), ctypes.byref(h)) == 6
    assert b"cannot open" in L.sal_last_error() and not h.value
    magic = b"SALGEOM\0"
    for name, blob in {
        "empty": b"",
        "foreign": b"not a geometry cache file at all",
        "version": magic + struct.pack("<II", 99, 0x01020304),
        "byte order": magic + struct.pack("<II", 1, 0x04030201),
        "no header": magic + struct.pack("<II", 1, 0x01020304),
    }.items():
        p = tmp_path / "bad.salgeom"
        p.write_bytes(blob)
        assert L.sal_geom_load(os.fsencode(str(p)), ctypes.byref(h)) == 6, name
        assert not h.value
