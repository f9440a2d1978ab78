"""Test infrastructure: ctypes bindings to the oracle libraries.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and only
as the checker.  Two back ends with one Python interface:

  Oracle()     the C restatement of the reference algorithm (oracle/liboracle.so)
  Reference()  the reference's own Fortran modules behind a bind(C) bridge (oracle/_ref/libsalref.so)

Both return 0-based ids (-1 = none) so results compare directly.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsalref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i64 = C.c_int64
_d = C.c_double
_i = C.c_int


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _tree_dict(level, is_leaf, parent, child, face, bounds, radius, count, offsets, points):
    return dict(level=level, is_leaf=is_leaf, parent=parent, child=child, face=face,
                bounds=bounds, radius=radius, count=count, offsets=offsets, points=points)


def _to0(a):
    """1-based reference ids to 0-based; -1 stays -1."""
    a = np.asarray(a, dtype=np.int32)
    return np.where(a > 0, a - 1, a).astype(np.int32)


class Oracle:
    """C restatement (oracle/sal_oracle.c)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle oracle)")
        L = self.L = C.CDLL(path)
        L.or_ctx_new.restype = C.c_void_p
        L.or_ctx_free.argtypes = [C.c_void_p]
        L.or_face.argtypes = [_d, _d, _d]
        L.or_face.restype = _i
        L.or_xieta.argtypes = [_d, _d, _d, _i, C.POINTER(_d), C.POINTER(_d)]
        L.or_xyz.argtypes = [_d, _d, _i, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d)]
        L.or_bli_coeffs.argtypes = [_i, _dp]
        L.or_bli_points.argtypes = [_d, _d, _i, _dp]
        L.or_interp_vals.argtypes = [_d, _d, _d, _d, _d, _d, _i, _dp]
        L.or_gfunc.argtypes = [_d] * 6 + [C.POINTER(_d), C.POINTER(_d)]
        L.or_tree_build.argtypes = [C.c_void_p, _i, _dp, _dp, _dp, _i]
        L.or_tree_build.restype = _i
        L.or_tree_n_panels.argtypes = [C.c_void_p]
        L.or_tree_n_panels.restype = _i
        L.or_tree_total_inside.argtypes = [C.c_void_p]
        L.or_tree_total_inside.restype = _i64
        L.or_tree_export.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _ip, _lp, _ip]
        L.or_assign.argtypes = [C.c_void_p, _i, _dp, _dp, _dp, _i, _ip]
        L.or_dtt.argtypes = [C.c_void_p, _d, _i]
        L.or_dtt.restype = _i64
        L.or_dtt_export.argtypes = [C.c_void_p, _ip, _ip, _ip]
        L.or_fast_sum.argtypes = [C.c_void_p, _i, _dp, _dp, _dp, _dp, _dp, _i, _i, _i, _dp, _dp]
        L.or_interact.argtypes = [C.c_void_p, _i, _i, _i, _dp, _dp, _dp, _dp, _dp, _i, _dp, _dp]
        L.or_stt.argtypes = [C.c_void_p, _d, _i, _i, _dp, _dp, _dp, C.POINTER(_i64), C.POINTER(_i64)]
        L.or_stt_export.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip]
        L.or_proxy_source.argtypes = [C.c_void_p, _i, _i, _ip, _dp, _dp, _dp, _dp, _dp, _i, _dp]
        L.or_mom6_sums.argtypes = [C.c_void_p, _i, _dp, _dp, _dp, _dp, _dp, _i, _dp, _dp, _dp]
        L.or_mom6_sums_sel.argtypes = [C.c_void_p, _i, _dp, _dp, _dp, _dp, _dp, _i, _dp, C.c_void_p, _dp, _dp]
        L.or_fast_sum_cart.argtypes = [C.c_void_p, _i, _dp, _dp, _dp, _dp, _dp, _i, _dp, _dp]
        L.or_direct_sum.argtypes = [_i, _dp, _dp, _dp, _dp, _dp, _i, _ip, _i, _dp, _dp]
        self.ctx = L.or_ctx_new()
        self.n_panels = 0

    def __del__(self):
        try:
            self.L.or_ctx_free(self.ctx)
        except Exception:
            pass

    # geometry -----------------------------------------------------------------------------
    def face(self, x, y, z):
        return self.L.or_face(x, y, z)

    def xieta(self, x, y, z, face=0):
        a, b = _d(), _d()
        self.L.or_xieta(x, y, z, face, C.byref(a), C.byref(b))
        return a.value, b.value

    def xyz(self, xi, eta, face):
        a, b, c = _d(), _d(), _d()
        self.L.or_xyz(xi, eta, face, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def bli_coeffs(self, d):
        c = np.zeros(d + 1)
        self.L.or_bli_coeffs(d, c)
        return c

    def bli_points(self, a, b, d):
        p = np.zeros(d + 1)
        self.L.or_bli_points(a, b, d, p)
        return p

    def interp_vals(self, xi, eta, a, b, c, d, deg):
        v = np.zeros((deg + 1) * (deg + 1))
        self.L.or_interp_vals(xi, eta, a, b, c, d, deg, v)
        return v.reshape(deg + 1, deg + 1).T  # [i_xi, j_eta]

    def gfunc(self, t, s):
        a, b = _d(), _d()
        self.L.or_gfunc(*t, *s, C.byref(a), C.byref(b))
        return a.value, b.value

    # tree ----------------------------------------------------------------------------------
    def tree_build(self, x, y, z, thresh):
        x, y, z = _f64(x), _f64(y), _f64(z)
        self.n_panels = self.L.or_tree_build(self.ctx, len(x), x, y, z, thresh)
        return self.n_panels

    def tree_export(self):
        P = self.L.or_tree_n_panels(self.ctx)
        tot = self.L.or_tree_total_inside(self.ctx)
        a = [np.zeros(P, np.int32) for _ in range(3)]
        child = np.zeros(4 * P, np.int32)
        face = np.zeros(P, np.int32)
        bounds = np.zeros(6 * P)
        radius = np.zeros(P)
        count = np.zeros(P, np.int32)
        offsets = np.zeros(P + 1, np.int64)
        pts = np.zeros(max(tot, 1), np.int32)
        self.L.or_tree_export(self.ctx, a[0], a[1], a[2], child, face, bounds, radius, count, offsets, pts)
        return _tree_dict(a[0], a[1], a[2], child.reshape(P, 4), face, bounds.reshape(P, 6), radius,
                          count, offsets, pts[:tot])

    def assign(self, x, y, z, nlev):
        x, y, z = _f64(x), _f64(y), _f64(z)
        pp = np.full(len(x) * nlev, -1, np.int32)
        self.L.or_assign(self.ctx, len(x), x, y, z, nlev, pp)
        return pp.reshape(len(x), nlev)

    def dtt(self, theta, thresh):
        n = self.L.or_dtt(self.ctx, theta, thresh)
        it, is_, ty = (np.zeros(max(n, 1), np.int32) for _ in range(3))
        self.L.or_dtt_export(self.ctx, it, is_, ty)
        return it[:n], is_[:n], ty[:n]

    def fast_sum(self, x, y, z, area, ssh, degree, rank=0, nproc=1):
        x, y, z, area, ssh = map(_f64, (x, y, z, area, ssh))
        gx, gy = np.zeros(len(x)), np.zeros(len(x))
        self.L.or_fast_sum(self.ctx, len(x), x, y, z, area, ssh, degree, rank, nproc, gx, gy)
        return gx, gy

    def fast_sum_cart(self, x, y, z, area, ssh, degree):
        x, y, z, area, ssh = map(_f64, (x, y, z, area, ssh))
        gx, gy = np.zeros(len(x)), np.zeros(len(x))
        self.L.or_fast_sum_cart(self.ctx, len(x), x, y, z, area, ssh, degree, gx, gy)
        return gx, gy

    def interact(self, kind, it, is_, x, y, z, area, ssh, degree):
        x, y, z, area, ssh = map(_f64, (x, y, z, area, ssh))
        gx, gy = np.zeros(len(x)), np.zeros(len(x))
        self.L.or_interact(self.ctx, kind, it, is_, x, y, z, area, ssh, degree, gx, gy)
        return gx, gy

    def stt(self, theta, thresh, x, y, z):
        x, y, z = _f64(x), _f64(y), _f64(z)
        npc, npp = _i64(), _i64()
        self.L.or_stt(self.ctx, theta, thresh, len(x), x, y, z, C.byref(npc), C.byref(npp))
        a = [np.zeros(max(npc.value, 1), np.int32) for _ in range(2)]
        b = [np.zeros(max(npp.value, 1), np.int32) for _ in range(2)]
        self.L.or_stt_export(self.ctx, a[0], a[1], b[0], b[1])
        return a[0][:npc.value], a[1][:npc.value], b[0][:npp.value], b[1][:npp.value]

    def proxy_source(self, pp, x, y, z, ssh, area, degree):
        x, y, z, ssh, area = map(_f64, (x, y, z, ssh, area))
        pp = np.ascontiguousarray(pp, dtype=np.int32)
        w = np.zeros((degree + 1) ** 2 * self.L.or_tree_n_panels(self.ctx))
        self.L.or_proxy_source(self.ctx, len(x), pp.shape[1], pp.reshape(-1), x, y, z, ssh, area, degree, w)
        return w

    def mom6_sums(self, x, y, z, area, ssh, degree, w, targets=None):
        """All targets' sums, or (targets given) only those targets' (zeros elsewhere)."""
        x, y, z, area, ssh, w = map(_f64, (x, y, z, area, ssh, w))
        gx, gy = np.zeros(len(x)), np.zeros(len(x))
        if targets is None:
            self.L.or_mom6_sums(self.ctx, len(x), x, y, z, area, ssh, degree, w, gx, gy)
        else:
            sel = np.zeros(len(x), np.uint8)
            sel[np.asarray(targets)] = 1
            self.L.or_mom6_sums_sel(self.ctx, len(x), x, y, z, area, ssh, degree, w, sel.ctypes.data, gx, gy)
        return gx, gy

    def direct_sum(self, x, y, z, area, ssh, targets, mode):
        x, y, z, area, ssh = map(_f64, (x, y, z, area, ssh))
        t = np.ascontiguousarray(targets, dtype=np.int32)
        gx, gy = np.zeros(len(t)), np.zeros(len(t))
        self.L.or_direct_sum(len(x), x, y, z, area, ssh, len(t), t, mode, gx, gy)
        return gx, gy


class Reference:
    """The reference's own Fortran modules (oracle/_ref/libsalref.so, built by oracle/Makefile)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        L = self.L = C.CDLL(path)
        L.ref_tree_build.argtypes = [_i, _dp, _dp, _dp, _i, C.POINTER(_i)]
        L.ref_tree_total_inside.argtypes = [C.POINTER(_i64)]
        L.ref_tree_export.argtypes = [_ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _ip, _lp, _ip]
        L.ref_dtt.argtypes = [_d, _i, C.POINTER(_i64)]
        L.ref_dtt_export.argtypes = [_ip, _ip, _ip]
        L.ref_fast_sum.argtypes = [_i, _dp, _dp, _dp, _dp, _dp, _i, _i, _i, _dp, _dp]
        L.ref_fast_sum_range.argtypes = [_i, _dp, _dp, _dp, _dp, _dp, _i, _i64, _i64, _dp, _dp]
        L.ref_interact.argtypes = [_i, _i, _i, _i, _dp, _dp, _dp, _dp, _dp, _i, _dp, _dp]
        L.ref_assign.argtypes = [_i, _dp, _dp, _dp, _i, _ip]
        L.ref_stt.argtypes = [_d, _i, _i, _dp, _dp, _dp, C.POINTER(_i64), C.POINTER(_i64)]
        L.ref_stt_export.argtypes = [_ip, _ip, _ip, _ip]
        L.ref_proxy_source.argtypes = [_i, _i, _ip, _dp, _dp, _dp, _dp, _dp, _i, _dp]
        L.ref_mom6_sums.argtypes = [_i, _dp, _dp, _dp, _dp, _dp, _i, _dp, _dp, _dp]
        L.ref_gfunc.argtypes = [_i] + [_d] * 6 + [C.POINTER(_d), C.POINTER(_d)]
        L.ref_bli_coeffs.argtypes = [_i, _dp]
        L.ref_bli_points.argtypes = [_d, _d, _i, _dp]
        L.ref_interp_vals.argtypes = [_d, _d, _d, _d, _d, _d, _i, _dp]
        L.ref_face.argtypes = [_d, _d, _d]
        L.ref_face.restype = _i
        L.ref_xieta.argtypes = [_d, _d, _d, _i, C.POINTER(_d), C.POINTER(_d)]
        L.ref_xyz.argtypes = [_d, _d, _i, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d)]
        L.ref_flush.argtypes = []
        self.n_panels = 0

    def flush(self):
        """Flushes the reference's buffered standard output."""
        self.L.ref_flush()

    def face(self, x, y, z):
        return self.L.ref_face(x, y, z)

    def xieta(self, x, y, z, face=0):
        a, b = _d(), _d()
        self.L.ref_xieta(x, y, z, face, C.byref(a), C.byref(b))
        return a.value, b.value

    def xyz(self, xi, eta, face):
        a, b, c = _d(), _d(), _d()
        self.L.ref_xyz(xi, eta, face, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def bli_coeffs(self, d):
        c = np.zeros(d + 1)
        self.L.ref_bli_coeffs(d, c)
        return c

    def bli_points(self, a, b, d):
        p = np.zeros(d + 1)
        self.L.ref_bli_points(a, b, d, p)
        return p

    def interp_vals(self, xi, eta, a, b, c, d, deg):
        v = np.zeros((deg + 1) * (deg + 1))
        self.L.ref_interp_vals(xi, eta, a, b, c, d, deg, v)
        return v.reshape(deg + 1, deg + 1).T

    def gfunc(self, t, s, which=0):
        a, b = _d(), _d()
        self.L.ref_gfunc(which, *t, *s, C.byref(a), C.byref(b))
        return a.value, b.value

    def tree_build(self, x, y, z, thresh):
        x, y, z = _f64(x), _f64(y), _f64(z)
        npan = _i()
        self.L.ref_tree_build(len(x), x, y, z, thresh, C.byref(npan))
        self.n_panels = npan.value
        return self.n_panels

    def tree_export(self):
        P = self.n_panels
        tot = _i64()
        self.L.ref_tree_total_inside(C.byref(tot))
        tot = tot.value
        level, is_leaf, ids, parent, face, count = (np.zeros(P, np.int32) for _ in range(6))
        child = np.zeros(4 * P, np.int32)
        bounds = np.zeros(6 * P)
        radius = np.zeros(P)
        offsets = np.zeros(P + 1, np.int64)
        pts = np.zeros(max(tot, 1), np.int32)
        self.L.ref_tree_export(level, is_leaf, ids, parent, child, face, bounds, radius, count, offsets, pts)
        assert np.array_equal(ids, np.arange(1, P + 1)), "reference panel ids are not 1..P"
        return _tree_dict(level, is_leaf, _to0(parent), _to0(child).reshape(P, 4), face,
                          bounds.reshape(P, 6), radius, count, offsets, (pts[:tot] - 1).astype(np.int32))

    def assign(self, x, y, z, nlev):
        x, y, z = _f64(x), _f64(y), _f64(z)
        pp = np.full(len(x) * nlev, -1, np.int32)
        self.L.ref_assign(len(x), x, y, z, nlev, pp)
        return _to0(pp).reshape(len(x), nlev)

    def dtt(self, theta, thresh):
        n = _i64()
        self.L.ref_dtt(theta, thresh, C.byref(n))
        n = n.value
        it, is_, ty = (np.zeros(max(n, 1), np.int32) for _ in range(3))
        self.L.ref_dtt_export(it, is_, ty)
        return it[:n] - 1, is_[:n] - 1, ty[:n]

    def fast_sum(self, x, y, z, area, ssh, degree, rank=0, nproc=1):
        x, y, z, area, ssh = map(_f64, (x, y, z, area, ssh))
        gx, gy = np.zeros(len(x)), np.zeros(len(x))
        self.L.ref_fast_sum(len(x), x, y, z, area, ssh, degree, rank, nproc, gx, gy)
        return gx, gy

    def fast_sum_range(self, x, y, z, area, ssh, degree, first, last, gx=None, gy=None):
        """Interactions first..last (0-based, inclusive of first, exclusive of last)."""
        x, y, z, area, ssh = map(_f64, (x, y, z, area, ssh))
        if gx is None:
            gx, gy = np.zeros(len(x)), np.zeros(len(x))
        self.L.ref_fast_sum_range(len(x), x, y, z, area, ssh, degree, first + 1, last, gx, gy)
        return gx, gy

    def interact(self, kind, it, is_, x, y, z, area, ssh, degree):
        x, y, z, area, ssh = map(_f64, (x, y, z, area, ssh))
        gx, gy = np.zeros(len(x)), np.zeros(len(x))
        self.L.ref_interact(kind, it + 1, is_ + 1, len(x), x, y, z, area, ssh, degree, gx, gy)
        return gx, gy

    def stt(self, theta, thresh, x, y, z):
        x, y, z = _f64(x), _f64(y), _f64(z)
        npc, npp = _i64(), _i64()
        self.L.ref_stt(theta, thresh, len(x), x, y, z, C.byref(npc), C.byref(npp))
        a = [np.zeros(max(npc.value, 1), np.int32) for _ in range(2)]
        b = [np.zeros(max(npp.value, 1), np.int32) for _ in range(2)]
        self.L.ref_stt_export(a[0], a[1], b[0], b[1])
        return (a[0][:npc.value] - 1, a[1][:npc.value] - 1, b[0][:npp.value] - 1, b[1][:npp.value] - 1)

    def proxy_source(self, pp, x, y, z, ssh, area, degree):
        """pp: 0-based (n, nlev) as returned by assign()."""
        x, y, z, ssh, area = map(_f64, (x, y, z, ssh, area))
        pp1 = np.where(pp >= 0, pp + 1, -1).astype(np.int32)
        pp1 = np.ascontiguousarray(pp1)
        w = np.zeros((degree + 1) ** 2 * self.n_panels)
        self.L.ref_proxy_source(len(x), pp1.shape[1], pp1.reshape(-1), x, y, z, ssh, area, degree, w)
        return w

    def mom6_sums(self, x, y, z, area, ssh, degree, w):
        x, y, z, area, ssh, w = map(_f64, (x, y, z, area, ssh, w))
        gx, gy = np.zeros(len(x)), np.zeros(len(x))
        self.L.ref_mom6_sums(len(x), x, y, z, area, ssh, degree, w, gx, gy)
        return gx, gy
