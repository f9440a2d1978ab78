"""Python host side of the salfort treecode: ctypes over the C ABI in include/salfort.h.

It mirrors the reference's operations (tree_module.f90, fast_sum_module.f90,
mom6_fast_sum_module.f90) for Python callers and for the tests:

    g = Geometry(x, y, z, cluster_thresh)          # tree_traverse
    g.dtt(theta)                                   # DUAL_TREE_TRAVERSAL
    gx, gy = g.fast_sum(area, ssh, degree)         # fast_sum (FP64, reference semantics)
    g.eval(area, ssh, out_x, out_y, opts)          # device tensors, R fields, any mode/precision

Device memory and streams come from torch; the library never falls back to the CPU: loading
fails loudly if libsalfort.so has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsalfort.so")

MODE_REFERENCE, MODE_ACCURATE = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
M2L_BF16X3, M2L_FP32_MFMA, M2L_BF16X6 = 1, 2, 3


class SalError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"salfort error {code}: {msg}")
        self.code = code


class EvalOpts(C.Structure):
    _fields_ = [("degree", C.c_int), ("mode", C.c_int), ("precision", C.c_int), ("m2l", C.c_int),
                ("rhs_tile", C.c_int), ("graph", C.c_int), ("store", C.c_int), ("near64", C.c_int)]


class ServerOpts(C.Structure):
    _fields_ = [("slots", C.c_int), ("max_fields", C.c_int), ("batching", C.c_int), ("graph", C.c_int),
                ("device", C.c_int), ("hbm_budget", C.c_int64), ("hbm_reserve", C.c_int64)]


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i64 = C.c_int64


def lib():
    """The loaded C library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python sal-fort_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        L.sal_last_error.restype = C.c_char_p
        sig = {
            "sal_geom_create": [_dp, _dp, _dp, _i64, C.c_int, C.POINTER(_vp)],
            "sal_geom_create_dev": [_vp, _vp, _vp, _i64, C.c_int, _vp, C.POINTER(_vp)],
            "sal_geom_destroy": [_vp],
            "sal_geom_counts": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(_i64)],
            "sal_geom_export": [_vp, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _ip, _lp, _ip],
            "sal_geom_ties": [_vp, C.POINTER(_i64), C.POINTER(_i64)],
            "sal_geom_device_bytes": [_vp, C.POINTER(_i64)],
            "sal_geom_save": [_vp, C.c_char_p],
            "sal_geom_load": [C.c_char_p, C.POINTER(_vp)],
            "sal_geom_size": [_vp, C.POINTER(_i64), C.POINTER(C.c_int)],
            "sal_set_tie_ulps": [C.c_double],
            "sal_assign_points": [_vp, _dp, _dp, _dp, _i64, C.c_int, _ip],
            "sal_dtt": [_vp, C.c_double, C.c_int, C.POINTER(_i64)],
            "sal_interact": [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_int],
            "sal_dtt_counts": [_vp, _lp],
            "sal_dtt_export": [_vp, _ip, _ip, _ip],
            "sal_dtt_import": [_vp, _ip, _ip, _ip, _i64],
            "sal_stt": [_vp, C.c_double, _dp, _dp, _dp, _i64, C.POINTER(_i64), C.POINTER(_i64)],
            "sal_stt_export": [_vp, _ip, _ip, _ip, _ip],
            "sal_stt_counts": [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)],
            "sal_fast_sum": [_vp, _dp, _dp, C.c_int, _dp, _dp, C.c_int],
            "sal_eval_opts_default": [C.POINTER(EvalOpts)],
            "sal_eval_dev": [_vp, C.POINTER(EvalOpts), _vp, _vp, C.c_int, _vp, _vp, C.c_int, _vp],
            "sal_work_create": [C.POINTER(_vp)],
            "sal_work_destroy": [_vp],
            "sal_work_device_bytes": [_vp, C.POINTER(_i64)],
            "sal_plan_prepare": [_vp, C.POINTER(EvalOpts), C.c_int, C.POINTER(_i64)],
            "sal_plan_stats": [_vp, C.POINTER(EvalOpts), _lp],
            "sal_plan_release": [_vp],
            "sal_plan_close_export": [_vp, C.POINTER(EvalOpts), _ip, _ip, np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")],
            "sal_eval_work": [_vp, C.POINTER(EvalOpts), _vp, _vp, _vp, C.c_int, _vp, _vp, C.c_int, _vp],
            "sal_server_opts_default": [C.POINTER(ServerOpts)],
            "sal_server_create": [C.POINTER(ServerOpts), C.POINTER(_vp)],
            "sal_server_add_geometry": [_vp, _vp, _dp, C.POINTER(EvalOpts), C.POINTER(C.c_int)],
            "sal_server_submit": [_vp, C.c_int, C.c_int, _vp, _vp, _vp, C.POINTER(_i64)],
            "sal_server_wait": [_vp, _i64, C.c_double, C.POINTER(C.c_int)],
            "sal_server_request_info": [_vp, _i64, C.POINTER(C.c_double), C.POINTER(_i64), C.POINTER(C.c_int)],
            "sal_server_stats": [_vp, _lp],
            "sal_server_destroy": [_vp],
            "sal_mom6_proxy_source": [_vp, _dp, _dp, C.c_int, _dp],
            "sal_mom6_sums": [_vp, _dp, _dp, C.c_int, C.c_int, _dp, _dp, C.c_int],
            "sal_mom6_pc_compute": [_vp, _dp, C.c_int, C.c_int, _dp, _dp, C.c_int],
            "sal_mom6_pp_compute": [_vp, _dp, _dp, _dp, _dp, C.c_int],
            "sal_work_counts": [_vp, C.c_int, _lp],
            "sal_prof_enable": [C.c_int],
            "sal_prof_reset": [],
            "sal_prof_count": [C.POINTER(C.c_int)],
            "sal_prof_get": [C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_double), C.POINTER(_i64)],
            "sal_direct_sum_dev": [_vp, _vp, _vp, _vp, _vp, _i64, C.c_int, _vp, _i64, C.c_int, _vp, _vp, _vp],
            "sal_direct_sum_dev_noalloc": [_vp, _vp, _vp, _vp, _vp, _i64, C.c_int, _vp, _i64, C.c_int, _vp, _vp, _vp],
            "sal_gradient_xyz_dev": [_vp, _vp, _vp, _i64, C.c_int, _vp, _vp, _vp, _vp],
            "sal_gen_fibonacci": [_i64, _dp, _dp, _dp],
            "sal_gen_mpas": [_i64, C.c_uint64, C.c_double, _dp, _dp, _dp, _dp],
            "sal_gen_mom6": [C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint64, C.c_double, C.c_double,
                             C.POINTER(_i64), _vp, _vp, _vp, _vp],
            "sal_gen_ssh": [_i64, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_uint64, C.c_int, _dp],
        }
        for name, args in sig.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise SalError(rc, lib().sal_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(t):
    """device pointer of a contiguous torch tensor"""
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return C.c_void_p(t.data_ptr())


def _check_tensor(t, name, shape, dtype="float64", cuda=True):
    """Raises ValueError unless t is a contiguous torch tensor of the dtype, on the device (CUDA) or the
    host, with the shape: the kernels read and write exactly that many elements of that type."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor, got {type(t).__name__}")
    if t.dtype != getattr(torch, dtype):
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if t.is_cuda != cuda:
        raise ValueError(f"{name} must be a {'CUDA' if cuda else 'CPU'} tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")


def _check_fields(n, area, ssh, gx, gy, cuda=True):
    """area [n]; ssh, gx, gy [R, n] or [n] (all alike); returns R."""
    if getattr(ssh, "dim", None) is None or ssh.dim() not in (1, 2):
        raise ValueError("ssh must be a tensor of shape [R, n] or [n]")
    shape = tuple(ssh.shape)
    if shape[-1] != n or (len(shape) == 2 and shape[0] < 1):
        raise ValueError(f"ssh must have shape [R, {n}] or [{n}], got {shape}")
    _check_tensor(ssh, "ssh", shape, cuda=cuda)
    _check_tensor(gx, "gx", shape, cuda=cuda)
    _check_tensor(gy, "gy", shape, cuda=cuda)
    _check_tensor(area, "area", (n,), cuda=cuda)
    return 1 if len(shape) == 1 else shape[0]


def _check_host(a, name, n):
    """A host array the library writes into in place: C-contiguous float64 numpy array of n elements."""
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous or a.size != n:
        raise ValueError(f"{name} must be a C-contiguous float64 numpy array of {n} elements")


def _stream(stream=None):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def set_tie_ulps(ulps):
    check(lib().sal_set_tie_ulps(float(ulps)))


def default_opts(**kw):
    o = EvalOpts()
    check(lib().sal_eval_opts_default(C.byref(o)))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Geometry:
    """HBM-resident grid: tree, interaction lists and evaluation plans (tree_module.f90)."""

    def __init__(self, x=None, y=None, z=None, cluster_thresh=65, device_xyz=None, stream=None):
        L = lib()
        h = _vp()
        if device_xyz is not None:
            dx, dy, dz = device_xyz
            n = dx.numel() if hasattr(dx, "numel") else -1
            for name, t in zip(("x", "y", "z"), (dx, dy, dz)):
                _check_tensor(t, name, (n,))
            check(L.sal_geom_create_dev(_ptr(dx), _ptr(dy), _ptr(dz), dx.numel(), cluster_thresh,
                                        _stream(stream), C.byref(h)))
            self.n = dx.numel()
        else:
            x, y, z = _f64(x), _f64(y), _f64(z)
            if x.ndim != 1 or y.shape != x.shape or z.shape != x.shape:
                raise ValueError("x, y, z must be 1-d arrays of one length")
            check(L.sal_geom_create(x, y, z, len(x), cluster_thresh, C.byref(h)))
            self.n = len(x)
        self.h = h
        self.thresh = cluster_thresh

    def save(self, path):
        """Writes the tree, the DTT list and the STT lists to `path` (the on-disk geometry cache)."""
        check(lib().sal_geom_save(self.h, os.fsencode(path)))

    @classmethod
    def load(cls, path):
        """A geometry read back from a file written by save(): same tree, lists and sums, no rebuild."""
        h = _vp()
        check(lib().sal_geom_load(os.fsencode(path), C.byref(h)))
        g = cls.__new__(cls)
        g.h = h
        n, thresh = _i64(), C.c_int()
        check(lib().sal_geom_size(h, C.byref(n), C.byref(thresh)))
        g.n, g.thresh = n.value, thresh.value
        g.n_interactions = int(sum(g.dtt_counts()))
        nt, npc, npp = _i64(), _i64(), _i64()
        check(lib().sal_stt_counts(h, C.byref(nt), C.byref(npc), C.byref(npp)))
        if nt.value:
            g.n_stt, g.n_pc, g.n_pp = nt.value, npc.value, npp.value
        return g

    def close(self):
        if getattr(self, "h", None):
            lib().sal_geom_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def counts(self):
        P, ml, tot = C.c_int(), C.c_int(), _i64()
        check(lib().sal_geom_counts(self.h, C.byref(P), C.byref(ml), C.byref(tot)))
        return P.value, ml.value, tot.value

    def ties(self):
        a, b = _i64(), _i64()
        check(lib().sal_geom_ties(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def device_bytes(self):
        b = _i64()
        check(lib().sal_geom_device_bytes(self.h, C.byref(b)))
        return b.value

    def export(self):
        P, _, tot = self.counts()
        level, is_leaf, parent, face, count = (np.zeros(P, np.int32) for _ in range(5))
        child = np.zeros(4 * P, np.int32)
        bounds = np.zeros(6 * P)
        radius = np.zeros(P)
        offsets = np.zeros(P + 1, np.int64)
        pts = np.zeros(max(tot, 1), np.int32)
        check(lib().sal_geom_export(self.h, level, is_leaf, parent, child, face, bounds, radius, count, offsets,
                                    pts))
        return dict(level=level, is_leaf=is_leaf, parent=parent, child=child.reshape(P, 4), face=face,
                    bounds=bounds.reshape(P, 6), radius=radius, count=count, offsets=offsets, points=pts[:tot])

    def assign_points(self, x, y, z, nlev):
        x, y, z = _f64(x), _f64(y), _f64(z)
        pp = np.zeros(len(x) * nlev, np.int32)
        check(lib().sal_assign_points(self.h, x, y, z, len(x), nlev, pp))
        return pp.reshape(len(x), nlev)

    def dtt(self, theta, cluster_thresh=0):
        """DUAL_TREE_TRAVERSAL; cluster_thresh types the entries (0: the tree's threshold)."""
        n = _i64()
        check(lib().sal_dtt(self.h, float(theta), int(cluster_thresh), C.byref(n)))
        self.n_interactions = n.value
        return n.value

    def dtt_counts(self):
        c = np.zeros(4, np.int64)
        check(lib().sal_dtt_counts(self.h, c))
        return c

    def dtt_export(self):
        n = self.n_interactions
        it, is_, ty = (np.zeros(max(n, 1), np.int32) for _ in range(3))
        check(lib().sal_dtt_export(self.h, it, is_, ty))
        return it[:n], is_[:n], ty[:n]

    def dtt_import(self, it, is_, ty):
        it, is_, ty = (np.ascontiguousarray(a, dtype=np.int32) for a in (it, is_, ty))
        check(lib().sal_dtt_import(self.h, it, is_, ty, len(it)))
        self.n_interactions = len(it)

    def stt(self, theta, x, y, z):
        x, y, z = _f64(x), _f64(y), _f64(z)
        a, b = _i64(), _i64()
        check(lib().sal_stt(self.h, float(theta), x, y, z, len(x), C.byref(a), C.byref(b)))
        self.n_pc, self.n_pp = a.value, b.value
        self.n_stt = len(x)
        return a.value, b.value

    def stt_export(self):
        a = [np.zeros(max(self.n_pc, 1), np.int32) for _ in range(2)]
        b = [np.zeros(max(self.n_pp, 1), np.int32) for _ in range(2)]
        check(lib().sal_stt_export(self.h, a[0], a[1], b[0], b[1]))
        return a[0][:self.n_pc], a[1][:self.n_pc], b[0][:self.n_pp], b[1][:self.n_pp]

    def fast_sum(self, area, ssh, degree, gx=None, gy=None):
        """fast_sum_module.f90:291 on host arrays; accumulates into gx, gy when given."""
        area, ssh = _f64(area), _f64(ssh)
        if area.size != self.n or ssh.size != self.n:
            raise ValueError(f"area and ssh must have {self.n} elements")
        acc = gx is not None
        if not acc:
            gx, gy = np.zeros(self.n), np.zeros(self.n)
        _check_host(gx, "gx", self.n)
        _check_host(gy, "gy", self.n)
        check(lib().sal_fast_sum(self.h, area, ssh, int(degree), gx, gy, int(acc)))
        return gx, gy

    def eval(self, area, ssh, gx, gy, opts=None, accumulate=False, stream=None):
        """Device tensors: area [n], ssh [R, n] or [n], gx/gy like ssh (float64)."""
        if opts is None:
            opts = default_opts()
        R = _check_fields(self.n, area, ssh, gx, gy)
        check(lib().sal_eval_dev(self.h, C.byref(opts), _ptr(area), _ptr(ssh), R, _ptr(gx), _ptr(gy),
                                 int(accumulate), _stream(stream)))

    def _grid_fields(self, area, ssh):
        area, ssh = _f64(area), _f64(ssh)
        if area.size != self.n or ssh.size != self.n:
            raise ValueError(f"area and ssh must have {self.n} elements")
        return area, ssh

    def mom6_proxy_source(self, area, ssh, degree):
        area, ssh = self._grid_fields(area, ssh)
        P = self.counts()[0]
        w = np.zeros((degree + 1) ** 2 * P)
        check(lib().sal_mom6_proxy_source(self.h, area, ssh, int(degree), w))
        return w

    def mom6_sums(self, area, ssh, degree, transposed=True):
        """proxy_source_compute + pc_compute + pp_compute for the STT targets (host arrays)."""
        area, ssh = self._grid_fields(area, ssh)
        gx, gy = np.zeros(self.n_stt), np.zeros(self.n_stt)
        check(lib().sal_mom6_sums(self.h, area, ssh, int(degree), int(transposed), gx, gy, 0))
        return gx, gy

    def mom6_pc_compute(self, w, degree, gx, gy, transposed=True):
        w = _f64(w)
        if w.size != (degree + 1) ** 2 * self.counts()[0]:
            raise ValueError("w must hold (degree + 1)^2 weights per panel")
        _check_host(gx, "gx", self.n_stt)
        _check_host(gy, "gy", self.n_stt)
        check(lib().sal_mom6_pc_compute(self.h, w, int(degree), int(transposed), gx, gy, 1))

    def mom6_pp_compute(self, area, ssh, gx, gy):
        area, ssh = self._grid_fields(area, ssh)
        _check_host(gx, "gx", self.n_stt)
        _check_host(gy, "gy", self.n_stt)
        check(lib().sal_mom6_pp_compute(self.h, area, ssh, gx, gy, 1))

    def work_counts(self, degree):
        K = np.zeros(4, np.int64)
        check(lib().sal_work_counts(self.h, int(degree), K))
        return K

    def prepare(self, opts, R=1):
        """Builds the mixed plan of `opts` now; returns the workspace bytes R fields need."""
        b = _i64()
        check(lib().sal_plan_prepare(self.h, C.byref(opts), int(R), C.byref(b)))
        return b.value

    def plan_stats(self, opts):
        st = np.zeros(13, np.int64)
        check(lib().sal_plan_stats(self.h, C.byref(opts), st))
        keys = ("m", "cluster_panels", "p2m_chunks", "proxy_target_panels", "leaves", "near_entries",
                "far_entries", "local_expansion_panels", "near_work_items", "l2p_leaves", "stored_near_bytes",
                "stored_far_bytes", "close_pairs")
        return dict(zip(keys, (int(v) for v in st)))

    def close_pairs(self, opts):
        """The mixed plan's close-pair list (summed in FP64): (target, source) point ids in the caller's
        order, and each panel's FP32 chord-squared cutoff."""
        nc = self.plan_stats(opts)["close_pairs"]
        t, s = np.zeros(max(nc, 1), np.int32), np.zeros(max(nc, 1), np.int32)
        cut2 = np.zeros(self.counts()[0], np.float32)
        check(lib().sal_plan_close_export(self.h, C.byref(opts), t, s, cut2))
        return t[:nc], s[:nc], cut2

    def release_plans(self):
        """Frees the cached evaluation plans (work lists, stored operators, work buffers)."""
        check(lib().sal_plan_release(self.h))

    def eval_work(self, work, area, ssh, gx, gy, opts, accumulate=False, stream=None):
        """Mixed-precision evaluation with the caller's Workspace (one per concurrent caller)."""
        R = _check_fields(self.n, area, ssh, gx, gy)
        check(lib().sal_eval_work(self.h, C.byref(opts), work.h, _ptr(area), _ptr(ssh), R, _ptr(gx), _ptr(gy),
                                  int(accumulate), _stream(stream)))


class Workspace:
    """Device buffers of one mixed-precision evaluation at a time (include/salfort.h, workspaces)."""

    def __init__(self):
        self.h = _vp()
        check(lib().sal_work_create(C.byref(self.h)))

    def device_bytes(self):
        b = _i64()
        check(lib().sal_work_device_bytes(self.h, C.byref(b)))
        return b.value

    def close(self):
        if getattr(self, "h", None):
            lib().sal_work_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop over HBM-resident geometries (sal_server_* in the C ABI).

    Requests are host arrays: ssh [R, n] float64 in, gx / gy [R, n] float64 out, kept alive by the
    caller until wait() reports them done (torch pinned tensors copy fastest)."""

    def __init__(self, slots=3, max_fields=32, batching=True, graph=False, hbm_budget=0, hbm_reserve=None):
        o = ServerOpts()
        check(lib().sal_server_opts_default(C.byref(o)))
        o.slots, o.max_fields, o.batching, o.graph = int(slots), int(max_fields), int(batching), int(graph)
        o.hbm_budget = int(hbm_budget)
        if hbm_reserve is not None:
            o.hbm_reserve = int(hbm_reserve)
        self.h = _vp()
        check(lib().sal_server_create(C.byref(o), C.byref(self.h)))
        self.geoms = []

    def add_geometry(self, geom, area, opts):
        area = _f64(area)
        if area.size != geom.n:
            raise ValueError(f"area must have {geom.n} elements")
        gid = C.c_int()
        check(lib().sal_server_add_geometry(self.h, geom.h, area, C.byref(opts), C.byref(gid)))
        self.geoms.append(geom)  # keeps the handle alive as long as the server
        return gid.value

    def submit(self, gid, ssh, gx, gy):
        """ssh, gx, gy: host tensors or arrays [R, n] float64, C-contiguous."""
        if not 0 <= gid < len(self.geoms):
            raise ValueError(f"no geometry {gid}")
        n = self.geoms[gid].n
        shape = tuple(ssh.shape)
        if len(shape) not in (1, 2) or shape[-1] != n:
            raise ValueError(f"ssh must have shape [R, {n}] or [{n}], got {shape}")
        for name, a in (("gx", gx), ("gy", gy)):
            if tuple(a.shape) != shape:
                raise ValueError(f"{name} must have the shape of ssh {shape}")
        R = 1 if len(shape) == 1 else shape[0]
        req = _i64()
        check(lib().sal_server_submit(self.h, int(gid), R, _host_ptr(ssh), _host_ptr(gx), _host_ptr(gy),
                                      C.byref(req)))
        return req.value

    def wait(self, req, timeout=-1.0):
        done = C.c_int()
        check(lib().sal_server_wait(self.h, int(req), float(timeout), C.byref(done)))
        return bool(done.value)

    def info(self, req):
        """(t_submit, t_admit, t_done) in seconds since the server started, batch id, batch fields."""
        t = (C.c_double * 3)()
        b, f = _i64(), C.c_int()
        check(lib().sal_server_request_info(self.h, int(req), t, C.byref(b), C.byref(f)))
        return tuple(t), b.value, f.value

    def stats(self):
        st = np.zeros(8, np.int64)
        check(lib().sal_server_stats(self.h, st))
        keys = ("submitted", "completed", "rejected", "batches", "merged_requests", "peak_busy_slots",
                "peak_held_bytes", "budget_bytes")
        return dict(zip(keys, (int(v) for v in st)))

    def close(self):
        if getattr(self, "h", None):
            check(lib().sal_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _host_ptr(a):
    """host address of a C-contiguous float64 numpy array or CPU torch tensor"""
    if isinstance(a, np.ndarray):
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise ValueError("host arrays must be C-contiguous float64")
        return C.c_void_p(a.ctypes.data)
    import torch
    if not isinstance(a, torch.Tensor) or a.is_cuda or not a.is_contiguous() or a.dtype != torch.float64:
        raise ValueError("host tensors must be contiguous float64 CPU tensors")
    return C.c_void_p(a.data_ptr())


def direct_sum(x, y, z, area, ssh, targets, mode, stream=None):
    """FP64 direct sum (device tensors); ssh [R, n] or [n]; returns (gx, gy) [R, nt] or [nt]."""
    import torch
    n = x.numel() if hasattr(x, "numel") else -1
    for name, t in zip(("x", "y", "z", "area"), (x, y, z, area)):
        _check_tensor(t, name, (n,))
    if getattr(ssh, "dim", None) is None or ssh.dim() not in (1, 2) or ssh.shape[-1] != n:
        raise ValueError(f"ssh must have shape [R, {n}] or [{n}]")
    _check_tensor(ssh, "ssh", tuple(ssh.shape))
    _check_tensor(targets, "targets", (targets.numel(),) if hasattr(targets, "numel") else (), dtype="int32")
    if mode not in (0, 1, 2):
        raise ValueError("mode must be 0, 1 or 2")
    nt = targets.numel()
    if nt and (int(targets.min()) < 0 or int(targets.max()) >= n):
        raise ValueError(f"targets must be point indices in [0, {n})")
    R = 1 if ssh.dim() == 1 else ssh.shape[0]
    gx = torch.empty((R, nt), dtype=torch.float64, device=x.device)
    gy = torch.empty_like(gx)
    check(lib().sal_direct_sum_dev(_ptr(x), _ptr(y), _ptr(z), _ptr(area), _ptr(ssh), x.numel(), R,
                                   _ptr(targets), nt, int(mode), _ptr(gx), _ptr(gy), _stream(stream)))
    if ssh.dim() == 1:
        return gx[0], gy[0]
    return gx, gy


def gradient_xyz(x, y, z, gx, gy, stream=None):
    """The gradient as a tangential Cartesian 3-vector per point (device tensors): gx, gy are the
    east/north components [R, n] or [n] in the points' order; returns [R, n, 3] or [n, 3]."""
    import torch
    n = x.numel() if hasattr(x, "numel") else -1
    for name, t in zip(("x", "y", "z"), (x, y, z)):
        _check_tensor(t, name, (n,))
    if getattr(gx, "dim", None) is None or gx.dim() not in (1, 2) or gx.shape[-1] != n:
        raise ValueError(f"gx must have shape [R, {n}] or [{n}]")
    R = 1 if gx.dim() == 1 else gx.shape[0]
    gx, gy = gx.contiguous(), gy.contiguous()
    _check_tensor(gx, "gx", tuple(gx.shape))
    _check_tensor(gy, "gy", tuple(gx.shape))
    g3 = torch.empty((R, n, 3), dtype=torch.float64, device=x.device)
    check(lib().sal_gradient_xyz_dev(_ptr(x), _ptr(y), _ptr(z), n, R, _ptr(gx), _ptr(gy), _ptr(g3),
                                     _stream(stream)))
    return g3[0] if gx.dim() == 1 else g3


# ---------------------------------------------------------------------------------------- inputs

def gen_fibonacci(n):
    x, y, z = (np.zeros(n) for _ in range(3))
    check(lib().sal_gen_fibonacci(n, x, y, z))
    return x, y, z


def gen_mpas(n, seed=0, total_area=3.61e14):
    x, y, z, a = (np.zeros(n) for _ in range(4))
    check(lib().sal_gen_mpas(n, seed, total_area, x, y, z, a))
    return x, y, z, a


def gen_mom6(nlon=1440, nlat=1080, lat0=-80.0, lat1=90.0, seed=0, land_frac=0.29, radius=6.371e6):
    n = _i64()
    check(lib().sal_gen_mom6(nlon, nlat, lat0, lat1, seed, land_frac, radius, C.byref(n), None, None, None, None))
    x, y, z, a = (np.zeros(n.value) for _ in range(4))
    check(lib().sal_gen_mom6(nlon, nlat, lat0, lat1, seed, land_frac, radius, C.byref(n),
                             x.ctypes.data_as(_vp), y.ctypes.data_as(_vp), z.ctypes.data_as(_vp),
                             a.ctypes.data_as(_vp)))
    return x, y, z, a


def gen_ssh(x, y, z, lmax=16, amp=0.1, noise=0.0, seed=0, R=1):
    x, y, z = _f64(x), _f64(y), _f64(z)
    out = np.zeros(R * len(x))
    check(lib().sal_gen_ssh(len(x), x, y, z, lmax, amp, noise, seed, R, out))
    return out.reshape(R, len(x)) if R > 1 else out


# the workloads of the benchmark and the parity tests, by name
EARTH_R = 6.371e6


def workload(name, n=None, seed=0, R=1):
    """Synthetic grid + fields: returns dict(x, y, z, area, ssh[R, n] or [n], thresh, degree, theta)."""
    if name == "fibonacci":   # N=40,962 Fibonacci lattice, l<=16 harmonics, d=4, leaf <= 64
        n = n or 40962
        x, y, z = gen_fibonacci(n)
        area = np.full(n, 4 * np.pi * EARTH_R ** 2 / n)
        ssh = gen_ssh(x, y, z, lmax=16, amp=0.1, noise=0.0, seed=seed, R=R)
        return dict(x=x, y=y, z=z, area=area, ssh=ssh, thresh=65, degree=4, theta=0.7)
    if name == "mpas":        # N=2,313,486 variable resolution, l<=64 + 0.02 m noise, d=8, leaf <= 128
        n = n or 2313486
        x, y, z, area = gen_mpas(n, seed=seed, total_area=3.61e14)
        ssh = gen_ssh(x, y, z, lmax=64, amp=0.1, noise=0.02, seed=seed + 1, R=R)
        return dict(x=x, y=y, z=z, area=area, ssh=ssh, thresh=129, degree=8, theta=0.7)
    if name == "mom6":        # 1440x1080 lat-lon ocean cells (~1.1M), fields as mpas
        x, y, z, area = gen_mom6(seed=seed)
        ssh = gen_ssh(x, y, z, lmax=64, amp=0.1, noise=0.02, seed=seed + 1, R=R)
        return dict(x=x, y=y, z=z, area=area, ssh=ssh, thresh=129, degree=8, theta=0.7)
    raise ValueError(name)
