"""ctypes bindings for the two CPU checkers (test infrastructure only).

``Oracle`` wraps ``liboracle.so`` (the C restatement, symbol prefix ``gzo_``) and ``Reference``
wraps ``_ref/libgzref.so`` (the reference's ``shmem/tree.cc``, prefix ``gzref_``). Both expose the
same methods so a test can run one input through either and compare.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libgzref.so")
REF_SO_V4 = os.path.join(HERE, "_ref", "libgzref_v4.so")


def _cpu_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = set(line.split(":", 1)[1].split())
                    return {"avx512f", "avx512bw", "avx512dq", "avx512vl"} <= fl
    except OSError:
        pass
    return False


def reference_path() -> str:
    """The reference build for this CPU: the AVX-512 one where the CPU has it, else AVX2."""
    if os.path.exists(REF_SO_V4) and _cpu_has_avx512():
        return REF_SO_V4
    return REF_SO

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_llp = C.POINTER(C.c_longlong)


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctype)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


@dataclass
class HostParticles:
    """Host SoA particle set (the reference's ``Particles``, tree.h:36-50)."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    m: np.ndarray
    soft: np.ndarray | None = None

    def __post_init__(self):
        self.x, self.y, self.z, self.m = (_f64(v) for v in (self.x, self.y, self.z, self.m))
        if self.soft is not None:
            self.soft = _f64(self.soft)

    @property
    def n(self) -> int:
        return int(self.m.shape[0])


class _GzoParticles(C.Structure):
    _fields_ = [("n", C.c_size_t), ("x", _dp), ("y", _dp), ("z", _dp), ("m", _dp),
                ("soft", _dp)]


class _GzoTreeView(C.Structure):
    _fields_ = [("nnodes", C.c_size_t), ("n_live", C.c_size_t), ("root", C.c_int),
                ("first", _ip), ("next", _ip), ("parent", _ip), ("plo", _ip), ("phi", _ip),
                ("mass", _dp), ("cx", _dp), ("cy", _dp), ("cz", _dp),
                ("size", _dp), ("gcx", _dp), ("gcy", _dp), ("gcz", _dp),
                ("wn_soft", _fp), ("vmax", _fp), ("xpad", _fp), ("tau", C.c_double),
                ("orderbuf", _up), ("leaf_of", _ip)]


@dataclass
class TreeArrays:
    """Node and particle arrays of one built tree, as numpy arrays."""

    nnodes: int
    root: int
    n_live: int
    first: np.ndarray
    next: np.ndarray
    plo: np.ndarray
    phi: np.ndarray
    parent: np.ndarray
    size: np.ndarray
    mass: np.ndarray
    cx: np.ndarray
    cy: np.ndarray
    cz: np.ndarray
    soft: np.ndarray
    delta: np.ndarray
    gcx: np.ndarray
    gcy: np.ndarray
    gcz: np.ndarray
    wn_s: np.ndarray
    wn_soft: np.ndarray
    wn_len: np.ndarray
    orderbuf: np.ndarray
    rank: np.ndarray
    leaf_of: np.ndarray


@dataclass
class ReuseState:
    """Per-node state of a tree reused between rebuilds (tree.h:266-317) and its walk view."""
    vmax: np.ndarray
    vcom: np.ndarray
    dp: np.ndarray
    dq: np.ndarray
    wcom: np.ndarray
    wmass: np.ndarray
    wn_s: np.ndarray
    wn_soft: np.ndarray
    wn_len: np.ndarray


def _reuse_arrays(nn):
    return ReuseState(np.zeros(nn, np.float32), np.zeros((nn, 3)), np.zeros((nn, 3)),
                      np.zeros((nn, 3)), np.zeros((nn, 3)), np.zeros(nn), np.zeros(nn),
                      np.zeros(nn, np.float32), np.zeros(nn, np.float32))


def _opt3(v):
    """Three float64 arrays the callee writes in place (or None)."""
    if v is None:
        return (None, None, None)
    for a in v:
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "v_last must be float64 arrays"
    return tuple(_ptr(a, _dp) for a in v)


class _Lib:
    prefix = ""

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
        self.lib = C.CDLL(path)
        self.path = path
        L, p = self.lib, self.prefix
        getattr(L, p + "morton").restype = C.c_uint64
        getattr(L, p + "morton").argtypes = [C.c_uint32] * 3
        getattr(L, p + "spline_force_over_r").restype = C.c_double
        getattr(L, p + "spline_force_over_r").argtypes = [C.c_double, C.c_double]
        getattr(L, p + "spline_potential").restype = C.c_double
        getattr(L, p + "spline_potential").argtypes = [C.c_double, C.c_double]
        getattr(L, p + "open_node").restype = C.c_int
        getattr(L, p + "open_node").argtypes = [C.c_double] * 11
        getattr(L, p + "tree_free").argtypes = [C.c_void_p]
        getattr(L, p + "tree_sizes").argtypes = [C.c_void_p, _llp]
        getattr(L, p + "tree_nodes").argtypes = [C.c_void_p] + [_ip] * 5 + [_dp] * 11 + [_fp, _fp]
        getattr(L, p + "tree_particles").argtypes = [C.c_void_p, _up, _up, _ip]

    # -- scalar kernels ------------------------------------------------------------------------
    def morton(self, a: int, b: int, c: int) -> int:
        return int(getattr(self.lib, self.prefix + "morton")(a, b, c))

    def spline_force_over_r(self, r: float, h: float) -> float:
        return getattr(self.lib, self.prefix + "spline_force_over_r")(r, h)

    def spline_potential(self, r: float, h: float) -> float:
        return getattr(self.lib, self.prefix + "spline_potential")(r, h)

    def open_node(self, *args) -> bool:
        return bool(getattr(self.lib, self.prefix + "open_node")(*[float(a) for a in args]))

    # -- trees ---------------------------------------------------------------------------------
    def _tree_arrays(self, handle) -> TreeArrays:
        p = self.prefix
        sz = np.zeros(3, dtype=np.int64)
        getattr(self.lib, p + "tree_sizes")(handle, _ptr(sz, _llp))
        nn, root, nl = (int(v) for v in sz)
        ints = {k: np.zeros(nn, dtype=np.int32) for k in ("first", "next", "plo", "phi", "parent")}
        dbls = {k: np.zeros(nn, dtype=np.float64) for k in
                ("size", "mass", "cx", "cy", "cz", "soft", "delta", "gcx", "gcy", "gcz", "wn_s")}
        flts = {k: np.zeros(nn, dtype=np.float32) for k in ("wn_soft", "wn_len")}
        getattr(self.lib, p + "tree_nodes")(
            handle, *[_ptr(ints[k], _ip) for k in ("first", "next", "plo", "phi", "parent")],
            *[_ptr(dbls[k], _dp) for k in ("size", "mass", "cx", "cy", "cz", "soft", "delta",
                                           "gcx", "gcy", "gcz", "wn_s")],
            _ptr(flts["wn_soft"], _fp), _ptr(flts["wn_len"], _fp))
        return TreeArrays(nnodes=nn, root=root, n_live=nl, orderbuf=None, rank=None,
                          leaf_of=None, **ints, **dbls, **flts)


class _EventsMixin:
    """Tree::mass_event, shift_node, raise_pad and raise_vmax (tree.cc:239-258, 276-285,
    323-330) for the leaf of each listed particle, in order, with per-event values; the same
    entry points in the oracle and the reference adapter (prefix gzo_ / gzref_)."""

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def mass_events(self, t, idx, dm, x, v, tau, box=0.0):
        ix = np.ascontiguousarray(idx, dtype=np.uint32)
        arrs = [_f64(dm)] + [_f64(a) for a in x] + [_f64(a) for a in v]
        self._fn("mass_events")(t, _ptr(ix, _up), ix.shape[0], *(_ptr(a, _dp) for a in arrs),
                                tau, box)

    def shift_nodes(self, t, idx, m, dx):
        ix = np.ascontiguousarray(idx, dtype=np.uint32)
        arrs = [_f64(m)] + [_f64(a) for a in dx]
        self._fn("shift_nodes")(t, _ptr(ix, _up), ix.shape[0], *(_ptr(a, _dp) for a in arrs))

    def raise_pad(self, t, idx, d):
        ix = np.ascontiguousarray(idx, dtype=np.uint32)
        self._fn("raise_pad")(t, _ptr(ix, _up), ix.shape[0], _ptr(_f64(d), _dp))

    def raise_vmax(self, t, idx, speed):
        ix = np.ascontiguousarray(idx, dtype=np.uint32)
        sp = np.ascontiguousarray(speed, dtype=np.float32)
        self._fn("raise_vmax")(t, _ptr(ix, _up), ix.shape[0], _ptr(sp, _fp))

    def pads(self, t):
        """(xpad float32, dmass float64) per node."""
        nn = self._tree_arrays(t).nnodes
        xpad, dmass = np.zeros(nn, np.float32), np.zeros(nn)
        self._fn("tree_pads")(t, _ptr(xpad, _fp), _ptr(dmass, _dp))
        return xpad, dmass


class Oracle(_EventsMixin, _Lib):
    """The C restatement, ``liboracle.so``."""

    prefix = "gzo_"

    def __init__(self, path: str = ORACLE_SO):
        super().__init__(path)
        L = self.lib
        L.gzo_keys.argtypes = [C.POINTER(_GzoParticles), C.c_longlong, _u64p, _dp]
        L.gzo_build.restype = C.c_void_p
        L.gzo_build.argtypes = [C.POINTER(_GzoParticles), C.c_longlong]
        L.gzo_build_shift.restype = C.c_void_p
        L.gzo_build_shift.argtypes = [C.POINTER(_GzoParticles), C.c_longlong, C.c_double,
                                      C.c_double, C.c_double]
        L.gzo_tree_skeys.argtypes = [C.c_void_p, _u64p]
        L.gzo_accel_grouped.argtypes = [C.c_void_p, C.POINTER(_GzoParticles), _up, C.c_size_t,
                                        C.c_double, C.c_double, C.c_int, _dp, _dp, _dp, _dp, _llp]
        L.gzo_accel.argtypes = [C.c_void_p, C.POINTER(_GzoParticles), _up, C.c_size_t,
                                C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.gzo_accel_brute.argtypes = [C.POINTER(_GzoParticles), _up, C.c_size_t, C.c_double,
                                      _dp, _dp, _dp]
        L.gzo_audit_tree.restype = C.c_longlong
        L.gzo_audit_tree.argtypes = [C.POINTER(_GzoTreeView), C.POINTER(_GzoParticles), _dp,
                                     _dp, _dp, C.c_int, C.c_char_p, C.c_size_t]
        L.gzo_direct_sum_f32.argtypes = [C.POINTER(_GzoParticles), _up, C.c_size_t,
                                         C.c_double, C.c_double, _dp, _dp, _dp]
        L.gzo_direct_sum_box.argtypes = [C.POINTER(_GzoParticles), _up, C.c_size_t,
                                         C.c_double, C.c_double, _dp, _dp, _dp]
        L.gzo_accel_grouped_tj.argtypes = [C.c_void_p, C.POINTER(_GzoParticles), _up,
                                           C.c_size_t, C.c_double, C.c_double, C.c_int, _dp,
                                           _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _llp]
        L.gzo_ngb_search.restype = C.c_size_t
        L.gzo_ngb_search.argtypes = [C.c_void_p, C.POINTER(_GzoParticles), _dp, _dp, _dp, _dp,
                                     C.c_size_t, C.c_double, C.POINTER(C.c_size_t), _up,
                                     C.c_size_t]
        L.gzo_init_hmax.argtypes = [C.c_void_p, _dp, C.c_size_t, C.c_size_t,
                                    C.POINTER(C.c_float)]
        L.gzo_raise_hmax.argtypes = [C.c_void_p, _up, C.c_size_t, _dp]
        L.gzo_tree_hmax.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
        L.gzo_ngb_search_reverse.restype = C.c_size_t
        L.gzo_ngb_search_reverse.argtypes = [C.c_void_p, C.POINTER(_GzoParticles), _dp,
                                             C.c_size_t, _dp, _dp, _dp, _dp, C.c_size_t,
                                             C.c_double, C.POINTER(C.c_size_t), _up, C.c_size_t]
        L.gzo_potential.argtypes = [C.c_void_p, C.POINTER(_GzoParticles), _up, C.c_size_t,
                                    C.c_double, C.c_double, _dp, _dp]
        L.gzo_build_vel.restype = C.c_void_p
        L.gzo_build_vel.argtypes = [C.POINTER(_GzoParticles), C.c_longlong, _dp, _dp, _dp]
        L.gzo_kick_nodes.argtypes = [C.c_void_p, _up, C.c_size_t, _dp, _dp, _dp, _dp, _dp, _dp,
                                     _dp, C.c_double]
        L.gzo_refresh_walk_nodes.argtypes = [C.c_void_p, C.c_double]
        L.gzo_raise_soft.argtypes = [C.c_void_p, _up, C.c_size_t, _dp]
        L.gzo_tree_reuse.argtypes = [C.c_void_p, _fp, _dp, _dp, _dp, _dp, _dp]
        L.gzo_mass_events.argtypes = [C.c_void_p, _up, C.c_size_t] + [_dp] * 7 + [C.c_double,
                                                                               C.c_double]
        L.gzo_shift_nodes.argtypes = [C.c_void_p, _up, C.c_size_t, _dp, _dp, _dp, _dp]
        L.gzo_raise_pad.argtypes = [C.c_void_p, _up, C.c_size_t, _dp]
        L.gzo_raise_vmax.argtypes = [C.c_void_p, _up, C.c_size_t, _fp]
        L.gzo_tree_pads.argtypes = [C.c_void_p, _fp, _dp]
        L.gzo_density.argtypes = [C.c_void_p, C.POINTER(_GzoParticles), C.c_void_p, _up,
                                  C.c_size_t, C.c_double, C.c_double, _dp, C.c_int, _dp,
                                  C.c_double, _dp, _dp, _ip, _ip]
        L.gzo_kernel_w.restype = C.c_double
        L.gzo_kernel_w.argtypes = [C.c_double, C.c_double, C.c_int]
        L.gzo_kernel_dwdh.restype = C.c_double
        L.gzo_kernel_dwdh.argtypes = [C.c_double, C.c_double, C.c_int]

    @staticmethod
    def _cp(P: HostParticles):
        return _GzoParticles(P.n, _ptr(P.x, _dp), _ptr(P.y, _dp), _ptr(P.z, _dp),
                             _ptr(P.m, _dp), _ptr(P.soft, _dp))

    def keys(self, P: HostParticles, seed: int = -1):
        keys = np.zeros(P.n, dtype=np.uint64)
        root = np.zeros(4, dtype=np.float64)
        cp = self._cp(P)
        self.lib.gzo_keys(C.byref(cp), seed, _ptr(keys, _u64p), _ptr(root, _dp))
        return keys, root

    def build(self, P: HostParticles, seed: int = -1, shift=None):
        """shift: SHMEM_TREE_SHIFT's fractions (tree.cc:400-407), or None."""
        cp = self._cp(P)
        if shift is not None:
            return self.lib.gzo_build_shift(C.byref(cp), seed, *(float(v) for v in shift))
        return self.lib.gzo_build(C.byref(cp), seed)

    def free(self, t):
        self.lib.gzo_tree_free(t)

    def tree(self, t, P: HostParticles) -> TreeArrays:
        ta = self._tree_arrays(t)
        ta.orderbuf = np.zeros(ta.n_live, dtype=np.uint32)
        ta.rank = np.zeros(P.n, dtype=np.uint32)
        ta.leaf_of = np.zeros(P.n, dtype=np.int32)
        self.lib.gzo_tree_particles(t, _ptr(ta.orderbuf, _up), _ptr(ta.rank, _up),
                                    _ptr(ta.leaf_of, _ip))
        return ta

    def skeys(self, t, n_live: int):
        out = np.zeros(n_live, dtype=np.uint64)
        self.lib.gzo_tree_skeys(t, _ptr(out, _u64p))
        return out

    def accel_grouped(self, t, P: HostParticles, targets, theta, G=1.0, batch=0, aold=None,
                      want_counts=False):
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        nt = tg.shape[0]
        ax, ay, az = (np.zeros(nt) for _ in range(3))
        counts = np.zeros(3, dtype=np.int64)
        ao = None if aold is None else _f64(aold)
        cp = self._cp(P)
        self.lib.gzo_accel_grouped(t, C.byref(cp), _ptr(tg, _up), nt, theta, G, batch,
                                   _ptr(ao, _dp), _ptr(ax, _dp), _ptr(ay, _dp), _ptr(az, _dp),
                                   _ptr(counts, _llp))
        if want_counts:
            return ax, ay, az, counts
        return ax, ay, az

    def accel_grouped_tj(self, t, P: HostParticles, targets, theta, G=1.0, batch=0, aold=None,
                         vel=None, jerk=False):
        """(ax, ay, az, tidal (6, nt), jerk (3, nt) or None)."""
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        nt = tg.shape[0]
        ax, ay, az = (np.zeros(nt) for _ in range(3))
        tid = np.zeros((6, nt))
        jk = np.zeros((3, nt)) if jerk else None
        ao = None if aold is None else _f64(aold)
        v = (None, None, None) if vel is None else tuple(_f64(a) for a in vel)
        cp = self._cp(P)
        self.lib.gzo_accel_grouped_tj(t, C.byref(cp), _ptr(tg, _up), nt, theta, G, batch,
                                      _ptr(ao, _dp), *(_ptr(a, _dp) for a in v), _ptr(ax, _dp),
                                      _ptr(ay, _dp), _ptr(az, _dp), _ptr(tid, _dp),
                                      _ptr(jk, _dp), None)
        return ax, ay, az, tid, jk

    def ngb_search(self, t, P: HostParticles, centres, radius, tau=0.0):
        """CSR (start, found) of the particles within radius[q] of centres[q] (n_q x 3)."""
        c = np.ascontiguousarray(centres, dtype=np.float64)
        cs = [np.ascontiguousarray(c[:, d]) for d in range(3)]
        r = _f64(radius)
        nq = r.shape[0]
        start = np.zeros(nq + 1, dtype=np.uint64)
        sp = start.ctypes.data_as(C.POINTER(C.c_size_t))
        cp = self._cp(P)
        total = self.lib.gzo_ngb_search(t, C.byref(cp), *(_ptr(a, _dp) for a in cs),
                                        _ptr(r, _dp), nq, tau, sp, None, 0)
        found = np.zeros(max(int(total), 1), dtype=np.uint32)
        self.lib.gzo_ngb_search(t, C.byref(cp), *(_ptr(a, _dp) for a in cs), _ptr(r, _dp), nq,
                                tau, sp, _ptr(found, _up), found.shape[0])
        return start.astype(np.int64), found[:int(total)]

    def init_hmax(self, t, h, n_gas):
        """Node hmax (float32 per node) from the gas kernel radii; kept for the reverse search."""
        hh = _f64(h)
        sz = np.zeros(3, dtype=np.int64)
        self.lib.gzo_tree_sizes(t, _ptr(sz, _llp))
        out = np.zeros(int(sz[0]), dtype=np.float32)
        self.lib.gzo_init_hmax(t, _ptr(hh, _dp), hh.shape[0], int(n_gas),
                               out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def raise_hmax(self, t, idx, h):
        ix = np.ascontiguousarray(idx, dtype=np.uint32)
        hh = _f64(h)
        self.lib.gzo_raise_hmax(t, _ptr(ix, _up), ix.shape[0], _ptr(hh, _dp))

    def hmax(self, t):
        """The node hmax as the oracle tree holds it (init_hmax, then raises)."""
        sz = np.zeros(3, dtype=np.int64)
        self.lib.gzo_tree_sizes(t, _ptr(sz, _llp))
        out = np.zeros(int(sz[0]), dtype=np.float32)
        self.lib.gzo_tree_hmax(t, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def ngb_search_reverse(self, t, P: HostParticles, h, n_gas, centres, hq, tau=0.0):
        """CSR (start, found) of the gas particles j with hq[q] <= r < h[j] (after init_hmax)."""
        c = np.ascontiguousarray(centres, dtype=np.float64)
        cs = [np.ascontiguousarray(c[:, d]) for d in range(3)]
        hh, r = _f64(h), _f64(hq)
        nq = r.shape[0]
        start = np.zeros(nq + 1, dtype=np.uint64)
        sp = start.ctypes.data_as(C.POINTER(C.c_size_t))
        cp = self._cp(P)
        args = (t, C.byref(cp), _ptr(hh, _dp), int(n_gas), *(_ptr(a, _dp) for a in cs),
                _ptr(r, _dp), nq, tau, sp)
        total = self.lib.gzo_ngb_search_reverse(*args, None, 0)
        found = np.zeros(max(int(total), 1), dtype=np.uint32)
        self.lib.gzo_ngb_search_reverse(*args, _ptr(found, _up), found.shape[0])
        return start.astype(np.int64), found[:int(total)]

    def potential(self, t, P: HostParticles, targets, theta, G=1.0, aold=None):
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        out = np.zeros(tg.shape[0])
        ao = None if aold is None else _f64(aold)
        cp = self._cp(P)
        self.lib.gzo_potential(t, C.byref(cp), _ptr(tg, _up), tg.shape[0], theta, G,
                               _ptr(ao, _dp), _ptr(out, _dp))
        return out

    def density(self, t, P: HostParticles, targets, des_ngb=32.0, ngb_tol=1.0, h_start=None,
                n_dims=3, type=None, cond_prev=None, tau=0.0):
        """shmem::density (hydro.cc:150-327): (h, rho, nngb, iters) per target."""
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        nt = tg.shape[0]
        out = (np.zeros(nt), np.zeros(nt), np.zeros(nt, np.int32), np.zeros(nt, np.int32))
        ty = None if type is None else np.ascontiguousarray(type, dtype=np.uint8)
        cp = self._cp(P)
        self.lib.gzo_density(t, C.byref(cp), None if ty is None else ty.ctypes.data_as(C.c_void_p),
                             _ptr(tg, _up), nt, des_ngb, ngb_tol,
                             None if h_start is None else _ptr(_f64(h_start), _dp), n_dims,
                             None if cond_prev is None else _ptr(_f64(cond_prev), _dp), tau,
                             _ptr(out[0], _dp), _ptr(out[1], _dp), _ptr(out[2], _ip),
                             _ptr(out[3], _ip))
        return out

    def kernel_w(self, r, h, dim=3):
        return self.lib.gzo_kernel_w(r, h, dim)

    def kernel_dwdh(self, r, h, dim=3):
        return self.lib.gzo_kernel_dwdh(r, h, dim)

    def build_vel(self, P: HostParticles, v, seed: int = -1):
        cp = self._cp(P)
        v = tuple(_f64(a) for a in v)
        t = self.lib.gzo_build_vel(C.byref(cp), seed, *(_ptr(a, _dp) for a in v))
        return t

    def kick_nodes(self, t, active, v, m, vlast, tau):
        act = np.ascontiguousarray(active, dtype=np.uint32)
        v = tuple(_f64(a) for a in v)
        self.lib.gzo_kick_nodes(t, _ptr(act, _up), act.shape[0], *(_ptr(a, _dp) for a in v),
                                _ptr(_f64(m), _dp), *_opt3(vlast), tau)

    def refresh_walk_nodes(self, t, tau):
        self.lib.gzo_refresh_walk_nodes(t, tau)

    def raise_soft(self, t, idx, soft):
        ix = np.ascontiguousarray(idx, dtype=np.uint32)
        self.lib.gzo_raise_soft(t, _ptr(ix, _up), ix.shape[0], _ptr(_f64(soft), _dp))

    def reuse(self, t) -> ReuseState:
        nn = self._tree_arrays(t).nnodes
        r = _reuse_arrays(nn)
        self.lib.gzo_tree_reuse(t, _ptr(r.vmax, _fp), _ptr(r.vcom, _dp), _ptr(r.dp, _dp),
                                _ptr(r.dq, _dp), _ptr(r.wcom, _dp), _ptr(r.wmass, _dp))
        self.lib.gzo_tree_nodes(t, *([None] * 15), _ptr(r.wn_s, _dp), _ptr(r.wn_soft, _fp),
                                _ptr(r.wn_len, _fp))
        return r

    def accel(self, t, P: HostParticles, targets, theta, G=1.0, aold=None):
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        nt = tg.shape[0]
        ax, ay, az = (np.zeros(nt) for _ in range(3))
        ao = None if aold is None else _f64(aold)
        cp = self._cp(P)
        self.lib.gzo_accel(t, C.byref(cp), _ptr(tg, _up), nt, theta, G, _ptr(ao, _dp),
                           _ptr(ax, _dp), _ptr(ay, _dp), _ptr(az, _dp))
        return ax, ay, az

    def accel_brute(self, P: HostParticles, targets, G=1.0):
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        nt = tg.shape[0]
        ax, ay, az = (np.zeros(nt) for _ in range(3))
        cp = self._cp(P)
        self.lib.gzo_accel_brute(C.byref(cp), _ptr(tg, _up), nt, G, _ptr(ax, _dp),
                                 _ptr(ay, _dp), _ptr(az, _dp))
        return ax, ay, az

    def audit_tree(self, T, P: HostParticles, vel=None, vmax=None, tau=0.0, check_com=True):
        """audit_tree (tree.cc:973-1119) over tree arrays T (a TreeArrays, or the dict a device
        tree downloads to): (violations, first message)."""
        g = (lambda k: T[k]) if isinstance(T, dict) else (lambda k: getattr(T, k))
        keep = {}

        def arr(k, dt, ptype):
            a = np.ascontiguousarray(g(k), dtype=dt)
            keep[k] = a
            return _ptr(a, ptype)

        vm = None if vmax is None else np.ascontiguousarray(vmax, dtype=np.float32)
        view = _GzoTreeView(
            int(g("nnodes")), int(g("n_live")), int(g("root")),
            *(arr(k, np.int32, _ip) for k in ("first", "next", "parent", "plo", "phi")),
            *(arr(k, np.float64, _dp) for k in ("mass", "cx", "cy", "cz", "size", "gcx", "gcy",
                                                 "gcz")),
            arr("wn_soft", np.float32, _fp), _ptr(vm, _fp), None, float(tau),
            arr("orderbuf", np.uint32, _up), arr("leaf_of", np.int32, _ip))
        v = (None, None, None) if vel is None else tuple(_f64(a) for a in vel)
        buf = C.create_string_buffer(512)
        cp = self._cp(P)
        bad = self.lib.gzo_audit_tree(C.byref(view), C.byref(cp), *(_ptr(a, _dp) for a in v),
                                      1 if check_com else 0, buf, 512)
        return int(bad), buf.value.decode()

    def direct_sum(self, P: HostParticles, targets, G=1.0, fp64=False, box=0.0):
        """grav_cuda.cu's direct sum: float-narrowed (fp64=False) or FP64, optional box."""
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        nt = tg.shape[0]
        ax, ay, az = (np.zeros(nt) for _ in range(3))
        cp = self._cp(P)
        fn = self.lib.gzo_direct_sum_box if fp64 else self.lib.gzo_direct_sum_f32
        fn(C.byref(cp), _ptr(tg, _up), nt, G, float(box), _ptr(ax, _dp), _ptr(ay, _dp),
           _ptr(az, _dp))
        return ax, ay, az


class Reference(_EventsMixin, _Lib):
    """The reference ``shmem/tree.cc``, compiled into ``_ref/libgzref.so``."""

    prefix = "gzref_"

    def __init__(self, path: str | None = None):
        path = path or reference_path()
        self.path = path
        super().__init__(path)
        L = self.lib
        L.gzref_particles_new.restype = C.c_void_p
        L.gzref_particles_new.argtypes = [C.c_size_t, _dp, _dp, _dp, _dp, _dp]
        L.gzref_particles_free.argtypes = [C.c_void_p]
        L.gzref_build.restype = C.c_void_p
        L.gzref_build.argtypes = [C.c_void_p, C.c_longlong]
        L.gzref_tree_times.argtypes = [C.c_void_p, _dp]
        L.gzref_accel_grouped.argtypes = [C.c_void_p, C.c_void_p, _up, C.c_size_t, C.c_double,
                                          C.c_double, C.c_int, _dp, _dp, _dp, _dp]
        L.gzref_accel.argtypes = [C.c_void_p, C.c_void_p, _up, C.c_size_t, C.c_double,
                                  C.c_double, _dp, _dp, _dp, _dp]
        L.gzref_potential.argtypes = [C.c_void_p, C.c_void_p, _up, C.c_size_t, C.c_double,
                                      C.c_double, _dp, _dp]
        L.gzref_accel_brute.argtypes = [C.c_void_p, _up, C.c_size_t, C.c_double, _dp, _dp, _dp]
        L.gzref_audit.restype = C.c_longlong
        L.gzref_audit.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t]
        L.gzref_accel_grouped_tj.argtypes = [C.c_void_p, C.c_void_p, _up, C.c_size_t,
                                             C.c_double, C.c_double, C.c_int, _dp, _dp, _dp, _dp,
                                             _dp, _dp, _dp, _dp, _dp]
        L.gzref_ngb_search.restype = C.c_size_t
        L.gzref_ngb_search.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, _dp, _dp, C.c_size_t,
                                       C.c_double, C.POINTER(C.c_size_t), _up, C.c_size_t]
        L.gzref_init_hmax.argtypes = [C.c_void_p, _dp, C.c_size_t, C.c_size_t,
                                      C.POINTER(C.c_float)]
        L.gzref_raise_hmax.argtypes = [C.c_void_p, _up, C.c_size_t, _dp, C.POINTER(C.c_float)]
        L.gzref_ngb_search_reverse.restype = C.c_size_t
        L.gzref_ngb_search_reverse.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_size_t, _dp, _dp,
                                               _dp, _dp, C.c_size_t, C.c_double,
                                               C.POINTER(C.c_size_t), _up, C.c_size_t]
        L.gzref_build_vel.restype = C.c_void_p
        L.gzref_build_vel.argtypes = [C.c_void_p, C.c_longlong, _dp, _dp, _dp]
        L.gzref_kick_nodes.argtypes = [C.c_void_p, _up, C.c_size_t, _dp, _dp, _dp, _dp, _dp, _dp,
                                       _dp, C.c_double]
        L.gzref_refresh_walk_nodes.argtypes = [C.c_void_p, C.c_double]
        L.gzref_raise_soft.argtypes = [C.c_void_p, _up, C.c_size_t, _dp]
        L.gzref_tree_reuse.argtypes = [C.c_void_p, _fp, _dp, _dp, _dp, _dp, _dp, _dp, _fp, _fp]
        L.gzref_mass_events.argtypes = [C.c_void_p, _up, C.c_size_t] + [_dp] * 7 + [
            C.c_double, C.c_double]
        L.gzref_shift_nodes.argtypes = [C.c_void_p, _up, C.c_size_t, _dp, _dp, _dp, _dp]
        L.gzref_raise_pad.argtypes = [C.c_void_p, _up, C.c_size_t, _dp]
        L.gzref_raise_vmax.argtypes = [C.c_void_p, _up, C.c_size_t, _fp]
        L.gzref_tree_pads.argtypes = [C.c_void_p, _fp, _dp]
        L.gzref_density.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _up, C.c_size_t,
                                    C.c_double, C.c_double, _dp, C.c_int, _dp, C.c_double, _dp,
                                    _dp, _ip, _ip]
        L.gzref_max_threads.restype = C.c_int
        L.gzref_set_threads.argtypes = [C.c_int]
        self._handles = {}

    def particles(self, P: HostParticles):
        return self.lib.gzref_particles_new(P.n, _ptr(P.x, _dp), _ptr(P.y, _dp), _ptr(P.z, _dp),
                                            _ptr(P.m, _dp), _ptr(P.soft, _dp))

    def free_particles(self, h):
        self.lib.gzref_particles_free(h)

    def build(self, rp, seed: int = -1):
        return self.lib.gzref_build(rp, seed)

    def free(self, t):
        self.lib.gzref_tree_free(t)

    def density(self, t, rp, targets, des_ngb=32.0, ngb_tol=1.0, h_start=None, n_dims=3,
                type=None, cond_prev=None, tau=0.0):
        """shmem::density (hydro.cc:150-327): (h, rho, nngb, iters) per target."""
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        nt = tg.shape[0]
        out = (np.zeros(nt), np.zeros(nt), np.zeros(nt, np.int32), np.zeros(nt, np.int32))
        ty = None if type is None else np.ascontiguousarray(type, dtype=np.uint8)
        self.lib.gzref_density(t, rp, None if ty is None else ty.ctypes.data_as(C.c_void_p),
                               _ptr(tg, _up), nt, des_ngb, ngb_tol,
                               None if h_start is None else _ptr(_f64(h_start), _dp), n_dims,
                               None if cond_prev is None else _ptr(_f64(cond_prev), _dp), tau,
                               _ptr(out[0], _dp), _ptr(out[1], _dp), _ptr(out[2], _ip),
                               _ptr(out[3], _ip))
        return out

    def build_vel(self, rp, v, seed: int = -1):
        v = tuple(_f64(a) for a in v)
        return self.lib.gzref_build_vel(rp, seed, *(_ptr(a, _dp) for a in v))

    def kick_nodes(self, t, active, v, m, vlast, tau):
        act = np.ascontiguousarray(active, dtype=np.uint32)
        v = tuple(_f64(a) for a in v)
        self.lib.gzref_kick_nodes(t, _ptr(act, _up), act.shape[0], *(_ptr(a, _dp) for a in v),
                                  _ptr(_f64(m), _dp), *_opt3(vlast), tau)

    def refresh_walk_nodes(self, t, tau):
        self.lib.gzref_refresh_walk_nodes(t, tau)

    def raise_soft(self, t, idx, soft):
        ix = np.ascontiguousarray(idx, dtype=np.uint32)
        self.lib.gzref_raise_soft(t, _ptr(ix, _up), ix.shape[0], _ptr(_f64(soft), _dp))

    def reuse(self, t) -> ReuseState:
        nn = self._tree_arrays(t).nnodes
        r = _reuse_arrays(nn)
        self.lib.gzref_tree_reuse(t, _ptr(r.vmax, _fp), _ptr(r.vcom, _dp), _ptr(r.dp, _dp),
                                  _ptr(r.dq, _dp), _ptr(r.wcom, _dp), _ptr(r.wmass, _dp),
                                  _ptr(r.wn_s, _dp), _ptr(r.wn_soft, _fp), _ptr(r.wn_len, _fp))
        return r

    def times(self, t):
        out = np.zeros(7)
        self.lib.gzref_tree_times(t, _ptr(out, _dp))
        return dict(zip(("bbox", "keys", "sort", "recurse", "links", "pack", "total"), out))

    def tree(self, t, n: int) -> TreeArrays:
        ta = self._tree_arrays(t)
        ta.orderbuf = np.zeros(ta.n_live, dtype=np.uint32)
        ta.rank = np.zeros(n, dtype=np.uint32)
        ta.leaf_of = np.zeros(n, dtype=np.int32)
        self.lib.gzref_tree_particles(t, _ptr(ta.orderbuf, _up), _ptr(ta.rank, _up),
                                      _ptr(ta.leaf_of, _ip))
        return ta

    def accel_grouped(self, t, rp, targets, theta, G=1.0, batch=0, aold=None):
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        nt = tg.shape[0]
        ax, ay, az = (np.zeros(nt) for _ in range(3))
        ao = None if aold is None else _f64(aold)
        self.lib.gzref_accel_grouped(t, rp, _ptr(tg, _up), nt, theta, G, batch, _ptr(ao, _dp),
                                     _ptr(ax, _dp), _ptr(ay, _dp), _ptr(az, _dp))
        return ax, ay, az

    def ngb_search(self, t, rp, centres, radius, tau=0.0):
        c = np.ascontiguousarray(centres, dtype=np.float64)
        cs = [np.ascontiguousarray(c[:, d]) for d in range(3)]
        r = _f64(radius)
        nq = r.shape[0]
        start = np.zeros(nq + 1, dtype=np.uint64)
        sp = start.ctypes.data_as(C.POINTER(C.c_size_t))
        total = self.lib.gzref_ngb_search(t, rp, *(_ptr(a, _dp) for a in cs), _ptr(r, _dp), nq,
                                          tau, sp, None, 0)
        found = np.zeros(max(int(total), 1), dtype=np.uint32)
        self.lib.gzref_ngb_search(t, rp, *(_ptr(a, _dp) for a in cs), _ptr(r, _dp), nq, tau, sp,
                                  _ptr(found, _up), found.shape[0])
        return start.astype(np.int64), found[:int(total)]

    def init_hmax(self, t, h, n_gas):
        hh = _f64(h)
        sz = np.zeros(3, dtype=np.int64)
        self.lib.gzref_tree_sizes(t, _ptr(sz, _llp))
        out = np.zeros(int(sz[0]), dtype=np.float32)
        self.lib.gzref_init_hmax(t, _ptr(hh, _dp), hh.shape[0], int(n_gas),
                                 out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def raise_hmax(self, t, idx, h):
        ix = np.ascontiguousarray(idx, dtype=np.uint32)
        hh = _f64(h)
        sz = np.zeros(3, dtype=np.int64)
        self.lib.gzref_tree_sizes(t, _ptr(sz, _llp))
        out = np.zeros(int(sz[0]), dtype=np.float32)
        self.lib.gzref_raise_hmax(t, _ptr(ix, _up), ix.shape[0], _ptr(hh, _dp),
                                  out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def ngb_search_reverse(self, t, rp, h, n_gas, centres, hq, tau=0.0):
        c = np.ascontiguousarray(centres, dtype=np.float64)
        cs = [np.ascontiguousarray(c[:, d]) for d in range(3)]
        hh, r = _f64(h), _f64(hq)
        nq = r.shape[0]
        start = np.zeros(nq + 1, dtype=np.uint64)
        sp = start.ctypes.data_as(C.POINTER(C.c_size_t))
        args = (t, rp, _ptr(hh, _dp), int(n_gas), *(_ptr(a, _dp) for a in cs), _ptr(r, _dp), nq,
                tau, sp)
        total = self.lib.gzref_ngb_search_reverse(*args, None, 0)
        found = np.zeros(max(int(total), 1), dtype=np.uint32)
        self.lib.gzref_ngb_search_reverse(*args, _ptr(found, _up), found.shape[0])
        return start.astype(np.int64), found[:int(total)]

    def accel_grouped_tj(self, t, rp, targets, theta, G=1.0, batch=0, aold=None, vel=None,
                         jerk=False):
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        nt = tg.shape[0]
        ax, ay, az = (np.zeros(nt) for _ in range(3))
        tid = np.zeros((6, nt))
        jk = np.zeros((3, nt)) if jerk else None
        ao = None if aold is None else _f64(aold)
        v = (None, None, None) if vel is None else tuple(_f64(a) for a in vel)
        self.lib.gzref_accel_grouped_tj(t, rp, _ptr(tg, _up), nt, theta, G, batch, _ptr(ao, _dp),
                                        *(_ptr(a, _dp) for a in v), _ptr(ax, _dp), _ptr(ay, _dp),
                                        _ptr(az, _dp), _ptr(tid, _dp), _ptr(jk, _dp))
        return ax, ay, az, tid, jk

    def accel(self, t, rp, targets, theta, G=1.0, aold=None):
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        nt = tg.shape[0]
        ax, ay, az = (np.zeros(nt) for _ in range(3))
        ao = None if aold is None else _f64(aold)
        self.lib.gzref_accel(t, rp, _ptr(tg, _up), nt, theta, G, _ptr(ao, _dp), _ptr(ax, _dp),
                             _ptr(ay, _dp), _ptr(az, _dp))
        return ax, ay, az

    def potential(self, t, rp, targets, theta, G=1.0, aold=None):
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        out = np.zeros(tg.shape[0])
        ao = None if aold is None else _f64(aold)
        self.lib.gzref_potential(t, rp, _ptr(tg, _up), tg.shape[0], theta, G, _ptr(ao, _dp),
                                 _ptr(out, _dp))
        return out

    def accel_brute(self, rp, targets, G=1.0):
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        nt = tg.shape[0]
        ax, ay, az = (np.zeros(nt) for _ in range(3))
        self.lib.gzref_accel_brute(rp, _ptr(tg, _up), nt, G, _ptr(ax, _dp), _ptr(ay, _dp),
                                   _ptr(az, _dp))
        return ax, ay, az

    def audit(self, t, rp) -> tuple[int, str]:
        buf = C.create_string_buffer(512)
        bad = self.lib.gzref_audit(t, rp, buf, 512)
        return int(bad), buf.value.decode()

    def max_threads(self) -> int:
        return int(self.lib.gzref_max_threads())

    def set_threads(self, n: int):
        self.lib.gzref_set_threads(int(n))


def have_reference() -> bool:
    return os.path.exists(REF_SO)


# -------------------------------------------------------------------------------------------------
# Canonical node order: DFS preorder with children in octant order. Two builds of the same tree
# agree node-for-node in this order whatever order their builders allocated ids in.
# -------------------------------------------------------------------------------------------------

def preorder_from_threaded(first: np.ndarray, nxt: np.ndarray, root: int) -> np.ndarray:
    """Preorder over a tree threaded the reference's way: the children of X are the chain
    first[X], next[first[X]], ... up to (excluding) next[X] (tree.cc:340-348, 1004)."""
    out = []
    stack = [root]
    while stack:
        x = stack.pop()
        out.append(x)
        if first[x] < 0:
            continue
        kids = []
        c = first[x]
        while c != nxt[x]:
            kids.append(c)
            c = nxt[c]
        stack.extend(reversed(kids))
    return np.asarray(out, dtype=np.int64)


def preorder_from_children(first: np.ndarray, nchild: np.ndarray, root: int) -> np.ndarray:
    """Preorder over a tree whose children of X are first[X] .. first[X] + nchild[X] - 1."""
    out = []
    stack = [root]
    while stack:
        x = stack.pop()
        out.append(x)
        if nchild[x] > 0:
            stack.extend(range(first[x] + nchild[x] - 1, first[x] - 1, -1))
    return np.asarray(out, dtype=np.int64)
