"""Host mirror of the shared-memory GIZMO tree call surface (shmem/tree.h:384-451) over the device
library. Device memory and streams come from PyTorch; every computation runs in libgzgrav.so.

    P = Particles(x, y, z, m, soft)          # torch tensors on a HIP device, float64
    T = build(P)                             # shmem::build            (tree.h:384-385)
    ax, ay, az = accel_grouped(T, P, targets, theta, G, batch=0, aold=None)
                                             # shmem::accel_grouped    (tree.h:428-440)
    ax, ay, az = accel(T, P, targets, theta, G, aold=None)
                                             # shmem::accel            (tree.h:392-394)
    ax, ay, az = accel_brute(P, targets, G)  # shmem::accel_brute      (tree.h:450-451)

As in the reference, accelerations include G, ``aold`` is per target slot in no-G units, and the
softening is the cubic-spline kernel extent. Errors raise ``GzError`` (the reference has no error
channel; its CUDA path exits the process, shmem/grav_cuda.cu:239-243).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _dev_f64(a, device):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)


def _dev_u32(a, device):
    """Particle indices as a contiguous uint32 device tensor."""
    if isinstance(a, torch.Tensor):
        if a.dtype == torch.uint32:
            return a.to(device).contiguous()
        if a.dtype != torch.int32:
            a = a.to(torch.int64).to(torch.int32)
        return a.to(device).contiguous().view(torch.uint32)
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.int64).astype(np.uint32).view(np.int32))
    return torch.as_tensor(arr, device=device).view(torch.uint32)


class Context:
    """One gz_ctx per device.

    The library orders its work on its own stream (the context's). Torch produces and consumes
    the tensors passed in, so every call here first waits for torch's current stream (inputs are
    complete) and, when the call returns before its kernels finish, waits for the library's
    stream before handing results back (outputs are complete)."""

    _by_device: dict = {}

    def __init__(self, device: int = 0):
        lib = L.load()
        self.device = device
        h = C.c_void_p()
        L.check(lib.gz_ctx_create(device, None, C.byref(h)), "gz_ctx_create")
        self.handle = h

    def enter(self):
        """Wait for torch's pending work on this device (the inputs)."""
        torch.cuda.current_stream(self.device).synchronize()

    @classmethod
    def get(cls, device: int = 0) -> "Context":
        ctx = cls._by_device.get(device)
        if ctx is None:
            ctx = cls._by_device[device] = Context(device)
        return ctx

    def synchronize(self):
        L.check(L.load().gz_ctx_synchronize(self.handle), "gz_ctx_synchronize")


@dataclass
class Particles:
    """SoA particle set on the device (shmem::Particles, tree.h:36-50)."""

    x: torch.Tensor
    y: torch.Tensor
    z: torch.Tensor
    m: torch.Tensor
    soft: torch.Tensor | None = None
    device: str | torch.device = "cuda:0"
    _c: L.gz_particles = field(default=None, repr=False)

    def __post_init__(self):
        dev = torch.device(self.device)
        self.x, self.y, self.z, self.m = (_dev_f64(v, dev) for v in (self.x, self.y, self.z,
                                                                        self.m))
        if self.soft is not None:
            self.soft = _dev_f64(self.soft, dev)
        self.device = dev
        self._c = L.gz_particles(self.n, self.x.data_ptr(), self.y.data_ptr(), self.z.data_ptr(),
                                 self.m.data_ptr(),
                                 None if self.soft is None else self.soft.data_ptr(), None, None)

    @property
    def n(self) -> int:
        return int(self.m.shape[0])

    @classmethod
    def from_numpy(cls, x, y, z, m, soft=None, device="cuda:0"):
        return cls(x, y, z, m, soft, device=device)

    def cstruct(self):
        return C.byref(self._c)


class Tree:
    """A device-resident tree (gz_tree). Node ids are level-major; see download()."""

    def __init__(self, handle, ctx: Context, particles: Particles):
        self.handle = handle
        self.ctx = ctx
        self.P = particles
        info = L.gz_tree_info()
        L.check(L.load().gz_tree_get_info(handle, C.byref(info)), "gz_tree_get_info")
        self.info = info

    def __del__(self):
        h = getattr(self, "handle", None)
        lib = getattr(L, "_lib", None) if L is not None else None
        if h is not None and lib is not None:
            lib.gz_tree_free(h)
            self.handle = None

    @property
    def nnodes(self) -> int:
        return int(self.info.nnodes)

    @property
    def n_live(self) -> int:
        return int(self.info.n_live)

    @property
    def root(self) -> int:
        return int(self.info.root)

    def times(self) -> dict:
        t = self.info.times
        return {k: getattr(t, k) for k, _ in L.gz_build_times._fields_}

    def device_orderbuf(self) -> torch.Tensor:
        """A device copy of orderbuf: every live particle in Morton order."""
        return order_targets(self, None)

    def download(self) -> dict:
        """Host copies of the node and particle arrays (numpy)."""
        lib = L.load()
        nn, n, nl = self.nnodes, self.P.n, self.n_live
        ints = {k: np.zeros(nn, dtype=np.int32) for k in
                ("first", "next", "nchild", "plo", "phi", "parent", "level")}
        dbls = {k: np.zeros(nn, dtype=np.float64) for k in
                ("size", "mass", "cx", "cy", "cz", "soft", "delta", "gcx", "gcy", "gcz", "wn_s")}
        flts = {k: np.zeros(nn, dtype=np.float32) for k in ("wn_soft", "wn_len")}
        ip = lambda a: a.ctypes.data_as(L._ip)  # noqa: E731
        dp = lambda a: a.ctypes.data_as(L._dp)  # noqa: E731
        fp = lambda a: a.ctypes.data_as(L._fp)  # noqa: E731
        L.check(lib.gz_tree_download_nodes(
            self.handle, *[ip(ints[k]) for k in ("first", "next", "nchild", "plo", "phi",
                                                  "parent", "level")],
            *[dp(dbls[k]) for k in ("size", "mass", "cx", "cy", "cz", "soft", "delta", "gcx",
                                    "gcy", "gcz", "wn_s")],
            fp(flts["wn_soft"]), fp(flts["wn_len"])), "gz_tree_download_nodes")
        orderbuf = np.zeros(nl, dtype=np.uint32)
        rank = np.zeros(n, dtype=np.uint32)
        leaf_of = np.zeros(n, dtype=np.int32)
        skey = np.zeros(nl, dtype=np.uint64)
        L.check(lib.gz_tree_download_particles(
            self.handle, orderbuf.ctypes.data_as(L._up), rank.ctypes.data_as(L._up),
            ip(leaf_of), skey.ctypes.data_as(L._u64p)), "gz_tree_download_particles")
        out = dict(nnodes=nn, root=self.root, n_live=nl, orderbuf=orderbuf, rank=rank,
                   leaf_of=leaf_of, skey=skey)
        out.update(ints)
        out.update(dbls)
        out.update(flts)
        return out

    def kick_nodes(self, active, v, m=None, vlast=None, tau: float = 0.0):
        """The kick feed (mfm.cc:3798-3811): raise_vmax for every active particle, and with
        vlast (three float64 device tensors, updated in place) kick_node with m (v - vlast)."""
        act = _dev_u32(active, self.P.device)
        vt = tuple(_dev_f64(a, self.P.device) for a in v)
        mt = self.P.m if m is None else _dev_f64(m, self.P.device)
        vl = (None, None, None)
        if vlast is not None:
            for a in vlast:
                assert a.dtype == torch.float64 and a.is_contiguous() and a.device == self.P.device
            vl = tuple(a.data_ptr() for a in vlast)
        self.ctx.enter()
        L.check(L.load().gz_tree_kick_nodes(self.ctx.handle, self.handle, _ptr(act),
                                            int(act.shape[0]), *(a.data_ptr() for a in vt),
                                            mt.data_ptr(), *vl, float(tau)),
                "gz_tree_kick_nodes")
        self.ctx.synchronize()

    def drift(self, dt: float):
        """gz_tree_drift: the sorted copy of the sources drifted by its sorted velocities (a tree
        built with vel=, kicks reported through kick_nodes)."""
        self.ctx.enter()
        L.check(L.load().gz_tree_drift(self.ctx.handle, self.handle, float(dt)),
                "gz_tree_drift")
        self.ctx.synchronize()

    def refresh_walk_nodes(self, tau: float):
        """Tree::refresh_walk_nodes (tree.cc:260-274)."""
        self.ctx.enter()
        L.check(L.load().gz_tree_refresh_walk_nodes(self.ctx.handle, self.handle, float(tau)),
                "gz_tree_refresh_walk_nodes")
        self.ctx.synchronize()

    def init_hmax(self, h, n_gas: int):
        """Tree::init_hmax (tree.cc:332-338): node hmax from the gas kernel radii h (by particle;
        the first n_gas particles are gas), for ngb_search_reverse."""
        hh = _dev_f64(h, self.P.device)
        self.ctx.enter()
        L.check(L.load().gz_tree_init_hmax(self.ctx.handle, self.handle, _ptr(hh),
                                           int(hh.shape[0]), int(n_gas)), "gz_tree_init_hmax")
        self.ctx.synchronize()

    def raise_hmax(self, idx, h):
        """Tree::raise_hmax (tree.cc:293-301) for the particles idx; h indexed by particle."""
        ix = _dev_u32(idx, self.P.device)
        hh = _dev_f64(h, self.P.device)
        if int(hh.shape[0]) < self.P.n:
            raise ValueError(f"raise_hmax: h holds {int(hh.shape[0])} values, the kernel reads "
                             f"it by particle index (n={self.P.n})")
        self.ctx.enter()
        L.check(L.load().gz_tree_raise_hmax(self.ctx.handle, self.handle, _ptr(ix),
                                            int(ix.shape[0]), _ptr(hh)), "gz_tree_raise_hmax")
        self.ctx.synchronize()

    def raise_soft(self, idx, soft):
        """Tree::raise_soft (tree.cc:311-319) for the particles idx; soft indexed by particle."""
        ix = _dev_u32(idx, self.P.device)
        st = _dev_f64(soft, self.P.device)
        self.ctx.enter()
        L.check(L.load().gz_tree_raise_soft(self.ctx.handle, self.handle, _ptr(ix),
                                            int(ix.shape[0]), st.data_ptr()),
                "gz_tree_raise_soft")
        self.ctx.synchronize()

    def mass_events(self, idx, dm, x, v, tau: float, box: float = 0.0):
        """Tree::mass_event (tree.cc:239-251) for each event k in order: a particle of mass dm[k]
        (negative removes one) at x[:, k] moving at v[:, k], appearing at tau since the build
        (tau < 0: the tree's age), in the build-time leaf of particle idx[k]."""
        dev = self.P.device
        ix = _dev_u32(idx, dev)
        arrs = [_dev_f64(dm, dev)] + [_dev_f64(a, dev) for a in x] + [_dev_f64(a, dev) for a in v]
        self.ctx.enter()
        L.check(L.load().gz_tree_mass_events(self.ctx.handle, self.handle, _ptr(ix),
                                             int(ix.shape[0]), *(a.data_ptr() for a in arrs),
                                             float(tau), float(box)),
                "gz_tree_mass_events")
        self.ctx.synchronize()

    def shift_nodes(self, idx, m, dx):
        """Tree::shift_node (tree.cc:253-258) for each event k in order: a particle of mass m[k]
        in the build-time leaf of particle idx[k] displaced by dx[:, k]."""
        dev = self.P.device
        ix = _dev_u32(idx, dev)
        arrs = [_dev_f64(m, dev)] + [_dev_f64(a, dev) for a in dx]
        self.ctx.enter()
        L.check(L.load().gz_tree_shift_nodes(self.ctx.handle, self.handle, _ptr(ix),
                                             int(ix.shape[0]), *(a.data_ptr() for a in arrs)),
                "gz_tree_shift_nodes")
        self.ctx.synchronize()

    def raise_pad(self, idx, d):
        """Tree::raise_pad (tree.cc:323-330) for each event k in order: a displacement d[k] of a
        particle in the build-time leaf of particle idx[k]."""
        ix = _dev_u32(idx, self.P.device)
        dt = _dev_f64(d, self.P.device)
        self.ctx.enter()
        L.check(L.load().gz_tree_raise_pad(self.ctx.handle, self.handle, _ptr(ix),
                                           int(ix.shape[0]), dt.data_ptr()),
                "gz_tree_raise_pad")
        self.ctx.synchronize()

    def raise_vmax(self, idx, speed):
        """Tree::raise_vmax (tree.cc:276-285) to speed[k] (float32, >= 0) for the build-time leaf
        of particle idx[k]."""
        ix = _dev_u32(idx, self.P.device)
        sp = torch.as_tensor(np.ascontiguousarray(speed, dtype=np.float32), device=self.P.device)
        self.ctx.enter()
        L.check(L.load().gz_tree_raise_vmax(self.ctx.handle, self.handle, _ptr(ix),
                                            int(ix.shape[0]), sp.data_ptr()),
                "gz_tree_raise_vmax")
        self.ctx.synchronize()

    def download_pads(self):
        """Host copies of xpad (nnodes float32) and dmass (nnodes float64)."""
        nn = self.nnodes
        xpad, dmass = np.zeros(nn, np.float32), np.zeros(nn)
        L.check(L.load().gz_tree_download_pads(self.handle, xpad.ctypes.data, dmass.ctypes.data),
                "gz_tree_download_pads")
        return xpad, dmass

    def download_reuse(self) -> dict:
        """Host copies of vmax (nnodes) and vcom, dp, dq (nnodes x 3)."""
        nn = self.nnodes
        out = dict(vmax=np.zeros(nn, np.float32), vcom=np.zeros((nn, 3)), dp=np.zeros((nn, 3)),
                   dq=np.zeros((nn, 3)))
        L.check(L.load().gz_tree_download_reuse(self.handle,
                                                *(out[k].ctypes.data for k in
                                                  ("vmax", "vcom", "dp", "dq"))),
                "gz_tree_download_reuse")
        return out

    def update_sources(self, P: Particles | None = None):
        """Re-read source positions/masses/softenings into the tree's sorted copy."""
        P = P or self.P
        self.ctx.enter()
        L.check(L.load().gz_tree_update_sources(self.ctx.handle, self.handle, P.cstruct()),
                "gz_tree_update_sources")
        self.ctx.synchronize()


def build(P: Particles, randomize_seed: int = -1, shift=(0.0, 0.0, 0.0), max_nodes: int = 0,
          clusters=None, vel=None, ctx: Context | None = None) -> Tree:
    """shmem::build (tree.h:384-385, tree.cc:350-526). clusters: particle offsets (length
    n_clusters + 1) of independent clusters, each given its own tree (a forest). vel: particle
    velocities (vx, vy, vz), kept as the reference does with want_vcom for reusing the tree."""
    ctx = ctx or Context.get(P.device.index or 0)
    cs, nc = None, 0
    if clusters is not None:
        cl = [int(v) for v in clusters]
        cs = (C.c_size_t * len(cl))(*cl)
        nc = len(cl) - 1
    vt = (None, None, None) if vel is None else tuple(_dev_f64(v, P.device) for v in vel)
    opts = L.gz_build_opts(int(randomize_seed), (C.c_double * 3)(*shift), int(max_nodes), cs, nc,
                           *(None if v is None else v.data_ptr() for v in vt))
    h = C.c_void_p()
    ctx.enter()
    L.check(L.load().gz_tree_build(ctx.handle, P.cstruct(), C.byref(opts), C.byref(h)),
            "gz_tree_build")
    T = Tree(h, ctx, P)
    T._vel = vt  # the build reads them on the stream; keep them alive until it has
    ctx.synchronize()
    return T


def _outputs(n, device):
    return tuple(torch.empty(n, dtype=torch.float64, device=device) for _ in range(3))


class Model:
    """The learned opening-parameter network (gz_model): 16 -> 64 -> 64 -> 1, GELU, bf16 weights
    and FP32 accumulation on the device, FP32 master weights trained with Adam."""

    N_PARAMS = L.MLP_PARAMS
    N_FEATURES = L.MLP_FEATURES

    def __init__(self, params=None, ctx: Context | None = None):
        self.ctx = ctx or Context.get(0)
        h = C.c_void_p()
        arr = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        if arr is not None and arr.shape != (self.N_PARAMS,):
            raise ValueError(f"expected {self.N_PARAMS} parameters, got {arr.shape}")
        L.check(L.load().gz_model_create(self.ctx.handle,
                                         None if arr is None else arr.ctypes.data_as(L._fp),
                                         C.byref(h)), "gz_model_create")
        self.handle = h
        self.calibration = None

    def __del__(self):
        h = getattr(self, "handle", None)
        lib = getattr(L, "_lib", None) if L is not None else None
        if h is not None and lib is not None:
            lib.gz_model_free(h)
            self.handle = None

    def params(self) -> np.ndarray:
        out = np.zeros(self.N_PARAMS, dtype=np.float32)
        L.check(L.load().gz_model_get_params(self.handle, out.ctypes.data_as(L._fp)),
                "gz_model_get_params")
        return out

    def set_params(self, params):
        arr = np.ascontiguousarray(params, dtype=np.float32)
        L.check(L.load().gz_model_set_params(self.handle, arr.ctypes.data_as(L._fp)),
                "gz_model_set_params")

    def set_calibration(self, feat_lo, feat_hi, label_lo, label_hi, margin, log10_alpha_ref):
        lo = np.ascontiguousarray(feat_lo, dtype=np.float32)
        hi = np.ascontiguousarray(feat_hi, dtype=np.float32)
        L.check(L.load().gz_model_set_calibration(self.handle, lo.ctypes.data_as(L._fp),
                                                  hi.ctypes.data_as(L._fp), float(label_lo),
                                                  float(label_hi), float(margin),
                                                  float(log10_alpha_ref)),
                "gz_model_set_calibration")
        self.calibration = dict(feat_lo=lo, feat_hi=hi, label_lo=float(label_lo),
                                label_hi=float(label_hi), margin=float(margin),
                                log10_alpha_ref=float(log10_alpha_ref))

    def set_input_map(self, mean, scale):
        """Inputs enter the first layer as (x - mean) * scale."""
        mu = np.ascontiguousarray(mean, dtype=np.float32)
        sc = np.ascontiguousarray(scale, dtype=np.float32)
        L.check(L.load().gz_model_set_input_map(self.handle, mu.ctypes.data_as(L._fp),
                                                sc.ctypes.data_as(L._fp)),
                "gz_model_set_input_map")

    def set_loss(self, quantile: float = 0.0):
        """0: mean squared error; 0 < q < 1: pinball loss of the q-quantile."""
        L.check(L.load().gz_model_set_loss(self.handle, float(quantile)), "gz_model_set_loss")

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        X = X.to(torch.float32).contiguous()
        out = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
        self.ctx.enter()
        L.check(L.load().gz_model_forward(self.ctx.handle, self.handle, _ptr(X), X.shape[0],
                                          _ptr(out)), "gz_model_forward")
        self.ctx.synchronize()
        return out

    def grad(self, X: torch.Tensor, T: torch.Tensor):
        X = X.to(torch.float32).contiguous()
        T = T.to(torch.float32).contiguous()
        g = torch.empty(self.N_PARAMS, dtype=torch.float32, device=X.device)
        loss = C.c_float(0.0)
        self.ctx.enter()
        L.check(L.load().gz_model_grad(self.ctx.handle, self.handle, _ptr(X), _ptr(T), X.shape[0],
                                       _ptr(g), C.byref(loss)), "gz_model_grad")
        return float(loss.value), g

    def train_step(self, X: torch.Tensor, T: torch.Tensor, lr: float) -> float:
        loss = C.c_float(0.0)
        self.ctx.enter()
        L.check(L.load().gz_model_train_step(self.ctx.handle, self.handle, _ptr(X), _ptr(T),
                                             X.shape[0], float(lr), C.byref(loss)),
                "gz_model_train_step")
        return float(loss.value)


def group_features(T: Tree, P: Particles, targets, aold, group_size: int = 0):
    """The groups the walk forms for these targets and their 16 features (device tensors)."""
    tg = _dev_u32(targets, P.device)
    nt = int(tg.shape[0])
    ao = _dev_f64(aold, P.device)
    feats = torch.empty((max(nt, 1), L.MLP_FEATURES), dtype=torch.float32, device=P.device)
    gstart = torch.empty(nt + 1, dtype=torch.int32, device=P.device)
    ng = C.c_size_t(0)
    T.ctx.enter()
    L.check(L.load().gz_tree_group_features(T.ctx.handle, T.handle, P.cstruct(), _ptr(tg), nt,
                                            int(group_size), _ptr(ao), _ptr(feats), _ptr(gstart),
                                            C.byref(ng)), "gz_tree_group_features")
    n = ng.value
    return feats[:n], gstart[: n + 1]


def accel_grouped(T: Tree, P: Particles, targets, theta: float, G: float = 1.0, batch: int = 0,
                  aold=None, *, precision: str = "fp64", group_size: int = 0,
                  persistent: bool = True, counts: bool = False, timing: bool = False,
                  stack_limit: int = 0, model: Model | None = None, out=None,
                  waves_per_cu: int = 0):
    """shmem::accel_grouped (tree.h:428-440) without tidal/jerk/lazy/sink terms.

    precision "fp64" is the parity walk; "fp32" evaluates interactions in FP32 on positions
    relative to the group centre (opening decisions and the per-target sum stay FP64).
    group_size > 0 groups targets under the highest node holding at most that many particles.
    persistent=True (the default) walks with a persistent grid fed by a work queue; False launches
    one workgroup per group.
    Returns (ax, ay, az) device tensors, plus a dict of walk counts when counts=True (node
    visits and interactions, SHMEM_COUNT_WALK's definition) or timing=True (walk and prep ms only,
    from the kernel variant without counters).
    """
    tg = _dev_u32(targets, P.device)
    nt = int(tg.shape[0])
    ax, ay, az = out if out is not None else _outputs(nt, P.device)
    ao = None if aold is None else _dev_f64(aold, P.device)
    prec = {"fp64": L.PREC_FP64, "fp32": L.PREC_FP32_REL}[precision]
    crit = L.CRIT_REFERENCE if model is None else L.CRIT_LEARNED
    launch = L.LAUNCH_QUEUE if persistent else L.LAUNCH_PER_GROUP
    opts = L.gz_walk_opts(prec, crit, int(group_size), launch, int(counts),
                          int(stack_limit), None if model is None else model.handle,
                          int(waves_per_cu))
    cnt = L.gz_walk_counts()
    want = counts or timing
    T.ctx.enter()
    L.check(L.load().gz_tree_accel_grouped(
        T.ctx.handle, T.handle, P.cstruct(), _ptr(tg), nt, float(theta), float(G), int(batch),
        _ptr(ao), _ptr(ax), _ptr(ay), _ptr(az), C.byref(opts),
        C.byref(cnt) if want else None), "gz_tree_accel_grouped")
    if want:
        c = {k: getattr(cnt, k) for k, _ in L.gz_walk_counts._fields_}
        return ax, ay, az, c
    return ax, ay, az


def autotune(T: Tree, P: Particles, targets, theta: float, G: float = 1.0, aold=None, *,
             precision: str = "fp32", model: Model | None = None, reps: int = 2,
             retune: bool = False):
    """gz_walk_autotune: the fastest launch parameters for this workload, timed on the device
    (group size, then node-stack depth and persistent launch width) and cached per workload in
    the context. Returns (choice, report): choice holds group_size, stack_limit and waves_per_cu,
    the keyword arguments of accel_grouped and WalkPlan; report lists the timed candidates."""
    tg = _dev_u32(targets, P.device)
    nt = int(tg.shape[0])
    ao = None if aold is None else _dev_f64(aold, P.device)
    prec = {"fp64": L.PREC_FP64, "fp32": L.PREC_FP32_REL}[precision]
    crit = L.CRIT_REFERENCE if model is None else L.CRIT_LEARNED
    base = L.gz_walk_opts(prec, crit, 0, L.LAUNCH_QUEUE, 0, 0,
                          None if model is None else model.handle, 0)
    best = L.gz_walk_opts()
    rep = L.gz_tune_report()
    T.ctx.enter()
    L.check(L.load().gz_walk_autotune(T.ctx.handle, T.handle, P.cstruct(), _ptr(tg), nt,
                                      float(theta), float(G), _ptr(ao), C.byref(base), int(reps),
                                      L.TUNE_RETUNE if retune else 0, C.byref(best),
                                      C.byref(rep)), "gz_walk_autotune")
    choice = {"group_size": best.group_size, "stack_limit": best.stack_limit,
              "waves_per_cu": best.waves_per_cu}
    pts = [{"group_size": rep.points[k].group_size, "stack_limit": rep.points[k].stack_limit,
            "waves_per_cu": rep.points[k].waves_per_cu, "ms": rep.points[k].ms}
           for k in range(rep.n_points)]
    return choice, {"from_cache": bool(rep.from_cache), "points": pts, "best_ms": rep.best.ms}


class WalkPlan:
    """Groups of a fixed target list formed once (gz_walk_plan); launch() only enqueues the walk,
    so launches can be captured into a graph. segments: slot offsets (length n_seg + 1) of
    independent target lists walked together, e.g. the active sets of several substeps."""

    def __init__(self, T: Tree, P: Particles, targets, *, segments=None, precision="fp32",
                 group_size=32, model: Model | None = None, stack_limit: int = 0,
                 waves_per_cu: int = 0):
        self.T, self.P = T, P
        self.targets = _dev_u32(targets, P.device)
        nt = int(self.targets.shape[0])
        prec = {"fp64": L.PREC_FP64, "fp32": L.PREC_FP32_REL}[precision]
        crit = L.CRIT_REFERENCE if model is None else L.CRIT_LEARNED
        opts = L.gz_walk_opts(prec, crit, int(group_size), L.LAUNCH_QUEUE, 0, int(stack_limit),
                              None if model is None else model.handle, int(waves_per_cu))
        self.model = model
        seg = None
        nseg = 0
        if segments is not None:
            arr = np.ascontiguousarray(segments, dtype=np.uint64)
            seg = (C.c_size_t * len(arr))(*[int(v) for v in arr])
            nseg = len(arr) - 1
        h = C.c_void_p()
        T.ctx.enter()
        L.check(L.load().gz_walk_plan_create(T.ctx.handle, T.handle, P.cstruct(),
                                             _ptr(self.targets), nt, seg, nseg, C.byref(opts),
                                             C.byref(h)), "gz_walk_plan_create")
        self.handle = h
        self.nt = nt
        self.ax, self.ay, self.az = _outputs(nt, P.device)

    def __del__(self):
        h = getattr(self, "handle", None)
        lib = getattr(L, "_lib", None) if L is not None else None
        if h is not None and lib is not None:
            lib.gz_walk_plan_free(h)
            self.handle = None

    def launch(self, theta, G=1.0, aold=None):
        """Enqueue the walk on the library stream (results in self.ax/ay/az once it finishes)."""
        self._aold = None if aold is None else _dev_f64(aold, self.P.device)
        L.check(L.load().gz_walk_plan_launch(self.T.ctx.handle, self.handle, float(theta),
                                             float(G), _ptr(self._aold), _ptr(self.ax),
                                             _ptr(self.ay), _ptr(self.az)),
                "gz_walk_plan_launch")

    def status(self) -> dict:
        cnt = L.gz_walk_counts()
        L.check(L.load().gz_walk_plan_status(self.T.ctx.handle, self.handle, C.byref(cnt)),
                "gz_walk_plan_status")
        return {k: getattr(cnt, k) for k, _ in L.gz_walk_counts._fields_}


class Graph:
    """Work enqueued on the library stream between begin() and end(), replayable (a hipGraph)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.handle = None

    def begin(self):
        self.ctx.enter()
        L.check(L.load().gz_graph_begin(self.ctx.handle), "gz_graph_begin")

    def end(self):
        h = C.c_void_p()
        L.check(L.load().gz_graph_end(self.ctx.handle, C.byref(h)), "gz_graph_end")
        self.handle = h

    def launch(self):
        L.check(L.load().gz_graph_launch(self.ctx.handle, self.handle), "gz_graph_launch")

    def __del__(self):
        h = getattr(self, "handle", None)
        lib = getattr(L, "_lib", None) if L is not None else None
        if h is not None and lib is not None:
            lib.gz_graph_free(h)
            self.handle = None


def accel_grouped_tidal(T: Tree, P: Particles, targets, theta: float, G: float = 1.0,
                        batch: int = 0, aold=None, vel=None, jerk: bool = False):
    """shmem::accel_grouped with the tidal tensor (and the jerk when vel is given and the tree
    was built with velocities), FP64 (tree.h:414-440). Returns (ax, ay, az, tidal (6, nt),
    jerk (3, nt) or None); the tidal tensor excludes G, as the reference's does."""
    tg = _dev_u32(targets, P.device)
    nt = int(tg.shape[0])
    ax, ay, az = _outputs(nt, P.device)
    tid = torch.empty((6, nt), dtype=torch.float64, device=P.device)
    jk = torch.empty((3, nt), dtype=torch.float64, device=P.device) if jerk else None
    ao = None if aold is None else _dev_f64(aold, P.device)
    v = (None, None, None) if vel is None else tuple(_dev_f64(a, P.device) for a in vel)
    T.ctx.enter()
    L.check(L.load().gz_tree_accel_grouped_tidal(
        T.ctx.handle, T.handle, P.cstruct(), _ptr(tg), nt, float(theta), float(G), int(batch),
        _ptr(ao), *(_ptr(a) for a in v), _ptr(ax), _ptr(ay), _ptr(az), _ptr(tid), _ptr(jk)),
        "gz_tree_accel_grouped_tidal")
    T.ctx.synchronize()
    return ax, ay, az, tid, jk


def density(T: Tree, P: Particles, targets, des_ngb: float = 32.0, ngb_tol: float = 1.0,
            h_start=None, n_dims: int = 3, type=None, cond_prev=None, tau: float = 0.0):
    """shmem::density (hydro.cc:150-327) for the target particles: (h, rho, nngb, iters) device
    tensors. type: GIZMO types by particle (0 = gas; None: all gas); cond_prev: condition numbers
    by particle; h_start: first guesses by target (None: the reference's own)."""
    dev = P.device
    tg = _dev_u32(targets, dev)
    nt = int(tg.shape[0])
    h = torch.empty(max(nt, 1), dtype=torch.float64, device=dev)
    rho = torch.empty_like(h)
    nngb = torch.empty(max(nt, 1), dtype=torch.int32, device=dev)
    iters = torch.empty_like(nngb)
    hs = None if h_start is None else _dev_f64(h_start, dev)
    cp = None if cond_prev is None else _dev_f64(cond_prev, dev)
    ty = None
    if type is not None:
        ty = torch.as_tensor(np.ascontiguousarray(type, dtype=np.uint8), device=dev)
    pc = L.gz_particles(P.n, P.x.data_ptr(), P.y.data_ptr(), P.z.data_ptr(), P.m.data_ptr(),
                        None if P.soft is None else P.soft.data_ptr(), None,
                        None if ty is None else ty.data_ptr())
    T.ctx.enter()
    L.check(L.load().gz_tree_density(T.ctx.handle, T.handle, C.byref(pc), _ptr(tg), nt,
                                     float(des_ngb), float(ngb_tol),
                                     None if hs is None else hs.data_ptr(), int(n_dims),
                                     None if cp is None else cp.data_ptr(), float(tau),
                                     h.data_ptr(), rho.data_ptr(), nngb.data_ptr(),
                                     iters.data_ptr()), "gz_tree_density")
    T.ctx.synchronize()
    return h[:nt], rho[:nt], nngb[:nt], iters[:nt]


def _ngb_lists(T: Tree, call, who: str, nq: int):
    """The lists of a search whose C call ends in (found, cap, &total): first into a buffer sized
    from the tree's last search (its mean list length and an eighth more, at least 32 entries per
    query, at most 2^28 entries in all), so that one pass of counting usually serves; a second
    call only when the lists need more."""
    dev = T.P.device
    total = C.c_size_t(0)
    per_query = max(32.0, 1.125 * getattr(T, "_ngb_per_query", 0.0))
    cap = max(1, min(int(per_query * nq) + 1, 1 << 28))
    found = torch.empty(cap, dtype=torch.int32, device=dev)
    L.check(call(found.data_ptr(), cap, C.byref(total)), who)
    if total.value > cap:
        cap = int(total.value)
        found = torch.empty(cap, dtype=torch.int32, device=dev)
        L.check(call(found.data_ptr(), cap, C.byref(total)), who)
    T._ngb_per_query = int(total.value) / max(nq, 1)
    T.ctx.synchronize()
    return found[:int(total.value)].view(torch.uint32)


def ngb_search(T: Tree, centres, radius, tau: float = 0.0):
    """shmem::ngb_search (hydro.h:73-75) for many centres: (start, found) device tensors, CSR —
    the particles within radius[q] of centres[q] (n_q x 3), each list in Morton order."""
    dev = T.P.device
    c = centres if isinstance(centres, torch.Tensor) else torch.as_tensor(np.asarray(centres))
    c = c.to(device=dev, dtype=torch.float64)
    cs = [c[:, d].contiguous() for d in range(3)]
    r = _dev_f64(radius, dev)
    nq = int(r.shape[0])
    start = torch.zeros(nq + 1, dtype=torch.int64, device=dev)
    lib = L.load()
    T.ctx.enter()
    args = (T.ctx.handle, T.handle, *(a.data_ptr() for a in cs), r.data_ptr(), nq, float(tau),
            start.data_ptr())
    found = _ngb_lists(T, lambda *out: lib.gz_tree_ngb_search(*args, *out),
                       "gz_tree_ngb_search", nq)
    return start, found


def ngb_search_reverse(T: Tree, h, n_gas: int, centres, h_q, tau: float = 0.0):
    """shmem::ngb_search_reverse (hydro.h:78-83) for many centres: (start, found) device tensors,
    CSR — the gas particles j with h_q[q] <= r < h[j] (their own kernel reaches centre q), each
    list in Morton order. Needs Tree.init_hmax."""
    dev = T.P.device
    c = centres if isinstance(centres, torch.Tensor) else torch.as_tensor(np.asarray(centres))
    c = c.to(device=dev, dtype=torch.float64)
    cs = [c[:, d].contiguous() for d in range(3)]
    hh = _dev_f64(h, dev)
    if int(hh.shape[0]) < int(n_gas):
        raise ValueError(f"ngb_search_reverse: h holds {int(hh.shape[0])} values for "
                         f"n_gas={int(n_gas)} gas particles")
    r = _dev_f64(h_q, dev)
    nq = int(r.shape[0])
    start = torch.zeros(nq + 1, dtype=torch.int64, device=dev)
    lib = L.load()
    T.ctx.enter()
    args = (T.ctx.handle, T.handle, hh.data_ptr(), int(n_gas), *(a.data_ptr() for a in cs),
            r.data_ptr(), nq, float(tau), start.data_ptr())
    found = _ngb_lists(T, lambda *out: lib.gz_tree_ngb_search_reverse(*args, *out),
                       "gz_tree_ngb_search_reverse", nq)
    return start, found


def potential(T: Tree, P: Particles, targets, theta: float, G: float = 1.0, aold=None):
    """shmem::potential (tree.h:446-447): spline-softened potential per target (FP64)."""
    tg = _dev_u32(targets, P.device)
    nt = int(tg.shape[0])
    phi = torch.empty(nt, dtype=torch.float64, device=P.device)
    ao = None if aold is None else _dev_f64(aold, P.device)
    T.ctx.enter()
    L.check(L.load().gz_tree_potential(T.ctx.handle, T.handle, P.cstruct(), _ptr(tg), nt,
                                       float(theta), float(G), _ptr(ao), _ptr(phi)),
            "gz_tree_potential")
    T.ctx.synchronize()
    return phi


def accel(T: Tree, P: Particles, targets, theta: float, G: float = 1.0, aold=None):
    """shmem::accel (tree.h:392-394): per-target walk, reached leaves always summed."""
    tg = _dev_u32(targets, P.device)
    nt = int(tg.shape[0])
    ax, ay, az = _outputs(nt, P.device)
    ao = None if aold is None else _dev_f64(aold, P.device)
    T.ctx.enter()
    L.check(L.load().gz_tree_accel(T.ctx.handle, T.handle, P.cstruct(), _ptr(tg), nt,
                                   float(theta), float(G), _ptr(ao), _ptr(ax), _ptr(ay),
                                   _ptr(az)), "gz_tree_accel")
    T.ctx.synchronize()
    return ax, ay, az


def accel_brute(P: Particles, targets, G: float = 1.0, ctx: Context | None = None):
    """shmem::accel_brute (tree.h:450-451): exact FP64 direct sum, includes G."""
    ctx = ctx or Context.get(P.device.index or 0)
    tg = _dev_u32(targets, P.device)
    nt = int(tg.shape[0])
    ax, ay, az = _outputs(nt, P.device)
    ctx.enter()
    L.check(L.load().gz_direct_sum(ctx.handle, P.cstruct(), _ptr(tg), nt, float(G), _ptr(ax),
                                   _ptr(ay), _ptr(az)), "gz_direct_sum")
    ctx.synchronize()
    return ax, ay, az


def direct_sum(P: Particles, targets, G: float = 1.0, fp64: bool = False, box: float = 0.0,
               ctx: Context | None = None):
    """shmem_cuda_accel_bruteforce (grav_cuda.cu:65-247): the device direct sum in the reference's
    default float-narrowed arithmetic (fp64=False: separations differenced in FP64, the rest in
    FP32) or its SHMEM_CUDA_FP64 form; box > 0 = nearest periodic image. Includes G."""
    ctx = ctx or Context.get(P.device.index or 0)
    tg = _dev_u32(targets, P.device)
    nt = int(tg.shape[0])
    ax, ay, az = _outputs(nt, P.device)
    ctx.enter()
    L.check(L.load().gz_direct_sum_ex(ctx.handle, P.cstruct(), _ptr(tg), nt, float(G),
                                      1 if fp64 else 0, float(box), _ptr(ax), _ptr(ay),
                                      _ptr(az)), "gz_direct_sum_ex")
    ctx.synchronize()
    return ax, ay, az


def order_targets(T: Tree, active=None) -> torch.Tensor:
    """Active particles in the tree's Morton order (mfm.cc:698-718). None = all active."""
    n_out = C.c_size_t(0)
    out = torch.empty(max(T.n_live, 1), dtype=torch.int32, device=T.P.device)
    T.ctx.enter()
    if active is None:
        L.check(L.load().gz_order_targets(T.ctx.handle, T.handle, None, T.P.n, _ptr(out),
                                          C.byref(n_out)), "gz_order_targets")
    else:
        act = _dev_u32(active, T.P.device)
        L.check(L.load().gz_order_targets(T.ctx.handle, T.handle, _ptr(act), int(act.shape[0]),
                                          _ptr(out), C.byref(n_out)), "gz_order_targets")
    return out[: n_out.value].view(torch.uint32)


def aold_from_accel(T: Tree, a_prev: torch.Tensor, targets, err_tol_force_acc: float,
                    G: float = 1.0) -> torch.Tensor:
    """aold[k] = ErrTolForceAcc * |a_prev[targets[k]]| / G (mfm.cc:743-751); a_prev is (N, 3)."""
    tg = _dev_u32(targets, T.P.device)
    a_prev = a_prev.to(T.P.device, torch.float64).contiguous()
    out = torch.empty(int(tg.shape[0]), dtype=torch.float64, device=T.P.device)
    T.ctx.enter()
    L.check(L.load().gz_aold_from_accel(T.ctx.handle, _ptr(a_prev), _ptr(tg), int(tg.shape[0]),
                                        float(err_tol_force_acc), float(G), _ptr(out)),
            "gz_aold_from_accel")
    T.ctx.synchronize()
    return out


def scatter_accel(targets, ax, ay, az, a_grav: torch.Tensor, ctx: Context | None = None):
    """a_grav[targets[k]] = (ax[k], ay[k], az[k]) (mfm.cc:828-831); a_grav is an (N, 3) float64
    device tensor, written in place."""
    ctx = ctx or Context.get(a_grav.device.index or 0)
    tg = _dev_u32(targets, a_grav.device)
    if a_grav.dtype != torch.float64 or not a_grav.is_contiguous() or a_grav.dim() != 2:
        raise ValueError("scatter_accel: a_grav must be a contiguous (N, 3) float64 tensor")
    a = [_dev_f64(v, a_grav.device) for v in (ax, ay, az)]
    ctx.enter()
    L.check(L.load().gz_scatter_accel(ctx.handle, _ptr(tg), int(tg.shape[0]), *(_ptr(v) for v in a),
                                      _ptr(a_grav)), "gz_scatter_accel")
    ctx.synchronize()


def drift(P: Particles, v, dt: float, ctx: Context | None = None):
    """The drift (mfm.cc:1287-1295, gravity part): P's positions += v dt in place for every
    particle with m > 0. v: three float64 device tensors by particle."""
    ctx = ctx or Context.get(P.device.index or 0)
    for t in v:
        if t.dtype != torch.float64 or not t.is_contiguous() or t.device != P.device \
                or int(t.shape[0]) != P.n:
            raise ValueError("drift: v must be three contiguous float64 tensors, one per particle")
    ctx.enter()
    L.check(L.load().gz_drift(ctx.handle, P.n, float(dt), *(_ptr(t) for t in v), _ptr(P.m),
                              _ptr(P.x), _ptr(P.y), _ptr(P.z)), "gz_drift")
    ctx.synchronize()


def kick(active, a_grav: torch.Tensor, dt: torch.Tensor, pending: torch.Tensor, v,
         a_extra: torch.Tensor | None = None, ctx: Context | None = None):
    """The gravity part of the KDK kick (mfm.cc:3749-3767) for the particles in active: with
    i = active[k], v[i] += a (pending[i] + dt[i] / 2) and pending[i] = dt[i] / 2, a = a_grav[i]
    (+ a_extra[i]); a_grav and a_extra are (N, 3), dt, pending and the three velocity tensors are
    by particle. pending and v (float64 device tensors) are updated in place."""
    ctx = ctx or Context.get(a_grav.device.index or 0)
    act = _dev_u32(active, a_grav.device)
    for t in (a_grav, dt, pending, *v) + (() if a_extra is None else (a_extra,)):
        if t.dtype != torch.float64 or not t.is_contiguous() or t.device != a_grav.device:
            raise ValueError("kick: every array must be a contiguous float64 device tensor")
    n = int(a_grav.shape[0])
    if any(int(t.shape[0]) != n for t in (dt, pending, *v)):
        raise ValueError("kick: dt, pending and v hold one value per particle, as a_grav does")
    ctx.enter()
    L.check(L.load().gz_kick(ctx.handle, _ptr(act), int(act.shape[0]), _ptr(a_grav),
                             _ptr(a_extra), _ptr(dt), _ptr(pending), *(_ptr(t) for t in v)),
            "gz_kick")
    ctx.synchronize()
