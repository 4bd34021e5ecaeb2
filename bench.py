#!/usr/bin/env python3
"""Force-evaluation throughput of the device tree walk at 8M particles.

Workload: the 8,388,608-particle Hernquist halo (a=1, 70%) + exponential disk (R_d=0.1, z_d=0.01,
30%), equal masses, cubic-spline softening h=0.005, all particles active, G=1. One step is one full
force evaluation: the grouped tree walk (target grouping, group bounds and the walk kernel) over all
8M targets in Morton order, with the reference's default opening criterion: the relative criterion
(ErrTolForceAcc = 0.0025) with aold from a bootstrap Barnes-Hut walk at theta = 0.7. The tree is
built once before the timed region (build time is reported separately); inputs are resident in HBM.

Reported beside the headline:
  * every walk variant's evals/s, kernel time and interactions/target (SHMEM_COUNT_WALK's count);
  * force error of each walk against the FP64 direct sum on 65,536 random probe particles;
  * roofline of the walk kernel: 20 flops per interaction over the kernel's HIP-event time;
  * cpu_baseline: the reference shmem CPU tree walk (oracle/_ref) timed on this host's cores on a
    bounded sample of whole leaves of the same tree, or the C restatement when the reference build
    is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

# The cores this process may run on, read before anything loads an OpenMP runtime: with
# OMP_PROC_BIND set, libgomp binds the initial thread to one place when it starts, after which
# the affinity mask shows that single core.
_AFFINITY_CPUS = len(os.sched_getaffinity(0))

# The reference's run recipe (shmem/README.md:49-57) binds OpenMP threads one per core, spread
# over the sockets. OpenMP reads these once, when its runtime is first loaded, so they are set
# before anything that links it is imported; a value already in the environment is kept.
os.environ.setdefault("OMP_PROC_BIND", "spread")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("force evals/s at 8M part., learned+FP32 vs plain GPU walk, p99 rel err ≤1e-2")
FLOPS_PER_INTERACTION = 20.0
PEAK_FP32_TFLOPS = 157.3
ERR_TOL_FORCE_ACC = 0.0025
THETA_BOOT = 0.7


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def rel_err(g, o):
    g = np.stack(g, 1)
    o = np.stack(o, 1)
    return np.linalg.norm(g - o, axis=1) / np.maximum(np.linalg.norm(o, axis=1), 1e-300)


def err_stats(e):
    return {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)),
            "max": float(np.max(e))}


PROFILE = os.path.join(ROOT, "profiles", "k_walk_fp32_learned_halo_disk_8m.md")


def profiled_traffic(path, kernel_ms):
    """HBM bytes per walk launch from the committed rocprofv3 counter summary (FETCH_SIZE and
    WRITE_SIZE scaled by a plain copy's readings), when that summary's kernel time matches this
    run's within 15% (otherwise the profile describes another build and None is returned)."""
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:") or line.startswith("kernel_avg_ms:"):
                k, v = line.split(":", 1)
                vals[k.strip()] = float(v.strip())
    if "traffic_bytes_per_launch" not in vals or "kernel_avg_ms" not in vals:
        return None
    if abs(vals["kernel_avg_ms"] - kernel_ms) > 0.15 * kernel_ms:
        return None
    return vals["traffic_bytes_per_launch"]


def time_walk(gz, T, P, targets, aold, steps, warmup, **kw):
    """Mean wall time per step (host clock around the calls, synchronised) and the walk kernel's
    mean HIP-event time over the same steps."""
    import torch
    out = None
    for _ in range(warmup):
        out = gz.accel_grouped(T, P, targets, 0.0, 1.0, 0, aold=aold, timing=True, **kw)
    torch.cuda.synchronize()
    walk_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        out = gz.accel_grouped(T, P, targets, 0.0, 1.0, 0, aold=aold, timing=True, **kw)
        walk_ms.append(out[3]["walk_ms"])
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return out, dt, float(np.mean(walk_ms))


def host_cores():
    """Threads for the CPU baseline and a description of the host. The count is the cores this
    process may run on (its affinity mask), capped by OMP_NUM_THREADS when that is set: a GPU
    box grants each GPU a share of its cores and says how many there."""
    avail = _AFFINITY_CPUS
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(avail, int(env)) if env.isdigit() and int(env) > 0 else avail
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    info = {"threads": threads, "affinity_cpus": avail, "nproc_online": os.cpu_count(),
            "omp_num_threads_env": env or None, "cpu_model": model,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES")}
    bi = os.path.join(ROOT, "oracle", "_ref", "build_info.txt")
    if os.path.exists(bi):
        with open(bi) as f:
            info["reference_build"] = f.read().strip()
    return threads, info


def cpu_baseline(x, y, z, m, soft, orderbuf, leaf_of, aold_all, seconds, threads, reps=3):
    """The reference walk (accel_grouped, batch 0, relative criterion) on whole leaves sampled
    uniformly along the Morton order, timed on `threads` host cores: one warm-up run, then the
    minimum of `reps` timed runs (shmem/bench_layout.cc:7-11)."""
    from oracle.bindings import HostParticles, Oracle, Reference, have_reference
    hp = HostParticles(x, y, z, m, soft)
    leaf_seq = leaf_of[orderbuf]
    run_start = np.concatenate([[True], leaf_seq[1:] != leaf_seq[:-1]])
    run_id = np.cumsum(run_start) - 1
    n_runs = int(run_id[-1]) + 1

    def sample(stride):
        keep = (run_id % stride) == 0
        return orderbuf[keep].astype(np.uint32), aold_all[keep]

    if have_reference():
        ref = Reference()
        ref.set_threads(threads)
        rp = ref.particles(hp)
        t0 = time.perf_counter()
        tr = ref.build(rp)
        build_s = time.perf_counter() - t0

        def walk(tg, ao):
            return ref.accel_grouped(tr, rp, tg, 0.0, 1.0, 0, ao)
        kind = "reference"
        cleanup = lambda: (ref.free(tr), ref.free_particles(rp))  # noqa: E731
    else:
        orc = Oracle()
        t0 = time.perf_counter()
        to = orc.build(hp)
        build_s = time.perf_counter() - t0

        def walk(tg, ao):
            return orc.accel_grouped(to, hp, tg, 0.0, 1.0, 0, ao)
        kind = "port"
        cleanup = lambda: orc.free(to)  # noqa: E731
    try:
        stride = 512
        tg, ao = sample(stride)
        t0 = time.perf_counter()
        walk(tg, ao)
        per = (time.perf_counter() - t0) / max(len(tg), 1)
        want = seconds / max(per, 1e-12)
        stride = int(max(1, min(n_runs, round(len(orderbuf) / max(want, 1.0)))))
        tg, ao = sample(stride)
        walk(tg, ao)  # warm-up
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            walk(tg, ao)
            times.append(time.perf_counter() - t0)
        dt = min(times)
    finally:
        cleanup()
    return {"value": len(tg) / dt, "unit": "force evals/s", "cores": threads, "kind": kind,
            "sample": f"{len(tg)} targets = every {stride}th leaf of the {n_runs} leaves in Morton "
                      f"order, relative criterion ErrTolForceAcc={ERR_TOL_FORCE_ACC}, aold from the "
                      f"theta={THETA_BOOT} bootstrap; min of {reps} timed walks after a warm-up: "
                      f"{', '.join(f'{t:.2f}' for t in times)} s; tree build {build_s:.2f} s",
            "build_s": build_s, "walk_s": times}


def widened_rows(gz, ics, args):
    """The path's other calls on the Hernquist 4M (a=1, eps=0.005), each beside the reference's
    own CPU code (oracle/_ref) on a sample: the potential, per-target acceleration and
    tidal-tensor walks (FP64, theta=0.7, every particle a target; the reference on every 512th
    particle, all host threads), and the forward and reverse neighbour searches (every 4th particle in Morton order
    as a query, radius and kernel radius = the side of the particle's leaf; the reference's
    per-centre calls, one thread, on every 64th query), and the density / smoothing-length solve
    (every particle a gas target, DesNumNgb 32, tolerance 1, first guesses from the mean density
    of the particle's leaf; the reference's density() on all host threads on every 64th target)."""
    import torch
    from oracle.bindings import HostParticles, Reference, have_reference
    x, y, z, m, soft = ics.hernquist()
    n = x.shape[0]
    P = gz.Particles(x, y, z, m, soft, device="cuda:0")
    T = gz.build(P)
    tg = T.device_orderbuf()
    td = T.download()
    out = {"workload": "hernquist N=4194304 eps=0.005"}

    def timed(fn):
        r = fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = fn()
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t0) / args.steps

    _, dt = timed(lambda: gz.potential(T, P, tg, THETA_BOOT))
    out["potential_fp64"] = {"ms_per_call": dt * 1e3, "evals_per_s": n / dt}
    _, dt = timed(lambda: gz.accel(T, P, tg, THETA_BOOT))
    out["accel_per_target_fp64"] = {"ms_per_call": dt * 1e3, "evals_per_s": n / dt}
    _, dt = timed(lambda: gz.accel_grouped_tidal(T, P, tg, THETA_BOOT))
    out["tidal_fp64"] = {"ms_per_call": dt * 1e3, "evals_per_s": n / dt}
    order = td["orderbuf"].astype(np.int64)
    side = td["size"][td["leaf_of"]]  # per particle: its leaf's side
    q = order[::4]
    c = np.stack([x[q], y[q], z[q]], 1)
    r = side[q]
    # the searches' inputs resident in HBM, as the walks' are
    c_d = torch.as_tensor(c, device=P.device)
    r_d = torch.as_tensor(r, device=P.device)
    side_d = torch.as_tensor(side, device=P.device)
    (s, _), dt = timed(lambda: gz.ngb_search(T, c_d, r_d))
    nq = q.shape[0]
    out["ngb_search"] = {"queries": nq, "per_query": int(s[-1]) / nq, "ms_per_call": dt * 1e3,
                         "queries_per_s": nq / dt}
    T.init_hmax(side_d, n)
    (s, _), dt = timed(lambda: gz.ngb_search_reverse(T, side_d, n, c_d, r_d))
    out["ngb_search_reverse"] = {"queries": nq, "per_query": int(s[-1]) / nq,
                                 "ms_per_call": dt * 1e3, "queries_per_s": nq / dt}
    # first guesses: DesNumNgb particles at the mean density of the particle's leaf
    leaf_n = (td["phi"] - td["plo"])[td["leaf_of"]].astype(np.float64)
    h_leaf = side * np.cbrt(3.0 * 32.0 / (4.0 * np.pi * leaf_n))
    h0 = h_leaf[order]
    h0_d = torch.as_tensor(h0, device=P.device)
    (_, rho_d, nngb_d, it_d), dt = timed(lambda: gz.density(T, P, tg, 32.0, 1.0, h_start=h0_d))
    it_h = it_d.cpu().numpy()
    out["density"] = {"targets": n, "ms_per_call": dt * 1e3, "targets_per_s": n / dt,
                      "mean_iterations": float(it_h.mean()),
                      "at_iteration_cap": int((it_h == 100).sum()),
                      "mean_neighbours_inside_h": float(nngb_d.float().mean())}
    if not args.no_cpu and have_reference():
        ref = Reference()
        threads = args.cpu_threads or host_cores()[0]
        ref.set_threads(threads)
        rp = ref.particles(HostParticles(x, y, z, m, soft))
        tr = ref.build(rp)
        try:
            smp = order[::512].astype(np.uint32)
            t0 = time.perf_counter()
            ref.potential(tr, rp, smp, THETA_BOOT)
            out["potential_fp64"]["cpu_reference_evals_per_s"] = smp.shape[0] / (
                time.perf_counter() - t0)
            t0 = time.perf_counter()
            ref.accel(tr, rp, smp, THETA_BOOT)
            out["accel_per_target_fp64"]["cpu_reference_evals_per_s"] = smp.shape[0] / (
                time.perf_counter() - t0)
            t0 = time.perf_counter()
            ref.accel_grouped_tj(tr, rp, smp, THETA_BOOT)
            out["tidal_fp64"]["cpu_reference_evals_per_s"] = smp.shape[0] / (
                time.perf_counter() - t0)
            cs, rs = c[::64], r[::64]
            t0 = time.perf_counter()
            ref.ngb_search(tr, rp, cs, rs)
            out["ngb_search"]["cpu_reference_queries_per_s_1thread"] = cs.shape[0] / (
                time.perf_counter() - t0)
            ref.init_hmax(tr, side, n)
            t0 = time.perf_counter()
            ref.ngb_search_reverse(tr, rp, side, n, cs, rs)
            out["ngb_search_reverse"]["cpu_reference_queries_per_s_1thread"] = cs.shape[0] / (
                time.perf_counter() - t0)
            ds = order[::64].astype(np.uint32)
            t0 = time.perf_counter()
            ref.density(tr, rp, ds, 32.0, 1.0, h_start=h_leaf[ds])
            out["density"]["cpu_reference_targets_per_s"] = ds.shape[0] / (
                time.perf_counter() - t0)
            out["cpu_threads_walks"] = threads
        finally:
            ref.free(tr)
            ref.free_particles(rp)
    log(f"rows: {json.dumps(out)}")
    del T, P, tg
    torch.cuda.empty_cache()


def other_configs(gz, ics, args):
    """The other named workloads, timed like the headline walk: the FP64 parity walk (per-leaf
    groups) and the FP32 walk (group size tuned), both with the relative criterion and aold from
    the theta=0.7 bootstrap, and their errors against the FP64 direct sum on up to 65,536 probe
    particles; for the Plummer case (the reference's own CPU-runnable size) also the reference
    CPU walk on every particle."""
    import torch
    for name, make in (("plummer N=65536 eps=0.01", ics.plummer),
                       ("hernquist N=4194304 eps=0.005", ics.hernquist),
                       ("clustered N=16777216 eps=0.001", ics.clustered)):
        x, y, z, m, soft = make()
        n = x.shape[0]
        P = gz.Particles(x, y, z, m, soft, device="cuda:0")
        T = gz.build(P)
        T = gz.build(P)
        tg = T.device_orderbuf()
        boot = gz.accel_grouped(T, P, tg, THETA_BOOT, 1.0, 0)
        aold = ERR_TOL_FORCE_ACC * torch.linalg.norm(torch.stack(boot, 1), dim=1)
        del boot
        probes = ics.probes(n, min(args.probes, n))
        pt = tg.view(torch.int32)[torch.as_tensor(probes.astype(np.int64), device="cuda:0")]
        ex = [v.cpu().numpy() for v in gz.accel_brute(P, pt.view(torch.uint32), 1.0)]
        out = {"workload": name, "build_ms": T.times()["total"]}
        # the FP32 walk's launch parameters from the library's autotuner (small systems have
        # too few 64-target groups to fill the GPU)
        ch32, _ = gz.autotune(T, P, tg, 0.0, 1.0, aold, precision="fp32")
        ch64 = {"group_size": 0, "stack_limit": 0, "waves_per_cu": 0}
        for label, prec, ch in (("fp64", "fp64", ch64), ("fp32", "fp32", ch32)):
            (ax, ay, az, _), dt, kms = time_walk(gz, T, P, tg, aold, args.steps, args.warmup,
                                                 precision=prec, **ch)
            e = rel_err(tuple(v.cpu().numpy()[probes] for v in (ax, ay, az)), ex)
            out[label] = {**ch, "ms_per_step": dt * 1e3, "kernel_ms": kms,
                          "evals_per_s": n / dt, "err": err_stats(e)}
            del ax, ay, az
        if n <= 65536 and not args.no_cpu:
            td = T.download()
            cpu = cpu_baseline(x, y, z, m, soft, td["orderbuf"], td["leaf_of"],
                               aold.cpu().numpy(), args.cpu_seconds,
                               args.cpu_threads or host_cores()[0])
            out["cpu_baseline"] = {k: cpu[k] for k in ("value", "unit", "cores", "kind",
                                                       "sample")}
            out["fp64_over_cpu"] = out["fp64"]["evals_per_s"] / cpu["value"]
        log(f"config: {json.dumps(out)}")
        del T, P, tg, aold
        torch.cuda.empty_cache()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=8388608)
    ap.add_argument("--probes", type=int, default=65536)
    ap.add_argument("--cpu-seconds", type=float, default=6.0,
                    help="CPU work per timed baseline walk (one warm-up and three timed walks)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--group-size", type=int, default=-1,
                    help="target-group size of the FP32 walks; -1 (default): the library's "
                         "autotuner picks group size, node-stack depth and launch width")
    ap.add_argument("--variants", default="fp64,fp32,learned,global",
                    help="walks to measure (the headline is the learned walk when measured, else "
                         "the FP32 walk); a profiler run can keep to one walk kernel")
    ap.add_argument("--model", default=os.path.join(ROOT, "gizmo_shared_memory_amd",
                                                     "learned_criterion.json"))
    ap.add_argument("--rows", action="store_true",
                    help="also time the path's other calls (potential, tidal tensor, forward and "
                         "reverse neighbour search) on the Hernquist 4M beside the reference's "
                         "CPU code, logged to stderr; the stdout line is unchanged")
    ap.add_argument("--rows-only", action="store_true",
                    help="time only the path's other calls (as --rows) and print nothing on "
                         "stdout: a short run for profiling those kernels")
    ap.add_argument("--configs", action="store_true",
                    help="also time the other named workloads (Plummer 64k, Hernquist 4M, "
                         "clustered 16M: build, FP64 and FP32 walks, errors against the direct "
                         "sum; the reference CPU walk beside the Plummer case) and log one JSON "
                         "object per workload to stderr; the stdout line is unchanged")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after the walks, run one plain device copy of this many MiB (a known "
                         "traffic reference for hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the walk shards "
              f"by target groups; no multi-GPU figure is produced)", file=sys.stderr)
        sys.exit(2)

    import torch

    from gizmo_shared_memory_amd import ics
    from gizmo_shared_memory_amd import tree as gz

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.rows_only:
        widened_rows(gz, ics, args)
        return
    t0 = time.perf_counter()
    x, y, z, m, soft = ics.halo_disk(n=args.n)
    log(f"generated halo+disk n={args.n} in {time.perf_counter() - t0:.1f} s")
    P = gz.Particles(x, y, z, m, soft, device="cuda:0")
    T = gz.build(P)
    T = gz.build(P)  # second build is the timed one (first touches the allocator)
    build_ms = T.times()["total"]
    log(f"tree: nodes={T.nnodes} leaves={T.info.nleaves} depth={T.info.depth} "
        f"build={build_ms:.1f} ms {T.times()}")
    targets = T.device_orderbuf()
    nt = int(targets.shape[0])

    # bootstrap walk (reference: theta 0.7, no aold), then aold = ErrTolForceAcc |a| / G
    boot = gz.accel_grouped(T, P, targets, THETA_BOOT, 1.0, 0, counts=True)
    a_boot = torch.stack(boot[:3], 1)
    aold = ERR_TOL_FORCE_ACC * torch.linalg.norm(a_boot, dim=1)
    log(f"bootstrap fp64 walk: {boot[3]}")

    # probes and the FP64 direct-sum judge
    probes = ics.probes(nt, args.probes)
    probe_t = targets.view(torch.int32)[torch.as_tensor(probes.astype(np.int64),
                                                        device="cuda:0")].view(torch.uint32)
    t0 = time.perf_counter()
    ex = gz.accel_brute(P, probe_t, 1.0)
    ex = [v.cpu().numpy() for v in ex]
    log(f"direct sum on {len(probes)} probes: {time.perf_counter() - t0:.2f} s")

    def measure(label, prec, choice, steps, warmup, scale=1.0, model=None):
        ao = aold if scale == 1.0 else aold * scale
        (ax, ay, az, c), dt, kms = time_walk(gz, T, P, targets, ao, steps, warmup, precision=prec,
                                             model=model, **choice)
        cnt = gz.accel_grouped(T, P, targets, 0.0, 1.0, 0, aold=ao, precision=prec,
                               group_size=choice["group_size"], counts=True, model=model)[3]
        inter = cnt["multipole"] + cnt["pairs"]
        e = rel_err(tuple(v.cpu().numpy()[probes] for v in (ax, ay, az)), ex)
        r = {**choice, "ms_per_step": dt * 1e3, "kernel_ms": kms,
             "evals_per_s": nt / dt, "interactions_per_target": inter / nt,
             "multipole_per_target": cnt["multipole"] / nt, "pairs_per_target": cnt["pairs"] / nt,
             "nodes_visited_per_target": cnt["nodes_visited"] / nt, "groups": cnt["groups"],
             "err": err_stats(e),
             "achieved_tflops": FLOPS_PER_INTERACTION * inter / (kms * 1e-3) / 1e12}
        if model is not None:
            r["fallback_groups"] = cnt["fallback_groups"]
        if scale != 1.0:
            r["err_tol_force_acc"] = ERR_TOL_FORCE_ACC * scale
        log(f"{label}: {r}")
        return r

    results = {}
    variants = set(args.variants.split(","))
    fixed_gs = int(args.group_size)

    def tune(label, prec="fp32", model=None):
        """The launch parameters for this walk: the library's autotuner (gz_walk_autotune:
        group size, then node-stack depth and launch width, each timed on this workload), or
        the one group size given with --group-size."""
        if fixed_gs >= 0:
            return {"group_size": fixed_gs, "stack_limit": 0, "waves_per_cu": 0}
        choice, rep = gz.autotune(T, P, targets, 0.0, 1.0, aold, precision=prec, model=model)
        log(f"autotune {label}: {choice} from {rep['points']}")
        return choice

    # plain GPU walks with the reference criterion: the parity walk with the reference's own
    # grouping (one group per leaf), and the FP32 walk
    if "fp64" in variants:
        results["fp64"] = measure("fp64", "fp64",
                                  {"group_size": 0, "stack_limit": 0, "waves_per_cu": 0},
                                  args.steps, args.warmup)
    ch32 = None
    if "fp32" in variants or "global" in variants:
        ch32 = tune("fp32")
    if "fp32" in variants:
        results["fp32"] = measure("fp32", "fp32", ch32, args.steps, args.warmup)

    # learned criterion (FP32 walk), its group size tuned on its own
    model = None
    if "learned" not in variants:
        pass
    elif os.path.exists(args.model):
        from gizmo_shared_memory_amd import learned
        model = learned.load_model(args.model)
        chl = tune("learned", model=model)
        results["learned_fp32"] = measure("learned_fp32", "fp32", chl, args.steps, args.warmup,
                                          model=model)
    else:
        log(f"no learned model at {args.model}")
    if "global" in variants and "learned_fp32" in results:
        # the method without the network: one global tolerance for every group, the largest
        # on a grid whose p99 error stays within the learned walk's
        goal = max(results["learned_fp32"]["err"]["p99"], 1e-12)
        chg = {k: results["learned_fp32"][k] for k in ("group_size", "stack_limit",
                                                        "waves_per_cu")}
        glob = None
        for k in range(0, 11):
            sc = 2.0 ** (k / 2.0)
            r = measure(f"global_tol_x{sc:.2f}", "fp32", chg, 1, 1, scale=sc)
            if r["err"]["p99"] <= goal:
                glob = sc
            else:
                break
        if glob is not None:  # the chosen tolerance, timed like the other walks
            glob = measure(f"global_tol_x{glob:.2f}", "fp32", chg, args.steps, args.warmup,
                           scale=glob)
        results["global_tolerance_fp32"] = glob

    if args.configs:
        other_configs(gz, ics, args)
    if args.rows:
        widened_rows(gz, ics, args)

    if args.copy_mb > 0:
        src = torch.ones(args.copy_mb << 18, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_mb} MiB read + {args.copy_mb} MiB written")
        del src, dst

    head = results.get("learned_fp32") or results.get("fp32") or results["fp64"]
    learned_ok = "learned_fp32" in results and head["err"]["p99"] <= 1e-2
    ref_e = rel_err(tuple(v.cpu().numpy()[probes] for v in boot[:3]), ex)
    cpu = None
    host = None
    if not args.no_cpu:
        threads, host = host_cores()
        threads = args.cpu_threads or threads
        td = T.download()
        # targets are orderbuf, so aold is already indexed by Morton slot
        cpu = cpu_baseline(x, y, z, m, soft, td["orderbuf"], td["leaf_of"], aold.cpu().numpy(),
                           args.cpu_seconds, threads)
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": METRIC,
        "value": head["evals_per_s"],
        "unit": "force evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"halo+disk N={args.n} (Hernquist a=1 70% + exp disk Rd=0.1 zd=0.01 30%),"
                        f" h=0.005, all active",
            "walk": ("FP32 group-relative, learned per-group tolerance with reference "
                     "fallback" if "learned_fp32" in results else
                     f"FP32 group-relative, reference criterion (relative {ERR_TOL_FORCE_ACC}, "
                     f"aold from theta={THETA_BOOT} bootstrap)") +
                    f", group_size={head['group_size']}, stack_limit={head['stack_limit']}, "
                    f"waves_per_cu={head['waves_per_cu']} (autotuned)",
            "learned_criterion": (os.path.relpath(args.model, ROOT)
                                  if "learned_fp32" in results else None),
            "n_probes": len(probes),
            "p99_rel_err": head["err"]["p99"],
            "p99_within_1e-2": bool(head["err"]["p99"] <= 1e-2),
        },
        "roofline": {"bound": "valu", "achieved": head["achieved_tflops"],
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": head["achieved_tflops"] / PEAK_FP32_TFLOPS,
                     "traffic": (profiled_traffic(PROFILE, head["kernel_ms"])
                                 if "learned_fp32" in results else None)},
        "cpu_baseline": None if cpu is None else {**{k: cpu[k] for k in
                                                     ("value", "unit", "cores", "kind", "sample")},
                                                  "host": host},
        "walks": results,
        "bootstrap_fp64": {"err": err_stats(ref_e), **{k: boot[3][k] for k in
                                                      ("nodes_visited", "multipole", "pairs")}},
        "build_ms": build_ms,
    }
    if cpu is not None:
        line["gpu_over_cpu"] = head["evals_per_s"] / cpu["value"]
    for k in ("fp32", "fp64"):
        if k in results:
            line[f"speedup_vs_plain_{k}"] = head["evals_per_s"] / results[k]["evals_per_s"]
    if results.get("global_tolerance_fp32"):
        line["speedup_vs_global_tolerance"] = (head["evals_per_s"] /
                                               results["global_tolerance_fp32"]["evals_per_s"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
