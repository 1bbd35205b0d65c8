"""Headline benchmark: refined points/s of ``generate`` with the graph resident in HBM.

Workload (the large-dynamic-range precision run): N = 16,777,216 points in 3-D drawn from a
mixture of 64 isotropic Gaussians (centres U[0,1]^3, widths log-uniform in [1e-4, 1e-1]),
n0 = 1000, k = 30, Matern-3/2 covariance (variance 1, scale 0.05) tabulated at r = 0 plus
2000 log bins over [1e-6, 2] with jitter 1e-5, xi ~ N(0,1) float32.  Synthetic data with
fixed seeds.

A step is one full ``generate`` call (FP64 seed block, fused refinement, scatter to the
caller's order) on the resident graph.  ``value`` is FP32-increment throughput (refined
points / s); the same line reports the plain FP64 and plain FP32 paths, their ratio, the
error of the FP32 paths against the FP64 path (itself checked point by point against the
FP64 oracle on a random sample), the refinement kernel's roofline figures from HIP
events on its own stream, and the CPU baseline (the FP64 oracle restatement of the
reference's fast_jit refinement on the host cores, bounded sample).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
# HBM bytes per launch of the fp32_increment weights kernel at N=16,777,216, k=30
# (profiles/refine_pmc_16m.md: FETCH_SIZE x2.0 + WRITE_SIZE x1.0, copy-calibrated)
TRAFFIC_WEIGHTS_16M_K30 = 2.5835e10
PEAK_FP64_TFLOPS = 78.6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def clustered(n, d, n_clusters, wmin, wmax, seed):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, (n_clusters, d))
    widths = np.exp(rng.uniform(np.log(wmin), np.log(wmax), n_clusters))
    lab = rng.integers(0, n_clusters, n)
    pts = np.empty((n, d), np.float32)
    step = 1 << 22
    for s in range(0, n, step):
        e = min(n, s + step)
        pts[s:e] = (centres[lab[s:e]] + widths[lab[s:e], None] * rng.standard_normal((e - s, d))).astype(np.float32)
    return pts


def algorithmic_bytes(k, d, sp, sv):
    """Per refined point: indices, neighbour coordinates and values, own coordinates, xi, output."""
    return 4 * k + k * d * sp + k * sv + d * sp + sv + sv


def algorithmic_flops(k, d):
    """Per refined point: pair distances, interpolation, Cholesky (k^3/3), two solves, the rest."""
    pairs = k * (k + 1) // 2
    return pairs * (3 * d) + pairs * 3 + k ** 3 / 3 + 2 * k * k + 4 * k


# Algorithmic bytes per unit of each timed stage (the unit is a refined point unless noted):
# the data the stage must read or write once, with neighbour coordinates and values counted
# per gather.  The weights phase reads the k + 1 points' coordinates and the k indices and
# writes k weights, sigma xi, the recomputation slot and the output position; the FP64
# list (per recomputed point) reads the same and writes k FP64 weights and sigma xi; the
# dataflow reads indices, weights, slot, output position, sigma xi and the k neighbour
# values and writes the value and the output.  Build stages (per point): the tree reads the
# points and writes them with the permutation and split dimensions; the kNN reads its query
# and writes k indices; depths read the k indices and write a depth; the ordering reads and
# writes points, index, neighbours and depth; the schedule (per refined point) reads indices
# and coordinates and writes them in slot order with the slot.
def stage_bytes(stage, k, d):
    return {
        "weights": (k + 1) * d * 4 + 4 * k + 4 * k + 12,
        "fp64_list": (k + 1) * d * 4 + 4 * k + 4 + 8 * k + 8,
        "dataflow": 12 * k + 20,
        "refine": algorithmic_bytes(k, d, 4, 8),
        "tree": 2 * d * 4 + 8,
        "knn": d * 4 + 4 * k,
        "depths": 4 * k + 4,
        "order": 2 * (d * 4 + 4 + 4 * k + 4),
        "schedule": 2 * (4 * k + d * 4) + 4,
    }.get(stage)


def stage_report(times, steps, units, flops, peak_tflops, k, d):
    """Per-stage mean device ms, algorithmic GB/s (and fraction of HBM peak) and, where the
    stage does floating-point work, TFLOP/s and its fraction of the vector peak."""
    out = {}
    for name, (ms, cnt) in times.items():
        m = ms / steps
        e = {"ms": round(m, 4)}
        b = stage_bytes(name, k, d)
        u = units.get(name)
        if b is not None and u and m > 0:
            gbs = b * u / (m * 1e-3) / 1e9
            e.update(bytes_per_unit=b, units=int(u), hbm_GBs=round(gbs, 1), hbm_frac=round(gbs / PEAK_HBM_GBS, 4))
        f = flops.get(name)
        if f and m > 0:
            tf = f / (m * 1e-3) / 1e12
            e.update(tflops=round(tf, 3), valu_frac=round(tf / peak_tflops[name], 4))
        out[name] = e
    return out


def cpu_quota():
    """CPUs this process may use by its cgroup quota (cpu.max), or None without one."""
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        return None if q == "max" else round(int(q) / int(per), 2)
    except (OSError, ValueError):
        return None


def run_differentiation(args):
    """The differentiation run: N=4,194,304 uniform in [0,1]^3, n0=500, k=20, RBF (variance 1,
    scale 0.1) on r=0 plus 1000 log bins over [1e-5, 2], jitter 1e-5.  L = sum w_i values_i
    with w ~ N(0,1).  Times generate and its vector-Jacobian product (adjoint kernel + FP64
    seed-block adjoint) in FP64 and FP32, and checks the gradients: the dot-product identity
    <w, J u> = <J^T w, u> (generate is linear in xi), central differences with h=1e-3 on 64
    random xi coordinates and with h=1e-8 x |value| on 64 table entries with non-negligible
    gradient (the loss is strongly curved in the table: at this length scale the local Schur
    complements are tiny, so h=1e-6 x |value| already carries ~1% truncation error and
    h=1e-4 x |value| makes some local systems indefinite), log-det, and the inverse round
    trip."""
    import torch

    import graphgp_amd as gp

    n, n0, k = args.n, 500, 20
    rng = np.random.default_rng(3)
    pts = torch.from_numpy(rng.uniform(0, 1, (n, 3)).astype(np.float32)).cuda()
    cov = gp.extras.rbf_kernel(variance=1.0, scale=0.1, r_min=1e-5, r_max=2.0, n_bins=1000, jitter=1e-5)
    bins, vals = cov
    torch.cuda.synchronize()
    t = time.time()
    g = gp.build_graph(pts, n0=n0, k=k)
    torch.cuda.synchronize()
    build_ms = (time.time() - t) * 1e3
    xi = torch.from_numpy(rng.standard_normal(n)).cuda()
    w = torch.from_numpy(rng.standard_normal(n)).cuda()
    n_ref = n - n0

    def timed(fn, steps):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(steps):
            out = fn()
        b.record()
        torch.cuda.synchronize()
        return out, a.elapsed_time(b) / steps

    res = {}
    for prec in ("fp64", "fp32"):
        x = xi if prec == "fp64" else xi.float()
        ww = w if prec == "fp64" else w.float()
        v, fwd_ms = timed(lambda: gp.generate(g, cov, x, precision=prec), args.steps)
        (gx, gv), vjp_ms = timed(lambda: gp.generate_vjp(g, cov, x, v, ww, precision=prec), args.steps)
        res[prec] = dict(v=v, gx=gx, gv=gv, fwd_ms=fwd_ms, vjp_ms=vjp_ms)
        log(f"{prec}: generate {fwd_ms:.2f} ms, vjp {vjp_ms:.2f} ms")
    gx64, gv64 = res["fp64"]["gx"], res["fp64"]["gv"]
    u = torch.from_numpy(np.random.default_rng(4).standard_normal(n)).cuda()
    lhs = float((w * gp.generate(g, cov, u, precision="fp64")).sum())
    rhs = float((gx64 * u).sum())
    loss = lambda x_, vv: float((w * gp.generate(g, (bins, vv), x_, precision="fp64")).sum())  # noqa: E731
    fd_xi = 0.0
    for j in np.random.default_rng(5).choice(n, 64, replace=False):
        e = torch.zeros_like(xi)
        e[j] = 1e-3
        fd = (loss(xi + e, vals) - loss(xi - e, vals)) / 2e-3
        fd_xi = max(fd_xi, abs(fd - float(gx64[j])) / max(abs(float(gx64[j])), 1e-3 * float(gx64.abs().max())))
    nz = torch.nonzero(gv64.abs() > 1e-3 * gv64.abs().max()).flatten().cpu().numpy()
    fd_tab_all = []
    for j in np.random.default_rng(6).choice(nz, min(64, len(nz)), replace=False):
        h = 1e-8 * max(abs(float(vals[j])), 1e-6)
        e = torch.zeros_like(vals)
        e[j] = h
        fd = (loss(xi, vals + e) - loss(xi, vals - e)) / (2 * h)
        fd_tab_all.append(abs(fd - float(gv64[j])) / abs(float(gv64[j])))
    fd_tab = max(fd_tab_all)
    ld64 = float(gp.generate_logdet(g, cov, precision="fp64"))
    ld32 = float(gp.generate_logdet(g, cov, precision="fp32"))
    xb = gp.generate_inv(g, cov, res["fp64"]["v"], precision="fp64")
    rt = float((xb - xi).abs().max())
    gx32, gv32 = res["fp32"]["gx"], res["fp32"]["gv"]
    line = {
        "metric": "refined points/s of the vector-Jacobian product of generate (adjoint kernel)",
        "value": round(n_ref / (res["fp32"]["vjp_ms"] * 1e-3), 1),
        "unit": "refined points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "higher_is_better": True,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "differentiation: N=4,194,304 uniform [0,1]^3, n0=500, k=20, RBF scale 0.1, "
                               "1000 log bins over [1e-5, 2], jitter 1e-5, L = sum w*values",
                   "n": n, "k": k, "n0": n0},
        "build_graph_ms": round(build_ms, 2),
        "fp64": {"generate_ms": round(res["fp64"]["fwd_ms"], 3), "vjp_ms": round(res["fp64"]["vjp_ms"], 3)},
        "fp32": {"generate_ms": round(res["fp32"]["fwd_ms"], 3), "vjp_ms": round(res["fp32"]["vjp_ms"], 3)},
        "checks": {
            "dot_product_rel": abs(lhs - rhs) / abs(lhs),
            "fd_xi_h1e-3_64coords_max_rel": fd_xi,
            "fd_table_h1e-8rel_64entries_max_rel": fd_tab,
            "fd_table_h1e-8rel_64entries_median_rel": float(np.median(fd_tab_all)),
            "fp32_vs_fp64_xi_bar_max_rel": float((gx32.double() - gx64).abs().max() / gx64.abs().max()),
            "fp32_vs_fp64_table_bar_rel_l2": float((gv32 - gv64).norm() / gv64.norm()),
            "logdet_fp64": ld64,
            "logdet_fp32_rel": abs(ld32 - ld64) / abs(ld64),
            "inverse_round_trip_max_abs_fp64": rt,
        },
    }
    print(json.dumps(line), flush=True)


def run_billion(args):
    """The billion-point run: N uniform in [0,1]^3 (default 1e9), n0=1000, k=16, RBF (variance
    1, scale 0.01) on r=0 plus 4000 log bins over [1e-6, 2], jitter 1e-5, xi ~ N(0,1) float32,
    int32 neighbour indices.  build_graph runs stage by stage (tree, kNN, depths, order,
    offsets, and the dataflow schedule that build_graph also computes when it fits) so that
    each stage's scratch is freed before the next (the graph alone is ~85 GB at 1e9), then
    one FP32 generate and one FP64 log-det.  Per stage: time, the stage's
    algorithmic bytes / time, and for the refinement its flops / time against the VALU peak."""
    import torch

    import graphgp_amd as gp
    from graphgp_amd import _lib
    from graphgp_amd.graph import _offsets, _schedule

    n, n0, k, d = args.n, 1000, 16, 3
    log(f"billion workload N={n} k={k}")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(2025)
    pts = torch.rand((n, d), generator=gen, device="cuda", dtype=torch.float32)
    cov = gp.extras.rbf_kernel(variance=1.0, scale=0.01, r_min=1e-6, r_max=2.0, n_bins=4001, jitter=1e-5)
    n_ref = n - n0
    stages = {}

    def stage(name, fn, nbytes):
        torch.cuda.synchronize()
        t = time.time()
        out = fn()
        torch.cuda.synchronize()
        dt = time.time() - t
        stages[name] = {"s": round(dt, 4), "algorithmic_GBs": round(nbytes / dt / 1e9, 1)}
        log(f"{name}: {dt:.3f} s, {torch.cuda.max_memory_allocated() / 2**30:.1f} GiB peak so far")
        return out

    levels = int(math.floor(math.log2(n))) + 1
    tp, sd, idx = stage("tree", lambda: gp.build_tree(pts), 2 * n * (d * 4 + 4) * levels)
    del pts
    torch.cuda.empty_cache()
    nb = stage("knn", lambda: gp.query_preceding_neighbors(tp, sd, n0=n0, k=k), n * d * 4 + n_ref * k * 4)
    del sd
    torch.cuda.empty_cache()
    dep = stage("depths", lambda: gp.compute_depths(nb, n0=n0), n_ref * k * 20)
    torch.cuda.empty_cache()
    po, io, no, do = stage("order", lambda: gp.order_by_depth(tp, idx, nb, dep),
                           2 * (n * (d * 4 + 8) + n_ref * k * 4) + n_ref * k * 4)
    del tp, idx, nb, dep
    torch.cuda.empty_cache()
    offsets = stage("offsets", lambda: _offsets(do, int(do[-1].item())), n * 4)
    g = gp.Graph(po, no, offsets, io)
    del do
    torch.cuda.empty_cache()
    # as build_graph does: kept on the graph, or skipped when it would not fit in half the
    # free memory (the generate then runs in graph order)
    sch = stage("schedule", lambda: _schedule(g, g.points, g.neighbors),
                2 * n_ref * k * 4 + 2 * n * d * 4 + 4 * n_ref * 8)
    log(f"schedule: {'computed' if sch is not None else 'skipped (memory)'}")
    del sch
    torch.cuda.empty_cache()
    xi = torch.randn(n, generator=gen, device="cuda", dtype=torch.float32)
    L = _lib.lib()
    ks, ke = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ks.record()
    ke.record()
    L.ggp_profile_events(ks.cuda_event, ke.cuda_event)
    v = stage("generate_fp32", lambda: gp.generate(g, cov, xi, precision="fp32"), algorithmic_bytes(k, d, 4, 4) * n_ref)
    L.ggp_profile_events(None, None)
    kernel_s = ks.elapsed_time(ke) * 1e-3
    nan = bool(torch.isnan(v).any())
    del v
    torch.cuda.empty_cache()
    ld = stage("logdet_fp64", lambda: float(gp.generate_logdet(g, cov, precision="fp64")), n_ref * (4 * k + 12 * k))
    tflops = algorithmic_flops(k, d) * n_ref / kernel_s / 1e12
    build_s = sum(stages[s]["s"] for s in ("tree", "knn", "depths", "order", "offsets", "schedule"))
    line = {
        "metric": "billion-point field: build_graph + FP32 generate + FP64 log-det",
        "value": round(n_ref / stages["generate_fp32"]["s"], 1),
        "unit": "refined points/s (generate)",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "N uniform [0,1]^3, n0=1000, k=16, RBF scale 0.01, 4000 log bins over [1e-6, 2], "
                               "jitter 1e-5", "n": n, "k": k, "n0": n0, "n_batches": len(offsets) - 1},
        "stages": stages,
        "build_graph_s": round(build_s, 3),
        "generate_refine_kernel_s": round(kernel_s, 4),
        "generate_valu_TFLOPs": round(tflops, 2),
        "generate_valu_frac": round(tflops / PEAK_FP32_TFLOPS, 4),
        "logdet_fp64": ld,
        "generate_has_nan": nan,
        "peak_device_GiB": round(torch.cuda.max_memory_allocated() / 2**30, 1),
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["headline", "differentiation", "billion"], default="headline",
                    help="headline: the generate benchmark (default); differentiation: the adjoint run; "
                         "billion: the billion-point field (use --n to scale it)")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16_777_216)
    ap.add_argument("--k", type=int, default=30)
    ap.add_argument("--n0", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--sample", type=int, default=2000, help="points checked against the oracle")
    ap.add_argument("--serve-requests", type=int, default=2000, help="requests of the serving run (0: skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB device-to-device (the reference dispatch for the FETCH_SIZE / "
                         "WRITE_SIZE counter scale) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); every box in this setup has one GPU",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        import torch

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read": 1 << 30, "bytes_written": 1 << 30}))
        return None
    if args.workload == "differentiation":
        if args.n == 16_777_216:
            args.n = 4_194_304
        return run_differentiation(args)
    if args.workload == "billion":
        if args.n == 16_777_216:
            args.n = 1_000_000_000
        return run_billion(args)

    import torch

    import graphgp_amd as gp
    from graphgp_amd import _lib
    import oracle as O

    torch.cuda.set_device(0)
    n, d, k, n0 = args.n, 3, args.k, args.n0
    log(f"workload N={n} d={d} k={k} n0={n0}")
    t = time.time()
    pts = clustered(n, d, 64, 1e-4, 1e-1, seed=2024)
    xi = np.random.default_rng(7).standard_normal(n).astype(np.float32)
    log(f"inputs generated in {time.time() - t:.1f}s")
    cov = gp.extras.matern_kernel(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2.0, n_bins=2001, jitter=1e-5)
    pts_d = torch.from_numpy(pts).cuda()
    xi_d = torch.from_numpy(xi).cuda()

    torch.cuda.synchronize()
    t = time.time()
    g = gp.build_graph(pts_d, n0=n0, k=k)
    torch.cuda.synchronize()
    build_s = time.time() - t
    log(f"build_graph {build_s * 1e3:.1f} ms, {len(g.offsets) - 1} batches")
    n_ref = n - n0
    # build stages: one more build (and schedule) with the library's stage timing on
    _lib.stage_profile(True)
    g2 = gp.build_graph(pts_d, n0=n0, k=k)
    build_stages = stage_report(_lib.stage_times(), 1, {"tree": n, "knn": n_ref, "depths": n_ref, "order": n,
                                                        "schedule": n_ref}, {}, {}, k, d)
    _lib.stage_profile(False)
    del g2
    log(f"build stages: {build_stages}")

    start_ev = torch.cuda.Event(enable_timing=True)
    stop_ev = torch.cuda.Event(enable_timing=True)
    L = _lib.lib()

    def timed(precision, steps, warmup):
        out = None
        for _ in range(warmup):
            out = gp.generate(g, cov, xi_d, precision=precision)
        torch.cuda.synchronize()
        kms = []
        t0 = time.time()
        for _ in range(steps):
            ks, ke = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ks.record()  # creates the underlying hipEvents; the library re-records them
            ke.record()  # around the refinement kernel on the same stream
            L.ggp_profile_events(ks.cuda_event, ke.cuda_event)
            out = gp.generate(g, cov, xi_d, precision=precision)
            L.ggp_profile_events(None, None)
            kms.append((ks, ke))
        torch.cuda.synchronize()
        wall = (time.time() - t0) / steps
        kernel_ms = float(np.mean([a.elapsed_time(b) for a, b in kms]))
        # stage times from a separate pass, so that the timed steps carry no stage events
        _lib.stage_profile(True)  # events around every stage of each step, on the call's stream
        stage_steps = max(2, steps // 4)
        for _ in range(stage_steps):
            gp.generate(g, cov, xi_d, precision=precision)
        torch.cuda.synchronize()
        stages = _lib.stage_times()
        _lib.stage_profile(False)
        return out, wall, kernel_ms, stages, stage_steps

    results = {}
    for prec, steps in (("fp32_increment", args.steps), ("fp64", max(2, args.steps // 2)),
                        ("fp32", max(2, args.steps // 2))):
        out, wall, kms, stages, stage_steps = timed(prec, steps, args.warmup)
        results[prec] = dict(out=out, ms=wall * 1e3, kernel_ms=kms, pts_per_s=n_ref / wall, stages=stages,
                             stage_steps=stage_steps,
                             steps=steps)
        log(f"{prec}: {wall * 1e3:.2f} ms/step, dominant refine kernel {kms:.2f} ms, {n_ref / wall / 1e9:.3f} G pts/s")

    # FP64 path checked point by point against the oracle (given its neighbours' FP64 values)
    v64 = results["fp64"]["out"].cpu().numpy()
    gpts = g.points.cpu().numpy()
    gnb = g.neighbors.cpu().numpy()
    gidx = g.indices.cpu().numpy()
    vg = v64[gidx]  # graph order
    xg = xi[gidx]
    ocov = (cov[0].cpu().numpy(), cov[1].cpu().numpy())
    rng = np.random.default_rng(1)
    sample = rng.choice(n_ref, min(args.sample, n_ref), replace=False)
    loc_err = 0.0
    for r in sample:
        mean, std = O.local_conditional(gpts, gnb[r], n0 + r, ocov, vg[gnb[r]])
        loc_err = max(loc_err, abs(mean + std * float(xg[n0 + r]) - vg[n0 + r]))
    errs = {}
    for prec in ("fp32_increment", "fp32"):
        dv = results[prec]["out"].cpu().numpy().astype(np.float64) - v64
        errs[prec] = dict(max_abs=float(np.nanmax(np.abs(dv))), rms=float(np.sqrt(np.nanmean(dv ** 2))))
    _, cnt = gp.generate(g, cov, xi_d, precision="fp32_increment", return_counters=True)
    fix_frac = cnt["fp64_recomputed"] / n_ref
    log(f"fp32_increment: {cnt['fp64_recomputed']} points recomputed in FP64 ({100 * fix_frac:.3f}%)")
    # local (per-point) error of the FP32 paths: invert the FP64 values and compare with xi
    v64_d = results["fp64"]["out"]
    for prec in ("fp32_increment", "fp32"):
        xb = gp.generate_inv(g, cov, v64_d.float(), precision=prec).cpu().numpy().astype(np.float64)
        dx = np.abs(xb - xi)
        errs[prec]["local_xi_max_abs"] = float(np.nanmax(dx))
        errs[prec]["local_xi_rms"] = float(np.sqrt(np.nanmean(dx ** 2)))
    # log-det of each FP32 path against the FP64 one
    ld64 = float(gp.generate_logdet(g, cov, precision="fp64"))
    for prec in ("fp32_increment", "fp32"):
        ld = float(gp.generate_logdet(g, cov, precision=prec))
        errs[prec]["logdet_abs"] = abs(ld - ld64)
        errs[prec]["logdet_rel"] = abs(ld - ld64) / abs(ld64)
    errs["logdet_fp64"] = ld64
    log(f"fp64 vs oracle (sampled, local): {loc_err:.3e}; errors vs fp64: {errs}")

    # every stage of the timed steps: device ms, algorithmic GB/s, TFLOP/s where it computes
    n_fix = cnt["fp64_recomputed"]
    units = {"weights": n_ref, "fp64_list": n_fix, "dataflow": n_ref, "refine": n_ref}
    fl = algorithmic_flops(k, d)
    stage_lines = {}
    for prec, peak in (("fp32_increment", PEAK_FP32_TFLOPS), ("fp64", PEAK_FP64_TFLOPS), ("fp32", PEAK_FP32_TFLOPS)):
        rr = results[prec]
        flops_ = {"weights": fl * n_ref, "fp64_list": fl * n_fix, "dataflow": (2 * k + 4) * n_ref,
                  "refine": fl * n_ref, "seed_block": n0 ** 3 / 3 + 2 * n0 ** 2}
        peaks = {"weights": PEAK_FP32_TFLOPS, "fp64_list": PEAK_FP64_TFLOPS, "dataflow": PEAK_FP32_TFLOPS,
                 "refine": peak, "seed_block": PEAK_FP64_TFLOPS}
        by = dict(units)
        if prec == "fp32":
            by["refine"] = None  # FP32 fused kernel: its own byte count below
        stage_lines[prec] = stage_report(rr["stages"], rr["stage_steps"], by, flops_, peaks, k, d)
        if prec == "fp32" and "refine" in stage_lines[prec]:
            e = stage_lines[prec]["refine"]
            b = algorithmic_bytes(k, d, 4, 4)
            gbs = b * n_ref / (e["ms"] * 1e-3) / 1e9
            e.update(bytes_per_unit=b, units=n_ref, hbm_GBs=round(gbs, 1), hbm_frac=round(gbs / PEAK_HBM_GBS, 4))
    log(f"stages: {stage_lines}")

    # roofline of the dominant kernel of the FP32-increment generate, its weights phase
    # (VALU bound: no MFMA in this kernel); the library's events bracket exactly that launch.
    # HBM traffic per launch comes from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs of the
    # same kernel and workload, scaled by a device copy of known size: the figure recorded
    # in profiles/refine_pmc_16m.md (a separate counter measurement; the summaries are not
    # shipped with the runtime tree, so the number is kept here).
    traffic = TRAFFIC_WEIGHTS_16M_K30
    r32 = results["fp32_increment"]
    kernel_s = r32["kernel_ms"] * 1e-3
    flops = algorithmic_flops(k, d) * n_ref
    bytes_ = algorithmic_bytes(k, d, 4, 4) * n_ref
    tflops = flops / kernel_s / 1e12
    gbs = bytes_ / kernel_s / 1e9
    r64 = results["fp64"]
    roofline = {
        "bound": "valu",
        "achieved": round(tflops, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(tflops / PEAK_FP32_TFLOPS, 4),
        "traffic": traffic if n == 16_777_216 and k == 30 else None,
        "traffic_source": "profiles/refine_pmc_16m.md (FETCH_SIZE x2.0 + WRITE_SIZE x1.0, copy-calibrated)",
        "kernel": "refine_kernel<float,32,4,3> weights phase of the fp32_increment generate",
        "kernel_ms": round(r32["kernel_ms"], 4),
        "flop_per_point": algorithmic_flops(k, d),
        "bytes_per_point": algorithmic_bytes(k, d, 4, 4),
        "hbm_achieved_GBs": round(gbs, 1),
        "hbm_frac": round(gbs / PEAK_HBM_GBS, 4),
        "fp64_kernel_ms": round(r64["kernel_ms"], 4),
        "fp64_tflops": round(algorithmic_flops(k, d) * n_ref / (r64["kernel_ms"] * 1e-3) / 1e12, 3),
        "fp64_frac": round(algorithmic_flops(k, d) * n_ref / (r64["kernel_ms"] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4),
    }

    cpu_baseline = None
    if not args.no_cpu:
        cores = len(os.sched_getaffinity(0))  # every CPU this process may run on
        O.set_threads(cores)
        opts = gpts.astype(np.float64)
        # calibrate, then a bounded sample of about args.cpu_seconds
        m = min(20000, n_ref)
        t = time.time()
        O.refine(opts[: n0 + m], gnb[:m], g.offsets, ocov, vg[:n0], xg[n0:n0 + m])
        rate = m / (time.time() - t)
        m = int(min(n_ref, max(m, rate * args.cpu_seconds)))
        t = time.time()
        O.refine(opts[: n0 + m], gnb[:m], g.offsets, ocov, vg[:n0], xg[n0:n0 + m])
        dt = time.time() - t
        cpu_baseline = {
            "value": round(m / dt, 1),
            "unit": "refined points/s",
            "cores": cores,
            "cpu_quota": cpu_quota(),
            "kind": "port",
            "sample": f"first {m} refined points of the same graph (FP64 oracle restatement of the fast_jit "
                      f"refinement, refine.py:123-145), {dt:.1f} s",
        }
        log(f"cpu baseline: {m / dt:.0f} pts/s on {cores} threads")

    serve = None
    if args.serve_requests > 0:
        from graphgp_amd.serving import serve_benchmark

        serve = serve_benchmark(args.serve_requests)
        log(f"serving: {serve}")

    line = {
        "metric": "refined points/sec at N=16M,k=30 (FP32 vs plain FP64); p50/p99 serve latency",
        "value": round(r32["pts_per_s"], 1),
        "unit": "refined points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "generate, N=16,777,216 3-D 64-cluster mixture (widths 1e-4..1e-1), n0=1000, k=30, "
                        "Matern-3/2 scale 0.05, 2001-entry table, jitter 1e-5, graph resident",
            "n": n, "d": d, "k": k, "n0": n0, "n_batches": len(g.offsets) - 1,
            "precision": "fp32_increment",
        },
        "fp64_points_per_s": round(r64["pts_per_s"], 1),
        "fp64_ms_per_step": round(r64["ms"], 4),
        "fp32_points_per_s": round(results["fp32"]["pts_per_s"], 1),
        "speedup_fp32inc_vs_fp64": round(r32["pts_per_s"] / r64["pts_per_s"], 3),
        "error_vs_fp64": errs,
        "fp32_increment_fp64_recomputed_fraction": fix_frac,
        "fp64_vs_oracle_sampled_max_abs": loc_err,
        "build_graph_ms": round(build_s * 1e3, 1),
        "build_stages": build_stages,
        "stages": stage_lines,
        "stages_note": ("stage times come from separate untimed steps with events around every stage; "
                        "fp32_increment: the seed block, the value buffer's fill and the FP64 table run on a "
                        "second stream beside the weights phase (their time overlaps the weights stage); "
                        "weights spans both weights launches"),
        "roofline": roofline,
        "cpu_baseline": cpu_baseline,
        "serve": serve,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
