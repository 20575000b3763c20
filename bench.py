#!/usr/bin/env python3
"""Benchmark of the raster3d path on one MI355X: Gaussian arrays -> node2 NanoVDB grid.

Workload (default): 2M Gaussians from a 64-cluster mixture (cluster centres U[-1,1]^3,
members N(0, 0.1^2) per axis), log-scales N(-5, 0.7^2), random quaternions, opacity logits
N(0,1), voxel size 0.005, RGBA8 values (the reference value semantics). One step is the
whole entry point on device-resident raw arrays: preprocessing, per-Gaussian setup, pair
emission, 64-bit key sort, per-voxel reduction and node2 grid build.

Prints one JSON line (see the keys below). ``--gpus N`` with N > 1 exits: this build runs on
one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nanovdb-editor_amd"))

METRIC = "Gaussians/s voxelized + fwd+bwd ms @2M vs plain GPU path; % HBM roofline"
HBM_PEAK_GBS = 8000.0
# FP32 VALU operations per second without packed or fused instructions (the predicate keeps the
# reference's unfused operation order and the build disables packed FP32): 256 CUs x 4 SIMDs x
# 16 lanes x 2.4 GHz
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
# FP32 operations of the walk (counted in the kernel's instruction stream): per candidate voxel,
# the new lattice plane's z term (convert, scale, offset: 3), per axis the plane term (1), the
# carried and the new corner (add, mul, mul: 3 each), the median with 0, its square and the
# subtraction (3), and the sign test (1) = 34; per kept pair, ordering the box extremes (6), the
# minimum sample distance (32) and the 8-bit alpha with its error bound (37) = 75
WALK_OPS_PER_CANDIDATE = 34
ALPHA_OPS_PER_PAIR = 75


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--gaussians", type=int, default=2_000_000)
    ap.add_argument("--value-mode", choices=["rgba8", "density"], default="rgba8")
    ap.add_argument("--cpu-sample", type=int, default=6_000,
                    help="Gaussians in the CPU-oracle baseline sample (0 disables it)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--batch-steps", type=int, default=2,
                    help="timed calls of the 256-cloud batched entry point (0 skips it)")
    ap.add_argument("--batch-workers", type=int, default=4)
    ap.add_argument("--big-steps", type=int, default=1,
                    help="timed calls of the 16M-Gaussian slabbed workload (0 skips it)")
    ap.add_argument("--fit-steps", type=int, default=1,
                    help="timed fit steps (forward + L2 + backward) of the 1M-Gaussian fit workload (0 skips it)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default sort shape instead of autotuning it for the workload")
    ap.add_argument("--fwd-bwd-steps", type=int, default=3,
                    help="timed density forward+backward steps (torch.autograd path); 0 disables")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (reads 1 GiB, writes 1 GiB per launch): the "
                         "reference kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    return ap.parse_args()


def copy_calibration():
    import torch
    dev = torch.device("cuda", 0)
    x = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    times = []
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        y = x + 0.0
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    nbytes = x.numel() * 4
    print(json.dumps({"copy_calibration": True, "read_bytes_per_launch": nbytes, "write_bytes_per_launch": nbytes,
                      "ms": times, "gbs": 2 * nbytes / (min(times) * 1e-3) / 1e9, "check": float(y[-1])}))


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` as measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on the
    default workload (the 'bench-traffic' line of the committed profile summary), or None."""
    path = os.path.join(ROOT, "profiles", "raster_clustered_2m_pmc.md")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "bench-traffic" and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


def fwd_bwd(cloud, raw, vox, steps):
    """Density voxelization forward + backward on the same workload through the autograd API:
    values = voxelize_density(...), loss = sum(w * values) with fixed random w, loss.backward()."""
    import torch
    from nanovdb_editor_amd.autograd import voxelize_density
    params = [raw[k].clone().requires_grad_(True) for k in ("means", "s", "q", "op")]
    w = None

    def step():
        nonlocal w
        for p in params:
            p.grad = None
        values, grid = voxelize_density(params[0], params[1], params[2], params[3], cloud.voxel_size, vox=vox)
        if w is None or w.numel() != values.numel():
            gen = torch.Generator(device=values.device).manual_seed(7)
            w = torch.randn(values.numel(), device=values.device, generator=gen)
        (values * w).sum().backward()
        return values.numel() - 2

    step()  # warm-up
    torch.cuda.synchronize()
    t_f = t_fb = 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        with torch.no_grad():
            voxelize_density(params[0], params[1], params[2], params[3], cloud.voxel_size, vox=vox)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        voxels = step()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_f += t1 - t0
        t_fb += t2 - t1
    return {"value_mode": "density", "forward_ms": 1e3 * t_f / steps, "fwd_bwd_ms": 1e3 * t_fb / steps,
            "voxels": voxels, "steps": steps,
            "backward_path": "sorted pair list" if vox.last_backward_used_pairs else "pair re-enumeration",
            "note": "float density grid forward (exact fixed-point voxel sums) and its adjoint w.r.t. means, "
                    "log-scales, quaternions, opacity logits via torch.autograd; loss = sum(w * values)"}


def fit_1m(vox, steps):
    """Fit step of the 1M-Gaussian workload: density forward of the fit cloud (seed 31) with voxel
    coordinates, the target grid (seed 32, built once) sampled at those voxels, L2 loss over the
    fit voxels, backward to all four parameter tensors."""
    import numpy as np
    import torch
    from nanovdb_editor_amd import synthetic as syn
    from nanovdb_editor_amd.autograd import voxelize_density
    from nanovdb_editor_amd.device import VALUE_DENSITY
    fit, tgt = syn.workload_fit_1m(31), syn.workload_fit_1m(32)
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32).ravel()).to(dev)
    tw = [t(a) for a in (tgt.means, tgt.quats, tgt.log_scales, tgt.sh0, tgt.opacity_logits)]
    tgrid, tst = vox.raster_node2(tw[0], tw[1], tw[2], torch.empty_like(tw[3]), tw[4], tgt.voxel_size,
                                  value_mode=VALUE_DENSITY, raw=True, sh0=tw[3])
    tcount = tst["voxel_count"] + 2
    tvals = tgrid[tgrid.numel() - 4 * tcount:].view(torch.float32)
    params = [t(a).requires_grad_(True) for a in (fit.means, fit.log_scales, fit.quats, fit.opacity_logits)]
    stats = {}

    def step():
        for p in params:
            p.grad = None
        values, grid, ijk = voxelize_density(*params, fit.voxel_size, vox=vox, with_coords=True)
        target = vox.node2_sample(tgrid, tvals, ijk[2:].contiguous())
        loss = ((values[2:] - target) ** 2).sum()
        loss.backward()
        stats["voxels"] = values.numel() - 2
        stats["loss"] = float(loss.detach())

    step()
    torch.cuda.synchronize()
    t_f = t_fb = 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        with torch.no_grad():
            voxelize_density(*params, fit.voxel_size, vox=vox)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_f += t1 - t0
        t_fb += t2 - t1
    _, st = vox.raster_node2(*(t(a) for a in (fit.means, fit.quats, fit.log_scales)), torch.empty_like(params[0]),
                             t(fit.opacity_logits), fit.voxel_size, value_mode=VALUE_DENSITY, raw=True,
                             sh0=torch.zeros_like(params[0]), copy_out=False)
    return {"workload": "fit_1m: 1M Gaussians, means U[-1,1]^3, log-scales N(-4,0.5^2), voxel size 0.005; "
                        "fit seed 31 against the target grid of seed 32, L2 over the fit voxels",
            "gaussians": fit.count, "pairs": st["pair_count"], "voxels": stats["voxels"], "target_voxels": tcount - 2,
            "forward_ms": 1e3 * t_f / steps, "fwd_bwd_ms": 1e3 * t_fb / steps, "steps": steps}


def clustered_16m(vox, steps):
    """16M Gaussians (256 clusters, log-scales N(-5.5, 0.6^2), voxel size 0.002): beyond the
    reference's 2 Mi-node and 3 GiB caps, built in voxel-x slabs of at most 2^31 pairs each."""
    import numpy as np
    import torch
    from nanovdb_editor_amd import synthetic as syn
    cloud = syn.workload_clustered_16m()
    dev = torch.device("cuda", 0)
    raw = {k: torch.from_numpy(np.ascontiguousarray(v, np.float32).ravel()).to(dev)
           for k, v in (("means", cloud.means), ("op", cloud.opacity_logits), ("q", cloud.quats),
                        ("s", cloud.log_scales), ("sh0", cloud.sh0))}
    del cloud.means, cloud.quats, cloud.log_scales, cloud.sh0
    work = {k: torch.empty_like(v) for k, v in raw.items()}
    colors = torch.empty_like(raw["sh0"])

    def step():
        for k in raw:
            work[k].copy_(raw[k])
        _, st = vox.raster_node2(work["means"], work["q"], work["s"], colors, work["op"], 0.002, raw=True,
                                 sh0=work["sh0"], copy_out=False, max_node_count=0xFFFFFFFF)
        return st

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = step()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    return {"workload": "clustered_16m: 16M Gaussians, 256 clusters (sigma 0.1), log-scales N(-5.5,0.6^2), "
                        "voxel size 0.002, RGBA8, slabs of <= 2^31 pairs",
            "gaussians": 16_000_000, "ms_per_step": ms, "gaussians_per_s": 16e6 / (ms * 1e-3),
            "pairs": st["pair_count"], "voxels": st["voxel_count"], "leaves": st["leaf_count"],
            "slabs": st["slab_count"], "grid_bytes": st["grid_size"], "steps": steps,
            "stage_ms": {k: st[k] for k in ("ms_setup", "ms_emit", "ms_sort", "ms_reduce", "ms_build")}}


def batched(vox, steps, workers):
    """The 256-cloud batch through vdbr_raster_node2_batch: the raw arrays are preprocessed in place
    by an untimed first call, then timed calls rerun the preprocessed clouds."""
    import numpy as np
    import torch
    from nanovdb_editor_amd import synthetic as syn
    clouds = syn.workload_batch_256()
    dev = torch.device("cuda", 0)
    items = []
    for c in clouds:
        d = {k: torch.from_numpy(np.ascontiguousarray(a, np.float32).ravel()).to(dev)
             for k, a in (("means", c.means), ("opacities", c.opacity_logits), ("quats", c.quats),
                          ("scales", c.log_scales), ("sh0", c.sh0))}
        d["colors"] = torch.empty_like(d["sh0"])
        d.update(voxel_size=c.voxel_size, raw=True)
        items.append(d)
    vox.raster_node2_batch(items, workers=workers, copy_out=False)
    for d in items:
        d["raw"] = False
    vox.raster_node2_batch(items, workers=workers, copy_out=False)
    torch.cuda.synchronize()

    def timed(w, k):
        t0 = time.perf_counter()
        for _ in range(k):
            _, st = vox.raster_node2_batch(items, workers=w, copy_out=False)
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t0) / k, st

    ms, st = timed(workers, steps)
    phases = dict(vox.last_batch_timing)
    graph_clouds = vox.last_batch_graph_clouds
    ms_seq, _ = timed(1, 1)
    n = sum(c.count for c in clouds)
    return {"workload": "batch_256: 256 clouds, sizes log-uniform in [1e3, 5e5], uniform or 64-cluster "
                        "(coin flip), voxel size U[0.004, 0.02], log-scales N(ln(1.5 vs), 0.4^2)",
            "clouds": len(clouds), "gaussians": n, "pairs": sum(x["pair_count"] for x in st),
            "voxels": sum(x["voxel_count"] for x in st), "workers": workers, "ms_per_batch": ms,
            "gaussians_per_s": n / (ms * 1e-3), "pairs_per_s": sum(x["pair_count"] for x in st) / (ms * 1e-3),
            "graph_clouds": graph_clouds, "last_call_phases_ms": phases,
            "ms_per_batch_one_stream": ms_seq, "steps": steps,
            "note": "one captured HIP graph per call over `workers` streams (setup of every cloud and one read "
                    "back first); an unchanged plan relaunches the graph captured by the first call"}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(cloud, sample, threads):
    """The oracle (CPU restatement of the reference path) on a fixed subsample of the workload, on
    every core this process may run on, and on one thread over a sixteenth of that sample."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    def run(count, nthreads):
        rng = np.random.default_rng(12345)
        idx = np.sort(rng.choice(cloud.count, size=min(count, cloud.count), replace=False))
        sub = cloud.subset(idx)
        t0 = time.perf_counter()
        m, op, q, s, col = oracle.preprocess(sub.means, sub.opacity_logits, sub.quats, sub.log_scales, sub.sh0)
        _, st = oracle.raster_node2(m, q, s, col, op, sub.voxel_size, nthreads=nthreads)
        return sub.count, time.perf_counter() - t0, st["pair_count"]

    # every core of this process's share: its affinity set, capped by the thread budget the host
    # grants one GPU's job (OMP_NUM_THREADS)
    affinity = len(os.sched_getaffinity(0))
    budget = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    cores = threads or (min(affinity, budget) if budget > 0 else affinity)
    n_all, dt_all, p_all = run(sample, cores)
    n_one, dt_one, p_one = run(max(sample // 16, 1), 1)
    return {"value": n_all / dt_all, "unit": "Gaussians/s", "cores": cores, "kind": "port",
            "sample": f"{n_all} Gaussians drawn (seed 12345) from the benchmark cloud, full oracle raster "
                      f"(preprocess, pair predicate, windowed node2 build, RGBA8 values) on {cores} threads in "
                      f"{dt_all:.2f} s; {p_all} pairs",
            "one_thread": {"value": n_one / dt_one, "unit": "Gaussians/s", "cores": 1,
                           "sample": f"{n_one} Gaussians (same seed), {dt_one:.2f} s, {p_one} pairs"},
            "nproc": os.cpu_count(), "affinity_cpus": affinity, "cpu_model": _cpu_model()}


def _progress(what):
    """One line on stderr as each leg of the run completes (stdout carries only the JSON line)."""
    print(f"[bench] {what} done", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    import numpy as np
    import torch
    from nanovdb_editor_amd import synthetic as syn
    from nanovdb_editor_amd.device import VALUE_DENSITY, VALUE_RGBA8, Voxelizer

    cloud = syn.clustered_cloud(args.gaussians, 64, 0.1, 0.005, -5.0, 0.7, seed=2)
    dev = torch.device("cuda", 0)
    raw = {k: torch.from_numpy(np.ascontiguousarray(v, np.float32).ravel()).to(dev)
           for k, v in (("means", cloud.means), ("op", cloud.opacity_logits), ("q", cloud.quats),
                        ("s", cloud.log_scales), ("sh0", cloud.sh0))}
    work = {k: torch.empty_like(v) for k, v in raw.items()}
    colors = torch.empty_like(raw["sh0"])
    vox = Voxelizer(0)
    mode = VALUE_RGBA8 if args.value_mode == "rgba8" else VALUE_DENSITY
    stream = torch.cuda.current_stream(dev)

    def step():
        for k in raw:
            work[k].copy_(raw[k])
        _, st = vox.raster_node2(work["means"], work["q"], work["s"], colors, work["op"], cloud.voxel_size,
                                 value_mode=mode, raw=True, sh0=work["sh0"], copy_out=False)
        return st

    tuned = None
    if args.autotune:
        # one probe step gives the pair count and sorted bit range; the sort autotuner then picks the
        # onesweep tile shape and digit width for sorts of that size (untimed, before the warmup)
        probe = step()
        torch.cuda.synchronize()
        best, rows = vox.autotune_sort(probe["pair_count"], probe["key_bits"] - probe["sort_bit_begin"])
        tuned = {"choice": best, "candidates": len(rows), "key_count": probe["pair_count"],
                 "key_bits": probe["key_bits"] - probe["sort_bit_begin"]}
        _progress("sort autotune")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(step())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = 1e3 * dt / args.steps
    _progress(f"{args.steps} timed steps")
    st = stats[-1]
    P = st["pair_count"]

    def avg(key):
        return sum(s[key] for s in stats) / len(stats)

    # the onesweep sort pass; algorithmic bytes per pass = P * (8 read + 8 write)
    passes = max(st["sort_passes"], 1)
    pass_ms = avg("ms_sort_passes") / passes
    pass_bytes = P * 16
    achieved = pass_bytes / (pass_ms * 1e-3) / 1e9 if pass_ms > 0 else 0.0
    default_workload = args.gaussians == 2_000_000 and args.value_mode == "rgba8"
    traffic = profiled_traffic("k_sort_pass") if default_workload else None
    # dominant kernel: the emission walk (k_emit_segments), one launch per step. Algorithmic bytes:
    # N * 44 (parameters) + N * 32 (per-Gaussian setup) + P * 8 (key writes); it is bound by FP32
    # issue, reported against the VALU peak from the candidate voxels it evaluates
    N = st["gaussian_count"]
    emit_bytes = N * 44 + N * 32 + P * 8
    walk_ms = avg("ms_walk_kernel")
    emit_achieved = emit_bytes / (walk_ms * 1e-3) / 1e9 if walk_ms > 0 else 0.0
    cand = st["candidate_voxels"]
    walk_ops = cand * WALK_OPS_PER_CANDIDATE + (P * ALPHA_OPS_PER_PAIR if args.value_mode == "rgba8" else 0)
    valu_tops = walk_ops / (walk_ms * 1e-3) / 1e12 if walk_ms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": args.gaussians / (ms_per_step * 1e-3),
        "unit": "Gaussians/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "clustered_2m: 2M Gaussians, 64 clusters (sigma 0.1), log-scales N(-5,0.7^2), "
                               "random quaternions, opacity logits N(0,1), voxel size 0.005",
                   "gaussians": args.gaussians, "voxel_size": cloud.voxel_size, "value_mode": args.value_mode,
                   "fragments": st["fragment_count"], "pairs": P, "voxels": st["voxel_count"],
                   "leaves": st["leaf_count"], "key_bits": st["key_bits"], "grid_bytes": st["grid_size"]},
        "roofline": {"bound": "hbm", "kernel": "k_emit_segments (pair predicate walk, staging, key writes)",
                     "achieved": emit_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": emit_achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic("k_emit_segments") if default_workload else None,
                     "avg_launch_ms": walk_ms, "algorithmic_bytes_per_launch": emit_bytes,
                     "valu": {"achieved": valu_tops, "peak": VALU_PEAK_TOPS, "unit": "T FP32 ops/s",
                              "frac": valu_tops / VALU_PEAK_TOPS, "candidate_voxels": cand,
                              "ops_per_candidate": WALK_OPS_PER_CANDIDATE,
                              "ops_per_pair": ALPHA_OPS_PER_PAIR if args.value_mode == "rgba8" else 0}},
        "sort_roofline": {"bound": "hbm", "kernel": "k_sort_pass (onesweep radix pass over 64-bit pair keys)",
                          "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                          "traffic": traffic, "avg_launch_ms": pass_ms, "algorithmic_bytes_per_launch": pass_bytes,
                          "passes": passes, "sorted_bits": st["key_bits"] - st["sort_bit_begin"],
                          "all_passes_ms": avg("ms_sort_passes")},
        "sort_autotune": tuned,
        "pairs_per_s": P / (ms_per_step * 1e-3),
        "stage_ms": {k: avg(k) for k in ("ms_setup", "ms_emit", "ms_sort", "ms_reduce", "ms_build", "ms_total",
                                        "ms_sort_passes", "ms_emit_kernel", "ms_row_kernel", "ms_walk_kernel",
                                        "ms_reduce_kernel")},
    }
    if args.fwd_bwd_steps > 0:
        out["differentiable"] = fwd_bwd(cloud, raw, vox, args.fwd_bwd_steps)
        _progress("density forward + backward")
    if args.fit_steps > 0 and default_workload:
        vox.trim()  # the fit workload needs most of HBM
        torch.cuda.empty_cache()
        fvox = Voxelizer(0)
        out["fit_1m"] = fit_1m(fvox, args.fit_steps)
        _progress("1M fit")
        fvox.close()
        torch.cuda.empty_cache()
    if args.big_steps > 0 and default_workload:
        bvox = Voxelizer(0)
        out["clustered_16m"] = clustered_16m(bvox, args.big_steps)
        _progress("16M clustered")
        bvox.close()
        torch.cuda.empty_cache()
    if args.batch_steps > 0 and default_workload:
        out["batched"] = batched(vox, args.batch_steps, args.batch_workers)
        _progress("batched clouds")
    if args.cpu_sample > 0:
        out["cpu_baseline"] = cpu_baseline(cloud, args.cpu_sample, args.cpu_threads)
        _progress("cpu baseline")
    else:
        out["cpu_baseline"] = None
    vox.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
