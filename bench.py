"""Traversal benchmark: incoherent closest-hit Grays/s on one MI355X.

Workload: the 1024 x 1024-cell fBm heightfield (2,097,152 triangles, SAH sweep
BVH with 2-triangle leaves) and 2^24 incoherent rays (cosine-distributed from
the hits of 1920 x 1080 primary rays), resident in HBM before timing.  A step
is one closest-hit launch over all 2^24 rays.

roofline: algorithmic bytes per ray B = 32 (ray) + 16 (hit) + 32 * N_box +
52 * N_tri, where N_box / N_tri are the reference escape walk's box and
triangle tests for that ray (counted on the device by the reference-walk
counting kernel, which the tests pin to the CPU oracle).  achieved = sum(B) /
average launch time (HIP events on the library stream), peak 8 TB/s.
cpu_baseline: the CPU oracle's escape walk (single thread) on a bounded
sample of the same rays.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--rays", type=int, default=1 << 24)
    p.add_argument("--kernel", default="wide", choices=["wide", "reference"])
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="bounded CPU-baseline sample duration")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--extra", action="store_true", help="also time coherent / shadow sets and the reference walk")
    p.add_argument("--no-autotune", action="store_true", help="keep the default wide-kernel configuration")
    p.add_argument("--sweep", action="store_true",
                   help="print Grays/s of every wide-kernel configuration on the three ray sets and exit")
    p.add_argument("--calibrate", action="store_true",
                   help="append 3 device-to-device copies of 1 GiB (counter calibration for rocprofv3 --pmc)")
    return p.parse_args()


def traffic_from_profile():
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", "trace_wide_heightfield_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        return float(d["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(scene_arrays, rays_host, seconds):
    from oracle import oracle
    sc = oracle.Scene(scene_arrays)
    done, t0 = 0, time.perf_counter()
    chunk = 4096
    while time.perf_counter() - t0 < seconds and done < len(rays_host):
        oracle.trace(sc, rays_host[done:done + chunk])
        done += chunk
    dt = time.perf_counter() - t0
    return {"value": done / dt / 1e9, "unit": "Grays/s", "cores": 1, "kind": "port",
            "sample": "%d incoherent heightfield rays (the first of the 2^24 set), CPU oracle escape walk, "
                      "1 thread, %.1f s" % (done, dt)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (rays shard with zero "
              "communication, see DESIGN.md)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    from gpugi_amd import rt, workloads

    t_setup = time.perf_counter()
    hs = rt.HostScene.heightfield(workloads.HF_CELLS, workloads.MASTER_SEED)
    ctx = rt.Context(0)
    ctx.load_scene(hs)
    n = args.rays
    R = workloads.HeightfieldRays(ctx, n=n)
    kernel = rt.KERNEL_WIDE if args.kernel == "wide" else rt.KERNEL_REFERENCE

    if args.sweep:
        sets = (("incoherent", R.incoherent, rt.CLOSEST_HIT), ("coherent", R.primary, rt.CLOSEST_HIT),
                ("shadow_any", R.shadow, rt.ANY_HIT))
        for leaf in (1, 2, 4, 8):
            for stack in (4, 8, 12, 16, 24):
                for refill in (8, 16, 32, 64):
                    ctx.set_wide_config(leaf, stack, refill)
                    row = {}
                    for name, buf, q in sets:
                        ctx.trace_device(buf, R.hits, n, q, rt.KERNEL_WIDE)
                        ctx.sync()
                        ctx.timer_begin()
                        for _ in range(3):
                            ctx.trace_device(buf, R.hits, n, q, rt.KERNEL_WIDE)
                        row[name] = round(n / (ctx.timer_end() / 3 * 1e-3) / 1e9, 3)
                    print(json.dumps(dict(ctx.wide_config(), **row)), flush=True)
        ctx.close()
        return

    # Autotune the wide kernel on a separate incoherent sample (other seed), untimed.
    tuned = None
    if kernel == rt.KERNEL_WIDE and not args.no_autotune:
        m = min(n, 1 << 22)
        T = workloads.HeightfieldRays(ctx, n=m, seed=(workloads.MASTER_SEED + 7) & 0xffffffff)
        tuned_ms = ctx.autotune_trace(T.incoherent, T.hits, m, rt.CLOSEST_HIT, repeats=2)
        tuned = dict(ctx.wide_config(), sample_rays=m, sample_ms=round(tuned_ms, 4))
        T.free()

    # Algorithmic bytes: reference escape-walk visit counts of every ray.
    nbox_b, ntri_b = ctx.alloc(n * 4), ctx.alloc(n * 4)
    ctx.trace_counts_device(R.incoherent, nbox_b, ntri_b, n, rt.CLOSEST_HIT)
    nbox = nbox_b.download(np.uint32, n).astype(np.float64)
    ntri = ntri_b.download(np.uint32, n).astype(np.float64)
    alg_bytes = float(n * (32 + 16) + 32.0 * nbox.sum() + 52.0 * ntri.sum())
    setup_s = time.perf_counter() - t_setup

    for _ in range(args.warmup):
        ctx.trace_device(R.incoherent, R.hits, n, rt.CLOSEST_HIT, kernel)
    ctx.sync()
    w0 = time.perf_counter()
    ctx.timer_begin()
    for _ in range(args.steps):
        ctx.trace_device(R.incoherent, R.hits, n, rt.CLOSEST_HIT, kernel)
    ms_total = ctx.timer_end()
    ctx.sync()
    wall = time.perf_counter() - w0
    ms = ms_total / args.steps
    grays = n / (ms * 1e-3) / 1e9
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    hits = R.hits.download(np.int32, n * 4).reshape(n, 4)

    extra = {"setup_s": round(setup_s, 2), "wall_s_timed": round(wall, 4),
             "hit_fraction": float((hits[:, 0] >= 0).mean()),
             "mean_box_tests": float(nbox.mean()), "mean_tri_tests": float(ntri.mean()),
             "alg_bytes_per_ray": alg_bytes / n, "wide_config": tuned or None}
    if args.extra:
        for name, buf, q in (("coherent_closest", R.primary, rt.CLOSEST_HIT), ("shadow_any", R.shadow, rt.ANY_HIT)):
            ctx.trace_device(buf, R.hits, n, q, kernel)
            ctx.sync()
            ctx.timer_begin()
            for _ in range(5):
                ctx.trace_device(buf, R.hits, n, q, kernel)
            extra[name + "_grays"] = n / (ctx.timer_end() / 5 * 1e-3) / 1e9
        ctx.timer_begin()
        ctx.trace_device(R.incoherent, R.hits, n, rt.CLOSEST_HIT, rt.KERNEL_REFERENCE)
        extra["reference_walk_incoherent_grays"] = n / (ctx.timer_end() * 1e-3) / 1e9

    if args.calibrate:
        a, b = ctx.alloc(1 << 30), ctx.alloc(1 << 30)
        for _ in range(3):
            b.copy_from(a, 1 << 30)
        ctx.sync()
        a.free()
        b.free()
    traffic = traffic_from_profile()
    result = {
        "metric": "incoherent closest-hit Grays/s",
        "value": round(grays, 4),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "heightfield 1024x1024 cells (2,097,152 triangles), 2^24 incoherent "
                               "cosine rays from 1920x1080 primary hits, closest hit",
                   "rays": n, "kernel": args.kernel},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "extra": extra,
    }
    if not args.no_cpu_baseline:
        rays_host = R.incoherent.download(np.float32, n * 8).reshape(-1, 8)
        result["cpu_baseline"] = cpu_baseline(hs.arrays(), rays_host, args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    R.free()
    nbox_b.free()
    ntri_b.free()
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
