"""Traversal benchmark: incoherent closest-hit Grays/s on one MI355X.

Workload: the 1024 x 1024-cell fBm heightfield (2,097,152 triangles, SAH sweep
BVH with 2-triangle leaves) and 2^24 incoherent rays (cosine-distributed from
the hits of 1920 x 1080 primary rays), resident in HBM before timing.  A step
is one closest-hit launch over all 2^24 rays.

roofline: the HBM bytes one launch of the dominant kernel moves (FETCH_SIZE +
WRITE_SIZE from the committed rocprofv3 --pmc summary, each scaled by what it
reads for a 1 GiB copy) over this run's average launch time (HIP events on
the library stream), against the 8 TB/s HBM peak.  The traversal is bound by
VALU issue and latency rather than HBM; roofline.valu gives its VALU
wave-instructions per CU cycle over those of the calibration kernel
(gpugi_probe_valu) and the active lanes per VALU instruction, both from the
committed counter summary.  extra.alg_bytes_*: the reference walk's
algorithmic bytes per ray, B = 32 (ray) + 16 (hit) + 32 * N_box + 52 * N_tri
(N_box / N_tri: the escape walk's box and triangle tests of that ray, counted
on the device by the counting kernel the tests pin to the CPU oracle), over
the launch time: the wide kernel visits far fewer nodes than that walk, and
caches serve the rest, so that rate exceeds the HBM peak.
cpu_baseline: the CPU oracle's escape walk on every host core the run may
use (one thread per core, OMP_NUM_THREADS or the CPU count) on a bounded
sample of the same rays.
"""
import argparse
import itertools
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
    p.add_argument("--autotune", action="store_true",
                   help="autotune the wide kernel on a separate incoherent sample before timing "
                        "(default: the library's default configuration)")
    p.add_argument("--summarize", nargs=4, metavar=("KT_DIR", "FETCH_DIR", "WRITE_DIR", "KT_LOG"),
                   help="write the profiles/ summaries from rocprofv3 output directories and exit")
    p.add_argument("--summarize-counters", nargs="+", metavar="NAME_THEN_DIRS",
                   help="write profiles/<NAME>.json from rocprofv3 --pmc output directories and exit")
    p.add_argument("--summarize-integrators", nargs="+", metavar="PMC_DIR",
                   help="write profiles/integrators_clutter_counters.json from rocprofv3 --pmc runs of --clutter "
                        "and exit")
    p.add_argument("--summarize-kernels", nargs=2, metavar=("NAME", "KT_DIR"),
                   help="write profiles/NAME.txt: per-kernel totals of a kernel trace of bench.py with the "
                        "arguments given by --profiled-args=..., and exit")
    p.add_argument("--profiled-args", default="", help="the bench.py arguments of the traced run (--summarize-kernels)")
    p.add_argument("--summarize-wavefront", metavar="KT_DIR",
                   help="write profiles/wavefront_pt_clutter_kernel_trace.txt from a kernel trace of --clutter and exit")
    p.add_argument("--summarize-mlp", nargs="+", metavar="KT_DIR_THEN_PMC_DIRS",
                   help="write profiles/mlp_kernels.json from a kernel trace and --pmc runs of --mlp-kernels and exit")
    p.add_argument("--sweep", action="store_true",
                   help="print Grays/s of every wide-kernel configuration on the three ray sets and exit "
                        "(heightfield; with --clutter, the clutter room)")
    p.add_argument("--mlp", action="store_true",
                   help="time-to-equal-RMSE of the cached path tracer against the plain one (heightfield, 1920x1080)")
    p.add_argument("--mlp-ref-spp", type=int, default=4096, help="samples per pixel of the plain reference render")
    p.add_argument("--mlp-frames", type=int, default=256, help="largest frame count of the two progressive renders")
    p.add_argument("--mlp-warmup", type=int, default=32, help="cached frames drawn before accumulation starts")
    p.add_argument("--mlp-kernels", action="store_true",
                   help="time the cache's training step (16,384 samples) and inference (1920x1080 queries) "
                        "on device-resident samples and report the BF16 matrix-core roofline")
    p.add_argument("--mlp-kernel-steps", type=int, default=50, help="timed training steps / inference launches")
    p.add_argument("--hierarchy", action="store_true",
                   help="HierarchyImportance passes on the 2,097,152-triangle heightfield at 1920x1080: "
                        "importance acquisition + propagation, then node materials")
    p.add_argument("--hierarchy-iterations", type=int, default=10, help="acquisition passes per update")
    p.add_argument("--batch", action="store_true",
                   help="batched rendering of 16 scenes: per-scene launches vs persistent launch vs hipGraph replay")
    p.add_argument("--batch-frames", type=int, default=10, help="frames timed per batched mode")
    p.add_argument("--clutter", action="store_true",
                   help="5.12M-triangle clutter room: trace roofline and PT/LT/BPT iteration times")
    p.add_argument("--clutter-iterations", type=int, default=4, help="timed iterations per integrator")
    p.add_argument("--clutter-incoherent-only", action="store_true", help="only the incoherent trace (profiling)")
    p.add_argument("--calibrate", action="store_true",
                   help="append 3 device-to-device copies of 1 GiB (counter calibration for rocprofv3 --pmc)")
    p.add_argument("--valu-probe", action="store_true",
                   help="append 3 launches of the VALU calibration kernel (gpugi_probe_valu) for rocprofv3 --pmc")
    return p.parse_args()


def traffic_from_profile(name="trace_wide_heightfield_pmc.json"):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        return float(d["hbm_bytes_per_launch"] if "hbm_bytes_per_launch" in d else d["derived"]["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


PROFILE_WORKLOAD = "heightfield 1024x1024 cells (2,097,152 triangles), 2^24 incoherent rays, closest hit"
DOMINANT_KERNEL = "void gpugi::k_trace_wide<false"
PROBE_KERNEL = "gpugi::k_probe_valu"
PROFILE_COMMANDS = {
    "kt": "cd /tmp && export TMPDIR=/tmp && timeout -k 10 300 rocprofv3 --kernel-trace --stats -d <out>/prof_kt "
          "-o kt --output-format csv -- python3 <repo>/bench.py > <out>/prof_kt.log",
    "fetch": "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 150 rocprofv3 --pmc FETCH_SIZE -d <out>/prof_fetch "
             "-o f --output-format csv -- python3 <repo>/bench.py --steps 3 --warmup 1 --no-cpu-baseline --calibrate",
    "write": "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 150 rocprofv3 --pmc WRITE_SIZE -d <out>/prof_write "
             "-o w --output-format csv -- python3 <repo>/bench.py --steps 3 --warmup 1 --no-cpu-baseline --calibrate",
    "summary": "python3 <repo>/bench.py --summarize <out>/prof_kt <out>/prof_fetch <out>/prof_write <out>/prof_kt.log",
}


def _csv_rows(directory, suffix):
    import csv
    import glob
    files = sorted(glob.glob(os.path.join(directory, "*" + suffix)))
    if not files:
        raise SystemExit("bench.py --summarize: no *%s under %s" % (suffix, directory))
    with open(files[0], newline="") as f:
        return list(csv.DictReader(f))


def _pmc_per_launch(directory, counter):
    """(average counter value of the dominant kernel, average of the three 1 GiB calibration copies)."""
    rows = [r for r in _csv_rows(directory, "_counter_collection.csv") if r["Counter_Name"] == counter]
    kern = [float(r["Counter_Value"]) for r in rows if r["Kernel_Name"].startswith(DOMINANT_KERNEL)]
    copies = [float(r["Counter_Value"]) for r in rows if r["Kernel_Name"] == "__amd_rocclr_copyBuffer"]
    copies = sorted(copies)[-3:]
    if not kern or len(copies) < 3:
        raise SystemExit("bench.py --summarize: %s rows missing in %s" % (counter, directory))
    return sum(kern) / len(kern), sum(copies) / len(copies)


def summarize(kt_dir, fetch_dir, write_dir, kt_log):
    """Writes profiles/trace_wide_heightfield_{kernel_stats.txt,pmc.json} from rocprofv3 CSV output.

    HBM bytes per launch = FETCH_SIZE * ff + WRITE_SIZE * fw (KiB -> bytes), where ff and fw make the
    counters read 1 GiB each for a 1 GiB device-to-device copy (the --calibrate copies of the same runs).
    """
    stats = _csv_rows(kt_dir, "_kernel_stats.csv")
    stats.sort(key=lambda r: -float(r["TotalDurationNs"]))
    dom = [r for r in stats if r["Name"].startswith(DOMINANT_KERNEL)]
    if not dom:
        raise SystemExit("bench.py --summarize: dominant kernel not in the kernel statistics")
    bench_line = None
    with open(kt_log) as f:
        for line in f:
            if line.startswith("{"):
                bench_line = json.loads(line)
    lines = ["rocprofv3 kernel-trace statistics: traversal benchmark (%s), MI355X (gfx950), one GPU." % PROFILE_WORKLOAD,
             "",
             "command: " + PROFILE_COMMANDS["kt"],
             "summary: " + PROFILE_COMMANDS["summary"]]
    if bench_line:
        lines.append("bench.py printed (same process): ms_per_step %.4f (HIP events over %d launches), value %.4f %s, "
                     "wide_config %s" % (bench_line["ms_per_step"], bench_line["steps"], bench_line["value"],
                                         bench_line["unit"], json.dumps(bench_line["extra"].get("wide_config"))))
    lines += ["", "%-72s %6s %14s %14s %14s" % ("kernel", "calls", "avg_ns", "min_ns", "max_ns")]
    for r in stats[:14]:
        lines.append("%-72s %6s %14d %14d %14d" % (r["Name"][:72], r["Calls"], round(float(r["AverageNs"])),
                                                   int(float(r["MinNs"])), int(float(r["MaxNs"]))))
    with open(os.path.join(ROOT, "profiles", "trace_wide_heightfield_kernel_stats.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")

    fk, fc = _pmc_per_launch(fetch_dir, "FETCH_SIZE")
    wk, wc = _pmc_per_launch(write_dir, "WRITE_SIZE")
    gib_kib = float(1 << 20)
    ff, fw = gib_kib / fc, gib_kib / wc
    hbm = (fk * ff + wk * fw) * 1024.0
    avg_ns = float(dom[0]["AverageNs"])
    out = {"kernel": dom[0]["Name"], "workload": PROFILE_WORKLOAD,
           "commands": [PROFILE_COMMANDS["fetch"], PROFILE_COMMANDS["write"], PROFILE_COMMANDS["summary"]],
           "calibration": {"copy_bytes": 1 << 30, "fetch_size_kib_read": fc, "write_size_kib_read": wc,
                           "fetch_factor": ff, "write_factor": fw},
           "fetch_size_kib_per_launch": fk, "write_size_kib_per_launch": wk,
           "hbm_bytes_per_launch": hbm, "launch_ns_avg_kernel_trace": avg_ns,
           "hbm_gbs": hbm / avg_ns}
    with open(os.path.join(ROOT, "profiles", "trace_wide_heightfield_pmc.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out))


def _has_counter(directory, counter):
    return any(r["Counter_Name"] == counter for r in _csv_rows(directory, "_counter_collection.csv"))


def summarize_counters(name, *dirs):
    """Writes profiles/<name>.json: per-launch averages of every counter that rocprofv3 --pmc collected for
    the dominant kernel in the given output directories, plus derived ratios when their inputs are there."""
    def per_kernel(prefix):
        per = {}
        for d in dirs:
            acc = {}
            for r in _csv_rows(d, "_counter_collection.csv"):
                if r["Kernel_Name"].startswith(prefix):
                    acc.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
            for k, v in acc.items():
                per[k] = sum(v) / len(v)
        return per

    per = per_kernel(DOMINANT_KERNEL)
    probe = per_kernel(PROBE_KERNEL)
    rays = float(1 << 24)
    derived = {}
    if "TCC_HIT_sum" in per and "TCC_MISS_sum" in per:
        derived["tcc_hit_rate"] = per["TCC_HIT_sum"] / (per["TCC_HIT_sum"] + per["TCC_MISS_sum"])
    if "SQ_INSTS_VALU" in per:
        derived["valu_wave_instructions_per_ray"] = per["SQ_INSTS_VALU"] / rays
    if "SQ_INSTS_VMEM_RD" in per:
        derived["vmem_read_wave_instructions_per_ray"] = per["SQ_INSTS_VMEM_RD"] / rays

    def per_cu_cycle(c):
        # GRBM_GUI_ACTIVE counts the launch's GPU-busy cycles summed over the 8 XCDs;
        # 256 CUs: VALU wave-instructions issued per CU per cycle
        return c["SQ_INSTS_VALU"] / (c["GRBM_GUI_ACTIVE"] / 8.0 * 256.0)

    if "SQ_INSTS_VALU" in per and "GRBM_GUI_ACTIVE" in per:
        derived["valu_wave_insts_per_cu_cycle"] = per_cu_cycle(per)
    if "SQ_INSTS_VALU" in probe and "GRBM_GUI_ACTIVE" in probe:
        derived["probe_valu_wave_insts_per_cu_cycle"] = per_cu_cycle(probe)
        if "valu_wave_insts_per_cu_cycle" in derived:
            derived["valu_issue_frac_of_probe"] = (derived["valu_wave_insts_per_cu_cycle"] /
                                                   derived["probe_valu_wave_insts_per_cu_cycle"])
    # vector-memory pipeline: one TA, TD and TCP (L1) per CU, 256 CUs
    cu_cycles = per["GRBM_GUI_ACTIVE"] / 8.0 * 256.0 if "GRBM_GUI_ACTIVE" in per else None
    if cu_cycles:
        for key, label in (("TA_TA_BUSY_sum", "ta_busy_frac"), ("TD_TD_BUSY_sum", "td_busy_frac"),
                          ("TD_TC_STALL_sum", "td_stalled_by_l1_frac"),
                          ("TCP_PENDING_STALL_CYCLES_sum", "l1_pending_stall_frac"),
                          ("TCP_READ_TAGCONFLICT_STALL_CYCLES_sum", "l1_tag_conflict_stall_frac"),
                          ("TCP_TOTAL_CACHE_ACCESSES_sum", "l1_accesses_per_cu_cycle")):
            if key in per:
                derived[label] = per[key] / cu_cycles
    if "TCP_TOTAL_CACHE_ACCESSES_sum" in per and "TCP_TCC_READ_REQ_sum" in per:
        derived["l1_hit_rate"] = 1.0 - per["TCP_TCC_READ_REQ_sum"] / per["TCP_TOTAL_CACHE_ACCESSES_sum"]
    if "SQ_WAIT_INST_ANY" in per and "SQ_WAVE_CYCLES" in per:
        derived["wave_wait_frac"] = per["SQ_WAIT_INST_ANY"] / per["SQ_WAVE_CYCLES"]
    if "TCP_TOTAL_CACHE_ACCESSES_sum" in per and "SQ_INSTS_VMEM_RD" in per:
        derived["l1_accesses_per_vmem_read_instruction"] = per["TCP_TOTAL_CACHE_ACCESSES_sum"] / per["SQ_INSTS_VMEM_RD"]
    if "SQ_THREAD_CYCLES_VALU" in per and "SQ_ACTIVE_INST_VALU" in per:
        derived["valu_thread_cycles_per_active_inst"] = per["SQ_THREAD_CYCLES_VALU"] / per["SQ_ACTIVE_INST_VALU"]
        if "SQ_THREAD_CYCLES_VALU" in probe and "SQ_ACTIVE_INST_VALU" in probe:
            # the calibration kernel runs all 64 lanes of every instruction
            full = probe["SQ_THREAD_CYCLES_VALU"] / probe["SQ_ACTIVE_INST_VALU"]
            derived["probe_thread_cycles_per_active_inst"] = full
            derived["valu_active_lanes"] = 64.0 * derived["valu_thread_cycles_per_active_inst"] / full
    clutter = "clutter" in name
    workload = ("clutter room, 4000 icospheres (5,120,016 triangles), 2^24 incoherent rays, closest hit"
                if clutter else PROFILE_WORKLOAD)
    run = ("--clutter --clutter-incoherent-only --steps 3 --calibrate --valu-probe" if clutter
           else "--steps 2 --warmup 1 --no-cpu-baseline --calibrate --valu-probe")
    if "FETCH_SIZE" in per and "WRITE_SIZE" in per:
        gib_kib = float(1 << 20)
        fc = _pmc_per_launch([d for d in dirs if _has_counter(d, "FETCH_SIZE")][0], "FETCH_SIZE")[1]
        wc = _pmc_per_launch([d for d in dirs if _has_counter(d, "WRITE_SIZE")][0], "WRITE_SIZE")[1]
        derived["calibration"] = {"fetch_factor": gib_kib / fc, "write_factor": gib_kib / wc}
        derived["hbm_bytes_per_launch"] = (per["FETCH_SIZE"] * gib_kib / fc + per["WRITE_SIZE"] * gib_kib / wc) * 1024.0
    out = {"kernel_prefix": DOMINANT_KERNEL, "probe_kernel_prefix": PROBE_KERNEL, "workload": workload,
           "commands": ["cd /tmp && export TMPDIR=/tmp && timeout -s KILL 120 rocprofv3 --pmc <counters of one pass> "
                        "-d <out>/<pass> -o <pass> --output-format csv -- python3 <repo>/bench.py " + run,
                        "python3 <repo>/bench.py --summarize-counters %s %s" % (name, " ".join(
                            "<out>/" + os.path.basename(os.path.normpath(d)) for d in dirs))],
           "per_launch": per, "probe_per_launch": probe, "derived": derived}
    with open(os.path.join(ROOT, "profiles", name + ".json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out))


INTEGRATOR_KERNELS = ("k_pathtracer<true>", "k_lighttracer<true, 0>", "k_lighttracer<true, 1>", "k_bpt_eye<true>",
                      "k_wf_shade<0>", "k_wf_shade<1>", "k_trace_wide<false, 256, 10, false>",
                      "k_trace_wide<true, 256, 10, false>")


def summarize_integrators(*dirs):
    """Writes profiles/integrators_clutter_counters.json: per-dispatch averages of the counters that
    rocprofv3 --pmc runs of `bench.py --clutter` collected for each integrator kernel (the one-kernel
    PT, LT, BPT light pass and eye pass, the wavefront shade passes and the trace kernels they launch),
    with the utilisation figures derived from them."""
    per = {k: {} for k in INTEGRATOR_KERNELS}
    calls = {k: 0 for k in INTEGRATOR_KERNELS}
    for d in dirs:
        acc = {k: {} for k in INTEGRATOR_KERNELS}
        for r in _csv_rows(d, "_counter_collection.csv"):
            k = _short_kernel(r["Kernel_Name"])
            if k in acc:
                acc[k].setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
        for k, counters in acc.items():
            for c, v in counters.items():
                per[k][c] = sum(v) / len(v)
                calls[k] = max(calls[k], len(v))
    out = {}
    for k in INTEGRATOR_KERNELS:
        c = per[k]
        if not c:
            continue
        dv = {}
        if "GRBM_GUI_ACTIVE" in c:
            cu_cycles = c["GRBM_GUI_ACTIVE"] / 8.0 * 256.0  # summed over the 8 XCDs; 256 CUs
            if "SQ_INSTS_VALU" in c:
                dv["valu_wave_insts_per_cu_cycle"] = c["SQ_INSTS_VALU"] / cu_cycles
            for key, label in (("TA_TA_BUSY_sum", "ta_busy_frac"), ("TD_TD_BUSY_sum", "td_busy_frac")):
                if key in c:
                    dv[label] = c[key] / cu_cycles
        if "SQ_THREAD_CYCLES_VALU" in c and "SQ_ACTIVE_INST_VALU" in c:
            # the VALU calibration kernel (all 64 lanes active) reads 64.0 here
            dv["valu_active_lanes"] = c["SQ_THREAD_CYCLES_VALU"] / c["SQ_ACTIVE_INST_VALU"]
        if "SQ_WAIT_INST_ANY" in c and "SQ_WAVE_CYCLES" in c:
            dv["wave_wait_frac"] = c["SQ_WAIT_INST_ANY"] / c["SQ_WAVE_CYCLES"]
        if "TCC_HIT_sum" in c and "TCC_MISS_sum" in c:
            dv["l2_hit_rate"] = c["TCC_HIT_sum"] / (c["TCC_HIT_sum"] + c["TCC_MISS_sum"])
        if "TCP_TOTAL_CACHE_ACCESSES_sum" in c and "TCP_TCC_READ_REQ_sum" in c:
            dv["l1_hit_rate"] = 1.0 - c["TCP_TCC_READ_REQ_sum"] / c["TCP_TOTAL_CACHE_ACCESSES_sum"]
        out[k] = {"dispatches": calls[k], "per_dispatch": c, "derived": dv}
    doc = {"workload": "clutter room (5,120,016 triangles), 1920x1080, path length 8, roulette from bounce 3, "
                       "wide traversal; per dispatch (the wavefront kernels and the trace kernels average over "
                       "every bounce's launch)",
           "commands": ["cd /tmp && export TMPDIR=/tmp && timeout -s KILL 300 rocprofv3 --pmc <counters of one pass> "
                        "-d <out>/<pass> -o <pass> --output-format csv -- python3 <repo>/bench.py --clutter "
                        "--clutter-iterations 1 --steps 1",
                        "python3 <repo>/bench.py --summarize-integrators " + " ".join(
                            "<out>/" + os.path.basename(os.path.normpath(d)) for d in dirs)],
           "kernels": out}
    with open(os.path.join(ROOT, "profiles", "integrators_clutter_counters.json"), "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    for k, v in out.items():
        print(k, json.dumps({a: round(b, 3) for a, b in v["derived"].items()}))


def batch_benchmark(frames):
    """16 heightfield scenes (10k..2M triangles, 256^2..1024^2 pixels) rendered by the
    path tracer: per-scene launches vs one persistent launch vs its hipGraph replay.
    Aggregate Grays/s = ray queries (closest + any hit) / GPU time."""
    from gpugi_amd import rt, workloads
    t0 = time.perf_counter()
    scenes = workloads.batch_scenes()
    b = rt.Batch(0)
    for hs, cam, w, h, lights in scenes:
        b.add(hs, cam, w, h, lights, workloads.BATCH_PATH_LENGTH)
    b.finalize()
    setup_s = time.perf_counter() - t0
    tuned = b.autotune(2)
    # the persistent wide frame at each LDS stack depth, the rest as tuned
    by_depth = {}
    for depth in (16, 24, 32):
        b.config(rt.KERNEL_WIDE, tuned["tiles_per_grab"], tuned["grid_multiplier"], depth)
        b.reset(0)
        b.run(1, 1)
        b.reset(0)
        by_depth[depth] = round(b.run(1, frames) / frames, 3)
    b.config(tuned["kernel"], tuned["tiles_per_grab"], tuned["grid_multiplier"], tuned["stack_depth"])
    out = {}
    for name, mode in (("separate", 0), ("persistent", 1), ("graph", 2)):
        b.reset(0)
        b.run(mode, 1)  # warm-up (and capture for the graph)
        b.reset(0)
        ms = b.run(mode, frames)
        rays = b.rays()
        out[name] = {"ms_per_frame": round(ms / frames, 3), "grays": round(rays / (ms * 1e-3) / 1e9, 4),
                     "rays_per_frame": rays // frames}
    b.close()
    return {
        "metric": "batched path tracing Grays/s (16 scenes, one persistent launch)",
        "value": out["graph"]["grays"], "unit": "Grays/s", "higher_is_better": True,
        "speedup_vs_separate": round(out["separate"]["ms_per_frame"] / out["graph"]["ms_per_frame"], 3),
        "modes": out, "autotuned": tuned, "wide_ms_per_frame_by_stack_depth": by_depth, "frames": frames,
        "config": {"workload": "16 heightfields, triangles log-uniform in [10k, 2M], square 256..1024 px, "
                               "4 point lights, path length %d" % workloads.BATCH_PATH_LENGTH,
                   "triangles": [int(s[0].info()["num_triangles"]) for s in scenes],
                   "pixels": [int(s[2] * s[3]) for s in scenes]},
        "setup_s": round(setup_s, 2),
    }


def hierarchy_benchmark(iterations, repeats=3):
    """HierarchyImportance::Draw's importance update (acquisition2.comp x `iterations`, one path
    of up to 8 vertices per pixel each, then the propagation passes) and
    ComputeHierarchyMaterials on the heightfield; GPU time by HIP events, best of `repeats`."""
    from gpugi_amd import rt, workloads
    W, H = workloads.HF_WIDTH, workloads.HF_HEIGHT
    hs = rt.HostScene.heightfield(workloads.HF_CELLS, workloads.MASTER_SEED)
    for pos in workloads.heightfield_lights():
        hs.add_point_light(pos, (0.5, 0.5, 0.5))
    ctx = rt.Context(0)
    ctx.load_scene(hs)
    ctx.set_screen(W, H)
    ctx.set_camera(workloads.heightfield_camera(W, H))
    ctx.set_light_samples(rt.LightSampler(hs, 123).generate(128))
    d = hs.desc()

    def timed(fn):
        fn()  # warm-up (allocates the buffers)
        best = None
        for _ in range(repeats):
            ctx.sync()
            ctx.timer_begin()
            fn()
            ms = ctx.timer_end()
            best = ms if best is None else min(best, ms)
        return best

    upd = timed(lambda: ctx.hierarchy_update_importance(iterations))
    mat = timed(ctx.hierarchy_compute_materials)
    imp, sub = ctx.hierarchy_read()
    ctx.close()
    paths = iterations * (W // 8 * 8) * (H // 8 * 8)
    return {"metric": "HierarchyImportance importance update, Mpaths/s", "value": round(paths / (upd * 1e3), 2),
            "unit": "Mpaths/s", "higher_is_better": True,
            "update_ms": round(upd, 3), "acquisition_iterations": iterations, "acquisition_paths": paths,
            "materials_ms": round(mat, 3), "nodes": int(d.num_nodes), "triangle_slots": int(d.num_leaves * 2),
            "root_importance": float(imp[0, 0]), "root_subtree": float(sub[0]),
            "config": {"workload": "heightfield 1024x1024 cells (2,097,152 triangles), 4 point lights, 128 light "
                                   "samples, 1920x1080, reference escape walk"}}


def run_valu_probe(ctx):
    """Three launches of the VALU calibration kernel (256 CUs x 32 blocks of 256 lanes,
    64 x 2048 FMAs per lane): the counter run reads its VALU issue rate."""
    blocks, iters = 256 * 32, 2048
    out = ctx.alloc(blocks * 256 * 4)
    probe = [ctx.probe_valu(out, blocks, iters) for _ in range(3)]
    out.free()
    return {"blocks": blocks, "iters": iters, "ms": [round(x, 4) for x in probe],
            "fma_tflops": round(blocks * 256 * iters * 64 * 2 / (min(probe) * 1e-3) / 1e12, 2)}


def sweep_wide_configs(ctx, R, n, grid):
    """Prints one JSON line per wide-kernel configuration (max_leaf_triangles, lds_stack_entries,
    refill_lanes, wait_lanes) of `grid`: Grays/s of the three ray sets of R, 1 warm-up + 3 timed launches."""
    from gpugi_amd import rt
    sets = (("incoherent", R.incoherent, rt.CLOSEST_HIT), ("coherent", R.primary, rt.CLOSEST_HIT),
            ("shadow_any", R.shadow, rt.ANY_HIT))
    for leaf, stack, refill, wait in grid:
        ctx.set_wide_config(leaf, stack, refill, wait)
        row = {}
        for name, buf, q in sets:
            ctx.trace_device(buf, R.hits, n, q, rt.KERNEL_WIDE)
            ctx.sync()
            ctx.timer_begin()
            for _ in range(3):
                ctx.trace_device(buf, R.hits, n, q, rt.KERNEL_WIDE)
            row[name] = round(n / (ctx.timer_end() / 3 * 1e-3) / 1e9, 3)
        print(json.dumps(dict(ctx.wide_config(), **row)), flush=True)


def clutter_sweep():
    """--clutter --sweep: the wide-kernel configurations on the clutter room's three 2^24-ray sets."""
    from gpugi_amd import rt, workloads
    hs = rt.HostScene.clutter(workloads.CLUTTER_SPHERES, workloads.MASTER_SEED)
    ctx = rt.Context(0)
    ctx.load_scene(hs)
    n = 1 << 24
    R = workloads.SceneRays(ctx, workloads.clutter_camera(workloads.HF_WIDTH, workloads.HF_HEIGHT),
                            workloads.CLUTTER_LIGHTS, n, workloads.MASTER_SEED, workloads.HF_WIDTH,
                            workloads.HF_HEIGHT)
    sweep_wide_configs(ctx, R, n, itertools.product((2, 4), (8, 10, 12, 16, 24), (8, 16), (16, 24)))
    R.free()
    ctx.close()


def clutter_benchmark(steps, iterations, incoherent_only=False, calibrate=False, valu_probe=False):
    """The 5,120,016-triangle clutter room: incoherent closest-hit Grays/s with its
    roofline (its BVH and triangles exceed the 256 MB MALL), then PT, LT and BPT
    iterations at 1920 x 1080, path length 8, roulette from bounce 3.
    incoherent_only: only the incoherent trace (the counter-profiling run)."""
    from gpugi_amd import rt, workloads
    t0 = time.perf_counter()
    hs = rt.HostScene.clutter(workloads.CLUTTER_SPHERES, workloads.MASTER_SEED)
    ctx = rt.Context(0)
    ctx.load_scene(hs)
    W, H = workloads.HF_WIDTH, workloads.HF_HEIGHT
    n = 1 << 24
    R = workloads.SceneRays(ctx, workloads.clutter_camera(W, H), workloads.CLUTTER_LIGHTS, n,
                            workloads.MASTER_SEED, W, H)
    nbox_b, ntri_b = ctx.alloc(n * 4), ctx.alloc(n * 4)
    ctx.trace_counts_device(R.incoherent, nbox_b, ntri_b, n, rt.CLOSEST_HIT)
    nbox = nbox_b.download(np.uint32, n).astype(np.float64)
    ntri = ntri_b.download(np.uint32, n).astype(np.float64)
    nbox_b.free()
    ntri_b.free()
    alg_bytes = float(n * (32 + 16) + 32.0 * nbox.sum() + 52.0 * ntri.sum())
    setup_s = time.perf_counter() - t0
    trace = {}
    sets = (("incoherent", R.incoherent, rt.CLOSEST_HIT), ("coherent", R.primary, rt.CLOSEST_HIT),
            ("shadow_any", R.shadow, rt.ANY_HIT))
    for name, buf, q in sets[:1] if incoherent_only else sets:
        ctx.trace_device(buf, R.hits, n, q, rt.KERNEL_WIDE)
        ctx.sync()
        ctx.timer_begin()
        for _ in range(steps):
            ctx.trace_device(buf, R.hits, n, q, rt.KERNEL_WIDE)
        ms = ctx.timer_end() / steps
        trace[name] = {"ms": round(ms, 4), "grays": round(n / (ms * 1e-3) / 1e9, 4)}
    hits = R.hits.download(np.int32, n * 4).reshape(n, 4)
    R.free()
    ms_inc = trace["incoherent"]["ms"]
    # measured HBM bytes of one launch (committed counter summary) over this run's launch time; the
    # reference walk's algorithmic bytes ride along under their own names
    alg_gbs = alg_bytes / (ms_inc * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": None,
            "alg_bytes_per_ray": alg_bytes / n, "alg_bytes_gbs": round(alg_gbs, 2),
            "alg_bytes_over_hbm_peak": round(alg_gbs / HBM_PEAK_GBS, 4), "mean_box_tests": float(nbox.mean()),
            "mean_tri_tests": float(ntri.mean())}
    traffic = traffic_from_profile("trace_wide_clutter_pmc.json")
    if traffic:
        hbm_gbs = traffic / (ms_inc * 1e-3) / 1e9
        roof.update(traffic=traffic, achieved=round(hbm_gbs, 2), frac=round(hbm_gbs / HBM_PEAK_GBS, 4))
    valu = valu_from_profile("trace_wide_clutter_pmc.json")
    if valu:
        roof["valu"] = valu
    if incoherent_only:
        if calibrate:  # 3 device-to-device copies of 1 GiB (counter calibration)
            a, b = ctx.alloc(1 << 30), ctx.alloc(1 << 30)
            for _ in range(3):
                b.copy_from(a, 1 << 30)
            ctx.sync()
            a.free()
            b.free()
        probe = run_valu_probe(ctx) if valu_probe else None
        ctx.close()
        return {"metric": "incoherent closest-hit Grays/s (clutter room)", "value": trace["incoherent"]["grays"],
                "unit": "Grays/s", "trace": trace, "roofline": roof, "valu_probe": probe}

    ctx.set_screen(W, H)
    ctx.set_camera(workloads.clutter_camera(W, H))
    ctx.set_max_path_length(workloads.CLUTTER_PATH_LENGTH)
    ctx.set_rr_start_depth(workloads.CLUTTER_RR_START)
    ctx.set_integrator_kernel(rt.KERNEL_WIDE)
    sampler = rt.LightSampler(hs, 123)
    integ = {}
    for name, integrator, nls in (("pt", rt.PT, 128), ("pt_wavefront", rt.PT, 128), ("lt", rt.LT, 256),
                                  ("bpt", rt.BPT, 128), ("bpt_wavefront", rt.BPT, 128)):
        # PT and BPT's eye paths under both schedules: one kernel, and per-bounce passes through the
        # trace kernels
        ctx.set_pt_schedule(rt.PT_WAVEFRONT if name.endswith("_wavefront") else rt.PT_MEGAKERNEL)
        sampler = rt.LightSampler(hs, 123)  # every integrator sees the same light-sample batches
        if integrator == rt.BPT:
            entries = 0
            for k in range(8):  # capacity estimate of bidirectionalpathtracer.cpp:65-112
                ctx.set_light_samples(sampler.generate(nls))
                entries += ctx.bpt_warmup(rt.wanghash(k))[0] & 0xffffffff
            mean = entries // 8
            ctx.bpt_set_light_cache(mean + mean // 10)
        ctx.reset_accum()
        ctx.set_light_samples(sampler.generate(nls))
        ctx.draw(integrator, rt.wanghash(0))
        ctx.sync()
        ms = 0.0
        for k in range(iterations):
            ctx.set_light_samples(sampler.generate(nls))
            ctx.timer_begin()
            ctx.draw(integrator, rt.wanghash(k + 1))
            ms += ctx.timer_end()
        img, _ = ctx.readback(W, H)
        integ[name] = {"ms_per_iteration": round(ms / iterations, 3),
                       "image_mean": float(img[..., :3].mean() / (iterations + 1))}
        if name == "pt_wavefront":
            # the same frame recorded into a HIP graph and replayed: no host launch per pass
            ctx.capture_draw(integrator, rt.wanghash(iterations + 1))
            ctx.replay_draw()
            ctx.sync()
            ms = 0.0
            for k in range(iterations):
                ctx.timer_begin()
                ctx.replay_draw()
                ms += ctx.timer_end()
            integ[name]["ms_per_iteration_graph_replay"] = round(ms / iterations, 3)
    ctx.close()
    info = hs.info()
    return {
        "metric": "incoherent closest-hit Grays/s (clutter room)", "value": trace["incoherent"]["grays"],
        "unit": "Grays/s", "higher_is_better": True,
        "config": {"workload": "clutter room, %d icospheres, %d triangles, 1920x1080, 2^24 rays per set"
                               % (workloads.CLUTTER_SPHERES, info["num_triangles"]),
                   "integrators": "path length %d, roulette from bounce %d, wide traversal"
                                  % (workloads.CLUTTER_PATH_LENGTH, workloads.CLUTTER_RR_START)},
        "trace": trace, "roofline": roof, "hit_fraction": float((hits[:, 0] >= 0).mean()),
        "integrators": integ, "setup_s": round(setup_s, 2),
    }


MLP_PATH_LENGTH = 8
MLP_LIGHT_INTENSITY = 0.5


def mlp_benchmark(ref_spp, max_frames, warmup=32):
    """Plain vs cached path tracing on the heightfield at 1920 x 1080.

    Both render progressively from frame 0; the cache is trained online from
    scratch and its first `warmup` frames are drawn but not accumulated (their
    time is charged to the cached render).  RMSE is taken against a plain render of
    `ref_spp` samples per pixel drawn with other frame seeds.  For each cached
    frame count F the plain sample count N with the same RMSE is interpolated
    (log-log) and the GPU-time ratio N * t_plain / (F * t_cached) reported.
    """
    from gpugi_amd import rt, workloads
    W, H = workloads.HF_WIDTH, workloads.HF_HEIGHT
    hs = rt.HostScene.heightfield(workloads.HF_CELLS, workloads.MASTER_SEED)
    for pos in workloads.heightfield_lights():
        hs.add_point_light(pos, (MLP_LIGHT_INTENSITY,) * 3)
    ctx = rt.Context(0)
    ctx.load_scene(hs)
    ctx.set_screen(W, H)
    ctx.set_camera(workloads.heightfield_camera(W, H))
    ctx.set_max_path_length(MLP_PATH_LENGTH)
    ctx.set_integrator_kernel(rt.KERNEL_WIDE)
    sampler = rt.LightSampler(hs, 123)
    # light-sample batches (128 per frame, as the path tracer uses) cycled through a fixed pool,
    # so frame k of every render sees the same batch
    pool = [sampler.generate(128) for _ in range(64)]

    def run(integrator, frames, first, checkpoints, warmup=0):
        """Frames `warmup`.. are accumulated; returns images at the checkpoints and ms per frame."""
        ctx.reset_accum()
        out, ms = {}, 0.0
        for k in range(warmup + frames):
            if k == warmup:
                ctx.reset_accum()
            ctx.set_light_samples(pool[(first + k) % len(pool)])
            ctx.sync()
            ctx.timer_begin()
            ctx.draw(integrator, rt.wanghash(first + k))
            ms += ctx.timer_end()
            f = k + 1 - warmup
            if f in checkpoints:
                img, _ = ctx.readback(W, H)
                out[f] = (img[..., :3] / np.float32(f)).copy()
        return out, ms / (warmup + frames)

    t0 = time.perf_counter()
    ref, ref_ms = run(rt.PT, ref_spp, 1 << 20, {ref_spp})
    ref = ref[ref_spp].astype(np.float64)
    checks = sorted({1 << i for i in range(0, 16) if (1 << i) <= max_frames} | {max_frames})
    plain, plain_ms = run(rt.PT, max_frames, 0, set(checks))

    def rmse(img):
        return float(np.sqrt(np.mean((img.astype(np.float64) - ref) ** 2)))

    rp = {n: rmse(plain[n]) for n in checks}
    ln = np.log([float(n) for n in checks])
    lp = np.log([rp[n] for n in checks])
    modes = {}
    for name, query in (("spread", rt.CACHE_SPREAD), ("primary_indirect", rt.CACHE_PRIMARY_INDIRECT)):
        ctx.mlp_init(1)
        ctx.cache_config()
        ctx.cache_set_query(query)
        cached, cached_ms = run(rt.PT_CACHE, max_frames, 0, set(checks), warmup)
        samples, loss = ctx.cache_stats()
        rc = {n: rmse(cached[n]) for n in checks}
        rows = []
        for f in checks:
            target = rc[f]
            # plain RMSE falls monotonically with spp; interpolate log(spp) at log(target)
            if target >= rp[checks[0]]:
                n_eq = float(checks[0]) * (rp[checks[0]] / target) ** 2
            elif target <= rp[checks[-1]]:
                n_eq = float(checks[-1]) * (rp[checks[-1]] / target) ** 2  # 1/sqrt(N) beyond the last point
            else:
                n_eq = float(np.exp(np.interp(-np.log(target), -lp, ln)))
            rows.append({"frames": f, "rmse_cached": round(rc[f], 6), "plain_spp_equal_rmse": round(n_eq, 2),
                         "gpu_time_ratio": round(n_eq * plain_ms / ((warmup + f) * cached_ms), 3)})
        best = max(rows, key=lambda r: r["gpu_time_ratio"])
        modes[name] = {"best_gpu_time_ratio": best["gpu_time_ratio"], "at_frames": best["frames"],
                       "ms_per_frame_cached": round(cached_ms, 3), "rows": rows, "last_train_samples": samples,
                       "last_loss": loss}
    ctx.cache_set_query(rt.CACHE_SPREAD)
    ctx.close()
    top = max(modes, key=lambda k: modes[k]["best_gpu_time_ratio"])
    return {
        "metric": "time-to-equal-RMSE, MLP cache vs plain PT (GPU-time ratio, higher is better)",
        "value": modes[top]["best_gpu_time_ratio"], "unit": "x", "higher_is_better": True,
        "config": {"workload": "heightfield 1024x1024 cells, 4 point lights, 1920x1080, path length %d, "
                               "wide traversal" % MLP_PATH_LENGTH,
                   "reference_spp": ref_spp, "frames": max_frames, "query": top,
                   "cache": "96-64x4-3 BF16 MLP, spread c 0.01, 1 training path per 32 pixels, 65536 samples "
                            "x 4 Adam steps per frame, lr 1e-3"},
        "ms_per_frame_plain": round(plain_ms, 3), "reference_ms_per_spp": round(ref_ms, 3), "warmup_frames": warmup,
        "rmse_plain": {str(n): round(rp[n], 6) for n in checks},
        "modes": modes, "wall_s": round(time.perf_counter() - t0, 1),
    }


BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix-core peak
# matrix-core FLOPs per sample as the kernels issue them (outputs padded to 16 in the
# forward pass and to 32 in the backward dX of the output layer):
# forward 2 * 19456; dW 2 * 19456; dX of layers 4..1: 2 * (64 * 32 + 3 * 64 * 64)
MLP_FWD_FLOP = 2 * 19456
MLP_TRAIN_FLOP = MLP_FWD_FLOP + 2 * 19456 + 2 * (64 * 32 + 3 * 64 * 64)


def mlp_kernel_benchmark(steps):
    """The cache's training step (gradient, its reduction and Adam over 16,384
    samples, the per-step batch of the cached path tracer) and inference over
    1920 x 1080 queries, on device-resident random features; HIP events over
    `steps` launches of each.  Matrix-core roofline from the FLOPs the kernels
    issue per sample."""
    import ctypes
    from gpugi_amd import rt
    rng = np.random.default_rng(5)

    def features(n):
        f = np.zeros((n, 16), np.float32)
        f[:, 0:3] = rng.random((n, 3))
        for a in (3, 6):
            d = rng.normal(size=(n, 3))
            f[:, a:a + 3] = d / np.linalg.norm(d, axis=1, keepdims=True)
        f[:, 9:12] = rng.random((n, 3))
        f[:, 12] = rng.random(n)
        return f

    n_train, n_infer = 16384, 1920 * 1080
    ctx = rt.Context(0)
    ctx.mlp_init(1)
    ft, fi = features(n_train), features(n_infer)
    tt = np.zeros((n_train, 4), np.float32)
    tt[:, :3] = 0.5 + 0.4 * np.sin(2 * np.pi * ft[:, :3])
    bft, btt, bfi, bout = ctx.alloc(ft.nbytes), ctx.alloc(tt.nbytes), ctx.alloc(fi.nbytes), ctx.alloc(n_infer * 16)
    bft.upload(ft)
    btt.upload(tt)
    bfi.upload(fi)
    vp = ctypes.c_void_p

    def train():
        ctx._call("gpugi_mlp_train_step", vp(bft.ptr), vp(btt.ptr), ctypes.c_size_t(n_train), ctypes.c_float(1e-3),
                  None, 1)

    def infer():
        ctx._call("gpugi_mlp_infer", vp(bfi.ptr), vp(bout.ptr), ctypes.c_size_t(n_infer), 1)

    out = {}
    for name, fn, n, flop in (("train_step", train, n_train, MLP_TRAIN_FLOP), ("inference", infer, n_infer,
                                                                                 MLP_FWD_FLOP)):
        for _ in range(3):
            fn()
        ctx.sync()
        ctx.timer_begin()
        for _ in range(steps):
            fn()
        ms = ctx.timer_end() / steps
        tflops = n * flop / (ms * 1e-3) / 1e12
        out[name] = {"samples": n, "ms": round(ms, 4), "flop_per_sample": flop, "tflops": round(tflops, 2),
                     "frac_bf16_peak": round(tflops / BF16_PEAK_TFLOPS, 4)}
    ctx.close()
    return {"metric": "radiance-cache MLP kernels, BF16 matrix cores", "unit": "TFLOP/s",
            "value": out["inference"]["tflops"], "higher_is_better": True, "peak_tflops": BF16_PEAK_TFLOPS,
            "steps": steps, "kernels": out,
            "config": {"workload": "96-64x4-3 MLP; training step = gradient + reduction + Adam over 16,384 "
                                   "samples; inference over 1920x1080 queries; random features on the device"}}


def _short_kernel(name):
    """Kernel name without namespaces, parameter list and return type."""
    for ns in ("(anonymous namespace)::", "gpugi::"):
        name = name.replace(ns, "")
    name = name.split("(")[0]
    if name.startswith("void "):
        name = name[5:]
    return name


def summarize_kernels(name, kt_dir, bench_args):
    """Writes profiles/<name>.txt: every kernel of a `rocprofv3 --kernel-trace --stats` run of
    `bench.py <bench_args>` with its calls, total and average duration."""
    stats = _csv_rows(kt_dir, "_kernel_stats.csv")
    stats.sort(key=lambda r: -float(r["TotalDurationNs"]))
    lines = ["Kernel trace of bench.py %s on one MI355X." % bench_args, "",
             "Commands (repo root; the first after cd /tmp && export TMPDIR=/tmp):",
             "  rocprofv3 --kernel-trace --stats -d <out>/kt -o kt --output-format csv -- python3 <repo>/bench.py "
             + bench_args,
             "  python3 bench.py --summarize-kernels %s <out>/kt --profiled-args='%s'" % (name, bench_args), "",
             "%-60s %6s %12s %12s %12s" % ("kernel", "calls", "total ms", "avg us", "max us")]
    for r in stats:
        lines.append("%-60s %6d %12.3f %12.1f %12.1f" % (_short_kernel(r["Name"])[:60], int(r["Calls"]),
                                                         float(r["TotalDurationNs"]) / 1e6,
                                                         float(r["AverageNs"]) / 1e3, float(r["MaxNs"]) / 1e3))
    text = "\n".join(lines) + "\n"
    with open(os.path.join(ROOT, "profiles", name + ".txt"), "w") as f:
        f.write(text)
    print(text)


def summarize_wavefront(kt_dir):
    """Writes profiles/wavefront_pt_clutter_kernel_trace.txt: per-kernel totals of a kernel trace of
    `bench.py --clutter`, and the kernels of its last wavefront path-tracer draw and its last
    wavefront BPT draw (eye paths) bounce by bounce."""
    stats = _csv_rows(kt_dir, "_kernel_stats.csv")
    stats.sort(key=lambda r: -float(r["TotalDurationNs"]))
    rows = _csv_rows(kt_dir, "_kernel_trace.csv")
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    names = [_short_kernel(r["Kernel_Name"]) for r in rows]
    dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in rows]
    lines = ["Wavefront schedule, clutter room (5,120,016 triangles), 1920x1080, path length 8,",
             "roulette from bounce 3, wide traversal; every pass reads its queue length on the device.",
             "Each bounce's shadow (PT) or light-cache connection (BPT) trace and its resolve pass run on",
             "a second stream, beside the next bounce's closest-hit trace, so kernel durations overlap",
             "(their sum exceeds the draw's span).  Mode <0> of the pass kernels is PT, <1> BPT.",
             "",
             "Commands (repo root; the first after cd /tmp && export TMPDIR=/tmp):",
             "  rocprofv3 --kernel-trace --stats -d <out>/wfprof -o run --output-format csv -- "
             "python3 <repo>/bench.py --clutter --clutter-iterations 2 --steps 1",
             "  python3 bench.py --summarize-wavefront <out>/wfprof",
             "",
             "Per-kernel totals over the run (k_trace_wide launches include the trace benchmark before the "
             "integrators):",
             "%-60s %6s %10s %10s" % ("kernel", "calls", "total ms", "avg us")]
    for r in stats[:20]:
        lines.append("%-60s %6d %10.3f %10.1f" % (_short_kernel(r["Name"])[:60], int(r["Calls"]),
                                                 float(r["TotalDurationNs"]) / 1e6, float(r["AverageNs"]) / 1e3))
    resets = {}
    for r, nm, d in zip(rows, names, dur):
        if nm == "k_wide_queue_reset":
            resets.setdefault(r.get("Queue_Id", "?"), []).append(d)
    if resets:
        lines += ["", "k_wide_queue_reset (one workgroup zeroing the batch counters before each k_trace_wide) "
                  "by hardware queue:"]
        for qid, ds in sorted(resets.items()):
            ds.sort()
            lines.append("  queue %s: %d launches, median %.1f us, max %.1f us" % (qid, len(ds), ds[len(ds) // 2],
                                                                               ds[-1]))
        lines += ["A reset queued beside the other stream's persistent trace waits for a free wave slot, which",
                  "opens only as that trace's waves exit; the trace launched behind it waits for the same slots,",
                  "so its long durations there are queueing, not work."]
    for mode, title in ((0, "path-tracer"), (1, "BPT")):
        raygen = [i for i, n in enumerate(names) if n == "k_wf_raygen<%d>" % mode]
        if not raygen:
            raise SystemExit("bench.py --summarize-wavefront: no wavefront %s draw in the trace" % title)
        first = raygen[-1]
        last = next(i for i in range(first, len(names)) if names[i] == "k_wf_finish<%d>" % mode)
        lines += ["", "The last wavefront %s draw, per bounce (kernel durations in us; late bounces launch over the"
                  % title, "full queue capacity and their kernels exit once the device count is reached):"]
        if mode == 1:  # the light pass that precedes the eye paths
            lp = max(j for j in range(first) if names[j].startswith("k_lighttracer<"))
            lines.append("  light pass (%s): %.1f us, up to the eye paths' ray generation %.2f ms" % (
                names[lp], dur[lp], (int(rows[first]["Start_Timestamp"]) - int(rows[lp]["Start_Timestamp"])) / 1e6))
        bounce = 0
        for i in range(first, last):
            if names[i] != "k_wf_shade<%d>" % mode:
                continue
            ch = max(j for j in range(first, i) if names[j].startswith("k_trace_wide<false") or
                     names[j].startswith("k_trace_ref<false"))
            sh = min((j for j in range(i, last) if names[j].startswith("k_trace_wide<true") or
                      names[j].startswith("k_trace_ref<true")), default=None)
            rs = min((j for j in range(i, last) if names[j] == "k_wf_resolve<%d>" % mode), default=None)
            lines.append("  bounce %d:  closest-hit trace %8.1f  shade %7.1f  %s trace %7.1f  resolve %6.1f" % (
                bounce, dur[ch], dur[i], "shadow" if mode == 0 else "connection",
                dur[sh] if sh is not None else 0.0, dur[rs] if rs is not None else 0.0))
            bounce += 1
        span = (int(rows[last]["End_Timestamp"]) - int(rows[first]["Start_Timestamp"])) / 1e6
        busy = sum(dur[first:last + 1]) / 1e3
        lines.append("  %s: %.2f ms from the first to the last kernel, %.2f ms of kernel time" % (
            "draw" if mode == 0 else "eye paths", span, busy))
    text = "\n".join(lines) + "\n"
    with open(os.path.join(ROOT, "profiles", "wavefront_pt_clutter_kernel_trace.txt"), "w") as f:
        f.write(text)
    print(text)


MLP_KERNELS = ("k_mlp_grad", "k_mlp_reduce", "k_mlp_infer")


def summarize_mlp(kt_dir, *pmc_dirs):
    """Writes profiles/mlp_kernels.json: per-kernel average durations (kernel trace) and per-launch
    counters (--pmc passes) of the MLP kernels in a --mlp-kernels run, with the matrix-core roofline:
    FLOPs per launch from the model (checked against SQ_INSTS_VALU_MFMA_MOPS_BF16 x 512) over the
    average duration, against the BF16 dense peak, and the MFMA busy fraction of all 1024 SIMDs."""
    def short(name):
        for k in MLP_KERNELS:
            for tail in ("(", "<"):
                if "::" + k + tail in name or name.startswith(k + tail):
                    return k
        return None

    dur = {}
    for r in _csv_rows(kt_dir, "_kernel_stats.csv"):
        k = short(r["Name"])
        if k:
            dur[k] = {"calls": int(r["Calls"]), "avg_us": float(r["AverageNs"]) / 1e3,
                      "min_us": float(r["MinNs"]) / 1e3, "max_us": float(r["MaxNs"]) / 1e3}
    counters = {}
    for d in pmc_dirs:
        acc = {}
        for r in _csv_rows(d, "_counter_collection.csv"):
            k = short(r["Kernel_Name"])
            if k:
                acc.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
        for k, cs in acc.items():
            for c, v in cs.items():
                counters.setdefault(k, {})[c] = sum(v) / len(v)
    model = {"k_mlp_grad": 16384 * (MLP_TRAIN_FLOP), "k_mlp_infer": 1920 * 1080 * MLP_FWD_FLOP}
    kernels = {}
    for k in MLP_KERNELS:
        e = {"duration": dur.get(k), "per_launch": counters.get(k, {})}
        c = e["per_launch"]
        if k in model and e["duration"]:
            flop = model[k]
            tflops = flop / (e["duration"]["avg_us"] * 1e-6) / 1e12
            e["roofline"] = {"bound": "mfma", "flop_per_launch": flop, "achieved_tflops": round(tflops, 2),
                             "peak_tflops": BF16_PEAK_TFLOPS, "frac": round(tflops / BF16_PEAK_TFLOPS, 4)}
            if "SQ_INSTS_VALU_MFMA_MOPS_BF16" in c:
                e["roofline"]["counted_flop_per_launch"] = c["SQ_INSTS_VALU_MFMA_MOPS_BF16"] * 512.0
        if "SQ_VALU_MFMA_BUSY_CYCLES" in c and "GRBM_GUI_ACTIVE" in c:
            # GRBM_GUI_ACTIVE sums the launch's busy cycles over the 8 XCDs; 1024 SIMDs
            e["mfma_busy_frac"] = c["SQ_VALU_MFMA_BUSY_CYCLES"] / (c["GRBM_GUI_ACTIVE"] / 8.0 * 1024.0)
        if "SQ_LDS_BANK_CONFLICT" in c and "SQ_LDS_IDX_ACTIVE" in c and c["SQ_LDS_IDX_ACTIVE"] > 0:
            e["lds_bank_conflict_frac"] = c["SQ_LDS_BANK_CONFLICT"] / c["SQ_LDS_IDX_ACTIVE"]
        if "SQ_WAVES" in c and c["SQ_WAVES"] > 0:
            for key in ("SQ_INSTS_MFMA", "SQ_INSTS_LDS", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_VMEM_RD",
                        "SQ_INSTS_VMEM_WR"):
                if key in c:
                    e.setdefault("per_wave", {})[key] = c[key] / c["SQ_WAVES"]
            if "SQ_WAVE_CYCLES" in c:
                e.setdefault("per_wave", {})["SQ_WAVE_CYCLES"] = c["SQ_WAVE_CYCLES"] / c["SQ_WAVES"]
            if "SQ_WAIT_INST_LDS" in c:
                e.setdefault("per_wave", {})["SQ_WAIT_INST_LDS"] = c["SQ_WAIT_INST_LDS"] / c["SQ_WAVES"]
            if "SQ_WAIT_INST_ANY" in c:
                e.setdefault("per_wave", {})["SQ_WAIT_INST_ANY"] = c["SQ_WAIT_INST_ANY"] / c["SQ_WAVES"]
        kernels[k] = e
    out = {"workload": "bench.py --mlp-kernels: training step over 16,384 samples (k_mlp_grad, k_mlp_reduce, "
                       "k_mlp_adam), inference over 1920x1080 queries (k_mlp_infer)",
           "commands": ["cd /tmp && export TMPDIR=/tmp && timeout -k 10 300 rocprofv3 --kernel-trace --stats -d <out>/kt "
                        "-o kt --output-format csv -- python3 <repo>/bench.py --mlp-kernels --mlp-kernel-steps 20",
                        "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 120 rocprofv3 --pmc <counters of one pass> "
                        "-d <out>/<pass> -o <pass> --output-format csv -- python3 <repo>/bench.py --mlp-kernels "
                        "--mlp-kernel-steps 5",
                        "python3 <repo>/bench.py --summarize-mlp <out>/kt <out>/<pass> ..."],
           "kernels": kernels}
    with open(os.path.join(ROOT, "profiles", "mlp_kernels.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out))


def host_threads():
    """Host cores this process may use: OMP_NUM_THREADS when set (the share a shared
    host grants the run), else the CPUs in the affinity mask."""
    try:
        return max(1, int(os.environ["OMP_NUM_THREADS"]))
    except (KeyError, ValueError):
        return max(1, len(os.sched_getaffinity(0)))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene_arrays, rays_host, seconds):
    """The CPU oracle's escape walk on `host_threads()` threads (its C calls release
    the GIL), each taking 4096-ray chunks of the sample in order until `seconds` pass."""
    import concurrent.futures
    import threading
    from oracle import oracle
    sc = oracle.Scene(scene_arrays)
    threads = host_threads()
    chunk = 4096
    lock = threading.Lock()
    state = {"next": 0, "done": 0}
    t0 = time.perf_counter()

    def worker():
        while time.perf_counter() - t0 < seconds:
            with lock:
                first = state["next"]
                if first >= len(rays_host):
                    return
                state["next"] = first + chunk
            part = rays_host[first:first + chunk]
            oracle.trace(sc, part)
            with lock:
                state["done"] += len(part)

    with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as ex:
        for f in [ex.submit(worker) for _ in range(threads)]:
            f.result()
    dt = time.perf_counter() - t0
    done = state["done"]
    return {"value": done / dt / 1e9, "unit": "Grays/s", "cores": threads, "kind": "port",
            "sample": "%d incoherent heightfield rays (the first of the 2^24 set), CPU oracle escape walk, "
                      "%d threads, %.1f s; host: %s, %d CPUs visible" % (done, threads, dt, cpu_model(),
                                                                     os.cpu_count() or 0)}


def valu_from_profile(name="trace_wide_heightfield_utilization.json"):
    """VALU issue fraction (trace kernel over the calibration kernel, wave-instructions
    per CU cycle) and active lanes per VALU instruction, from a committed counter summary."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)["derived"]
    except (OSError, KeyError, ValueError):
        return None
    out = {k: d[k] for k in ("valu_issue_frac_of_probe", "valu_wave_insts_per_cu_cycle",
                             "probe_valu_wave_insts_per_cu_cycle", "valu_active_lanes",
                             "valu_wave_instructions_per_ray", "tcc_hit_rate") if k in d}
    return out or None


def main():
    args = parse()
    if args.summarize:
        summarize(*args.summarize)
        return
    if args.summarize_counters:
        summarize_counters(*args.summarize_counters)
        return
    if args.summarize_integrators:
        summarize_integrators(*args.summarize_integrators)
        return
    if args.summarize_mlp:
        summarize_mlp(*args.summarize_mlp)
        return
    if args.summarize_wavefront:
        summarize_wavefront(args.summarize_wavefront)
        return
    if args.summarize_kernels:
        summarize_kernels(*args.summarize_kernels, args.profiled_args)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (ray sets and pixels would "
              "shard across GPUs with no communication during traversal)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.mlp_kernels:
        print(json.dumps(mlp_kernel_benchmark(args.mlp_kernel_steps)))
        return
    if args.mlp:
        print(json.dumps(mlp_benchmark(args.mlp_ref_spp, args.mlp_frames, args.mlp_warmup)))
        return
    if args.hierarchy:
        print(json.dumps(hierarchy_benchmark(args.hierarchy_iterations)))
        return
    if args.batch:
        print(json.dumps(batch_benchmark(args.batch_frames)))
        return
    if args.clutter and args.sweep:
        clutter_sweep()
        return
    if args.clutter:
        print(json.dumps(clutter_benchmark(args.steps, args.clutter_iterations, args.clutter_incoherent_only,
                                           args.calibrate, args.valu_probe)))
        return
    from gpugi_amd import rt, workloads

    t_setup = time.perf_counter()
    hs = rt.HostScene.heightfield(workloads.HF_CELLS, workloads.MASTER_SEED)
    ctx = rt.Context(0)
    ctx.load_scene(hs)
    n = args.rays
    R = workloads.HeightfieldRays(ctx, n=n)
    kernel = rt.KERNEL_WIDE if args.kernel == "wide" else rt.KERNEL_REFERENCE

    if args.sweep:
        sweep_wide_configs(ctx, R, n, itertools.product((1, 2, 4), (8, 10, 12, 16), (8, 16, 32), (1, 8, 16, 32)))
        ctx.close()
        return

    # Autotune the wide kernel on a separate incoherent sample (other seed), untimed.
    tuned = None
    if kernel == rt.KERNEL_WIDE and args.autotune:
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
             "alg_bytes_per_ray": alg_bytes / n,
             "wide_config": tuned or ctx.wide_config()}
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
        # the same walk over oriented boxes (bvhType 1), then back to axis-aligned boxes
        ctx.set_bvh_type(rt.BVH_OBOX)
        ctx.load_scene(hs)
        ctx.timer_begin()
        ctx.trace_device(R.incoherent, R.hits, n, rt.CLOSEST_HIT, rt.KERNEL_REFERENCE)
        extra["reference_walk_obox_incoherent_grays"] = n / (ctx.timer_end() * 1e-3) / 1e9
        ctx.set_bvh_type(rt.BVH_AABOX)
        ctx.load_scene(hs)  # the wide configuration is the context's and survives the reload
        # host buffers through the same entry point: upload of the rays, the launch and the hit download
        rays_np = R.incoherent.download(np.float32, n * 8).view(rt.RAY_DTYPE)
        ctx.trace(rays_np[: 1 << 20], rt.CLOSEST_HIT, kernel)
        t0 = time.perf_counter()
        ctx.trace(rays_np, rt.CLOSEST_HIT, kernel)
        extra["host_buffers_pcie_inclusive_grays"] = n / (time.perf_counter() - t0) / 1e9
        # the same from pinned host buffers (gpugi_host_alloc): copies overlap the traversal
        pr = ctx.host_empty(n, rt.RAY_DTYPE)
        ph = ctx.host_empty(n, rt.HIT_DTYPE)
        pr[:] = rays_np
        del rays_np
        ctx.trace(pr[: 1 << 23], rt.CLOSEST_HIT, kernel, out=ph)
        t0 = time.perf_counter()
        ctx.trace(pr, rt.CLOSEST_HIT, kernel, out=ph)
        extra["pinned_host_buffers_pcie_inclusive_grays"] = n / (time.perf_counter() - t0) / 1e9
        ctx.host_free(pr)
        ctx.host_free(ph)

    if args.calibrate:
        a, b = ctx.alloc(1 << 30), ctx.alloc(1 << 30)
        for _ in range(3):
            b.copy_from(a, 1 << 30)
        ctx.sync()
        a.free()
        b.free()
    if args.valu_probe:
        extra["valu_probe"] = run_valu_probe(ctx)
    # the reference walk's algorithmic bytes over the launch time (the contract's per-ray figure)
    extra["alg_bytes_gbs"] = round(achieved, 2)
    extra["alg_bytes_over_hbm_peak"] = round(achieved / HBM_PEAK_GBS, 4)
    traffic = traffic_from_profile()
    roofline = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": traffic}
    if traffic:
        # HBM traffic the counters measured for one launch, over this run's launch time
        hbm_gbs = traffic / (ms * 1e-3) / 1e9
        roofline.update(achieved=round(hbm_gbs, 2), frac=round(hbm_gbs / HBM_PEAK_GBS, 4))
    valu = valu_from_profile()
    if valu:
        roofline["valu"] = valu
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
        "roofline": roofline,
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
