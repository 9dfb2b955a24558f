"""Benchmark: Mrays/s over all bounces of the full path trace on one MI355X.

Workload: the 1920x1080 render of the full procedural scene (room + 500k-tri
water grid + 200k-tri textured surface of revolution + 20k-tri mirror
sphere, 721 316 triangles), 2x2 supersampling (8 294 400 camera rays),
depth 8, FP32.  One step is one sample pass over every camera ray with all
bounces; K steps are one render of K samples per pixel (the primary hits are
cast once per render and cached, as the reference does).  Rays are counted
like the reference's intersect launches: primaries once plus every
continuation ray of every pass.

Besides the timed render the script reports:
  roofline      the cast kernel (dominant): `achieved` is its algorithmic
                bytes -- the information content of the reference's search:
                32 B per ray (origin and direction in, hit id and t out),
                40 B per node box-tested (box and links), 4 B per leaf
                collected (list entry), 56 B per triangle test (list entry,
                int4, three vertices) -- over its average launch time from
                HIP events on the render stream, against 8 TB/s HBM.
                `traffic` is the HBM bytes per launch rocprofv3 counted
                (FETCH_SIZE, WRITE_SIZE, scaled by a copy of known size) and
                `issue` the VALU side of the same kernel (busy fraction,
                active lanes, instructions per node visit and per triangle
                test), both read from profiles/cast_counters_full1080p.json
                when that file was measured on the same kernel sources,
                tuning and workload (else null).
  cpu_baseline  the float32 C restatement of the reference kernels (oracle/)
                on the host cores: camera rays sampled in 32 bands spread
                evenly over the frame, at the same spp; one warm-up, then
                the median of three timed runs.
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]

Profiling helpers (not benchmark runs):
  python bench.py --calibrate-copy
        3 device copies of 1 GiB: what FETCH_SIZE / WRITE_SIZE read per byte
        moved (run under rocprofv3 --pmc).
  python bench.py --summarize-counters DIR
        reads the rocprofv3 outputs of the counter runs under DIR (fetch/,
        write/, sq/, l2/, cal_fetch/, cal_write/, trace/, and bench_config.json
        written by --profile-record DIR) and writes
        profiles/cast_counters_full1080p.json.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--water", type=int, default=500, help="water grid side (500 -> 500k tris)")
    ap.add_argument("--no-supersample", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=4,
                    help="passes of the kernel-timed and work-counted renders")
    ap.add_argument("--scene-dir", default=os.path.join(tempfile.gettempdir(), "hpt_scenes"))
    ap.add_argument("--batch-views", type=int, default=64, help="0 skips the batching leg")
    ap.add_argument("--serve-requests", type=int, default=40, help="0 skips the serving leg")
    ap.add_argument("--serve-rates", type=float, nargs="*", default=[2.0, 8.0])
    ap.add_argument("--literal-steps", type=int, default=None,
                    help="passes of the literal-arithmetic comparison render (default: --steps; "
                         "0 skips it)")
    ap.add_argument("--cast-tris", type=int, default=200_000,
                    help="triangles of the cast-only soup (0 skips the cast-only leg)")
    ap.add_argument("--profile-record", metavar="DIR",
                    help="write DIR/bench_config.json (kernel sources, tuning, workload)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run the copy of known size for the byte counters")
    ap.add_argument("--summarize-counters", metavar="DIR",
                    help="only write the counter record from the rocprofv3 runs under DIR")
    ap.add_argument("--serve-profile", action="store_true",
                    help="only the serving burst for a kernel trace: a warm burst, a 1 s "
                         "pause, the measured burst (default launch forms)")
    args = ap.parse_args()

    if args.summarize_counters:
        summarize_counters(args.summarize_counters)
        return
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.serve_profile:
        serve_profile(args)
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(sample-space or tile sharding across GPUs is a design, not code)"}))
        sys.exit(2)

    import numpy as np
    import torch

    from hptracing_amd import scenes
    from hptracing_amd.runner import TraceRunner, make_view
    from hptracing_amd.scene import KDTree

    t0 = time.time()
    path = scenes.cached(args.scene_dir, scenes.write_full, 0, water_side=args.water)
    log(f"[bench] scene written {time.time() - t0:.1f}s: {path}")
    t0 = time.time()
    kd = KDTree(path, os.path.dirname(path) + "/")
    log(f"[bench] scene loaded, KD built in {kd.build_seconds:.1f}s "
        f"({kd.n_triangles} tris, {len(kd.nodes)} nodes), total {time.time() - t0:.1f}s")
    runner = TraceRunner(kd)
    log(f"[bench] device scene {runner.stats['device_bytes'] / 1e9:.2f} GB")

    ss = 1 if args.no_supersample else 2
    view = dict(**scenes.ROOM_VIEW, width=args.width / args.height, height=1.0,
                sample_x=args.width, sample_y=args.height, supersample_x=ss, supersample_y=ss,
                background_color=(0.0, 0.0, 0.0), depth=args.depth, disable_diffuse=0,
                brightness=1.0)

    if args.warmup > 0:
        runner.render(make_view(**view, sample=args.warmup, seed=12345))
        log(f"[bench] warmup {args.warmup} pass(es) done")

    torch.cuda.synchronize()
    t_start = time.perf_counter()
    runner.render(make_view(**view, sample=args.steps, seed=1))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    st = runner.last_stats
    rays = int(st.rays_cast)
    value = rays / elapsed / 1e6
    live = [int(st.live_per_depth[i]) for i in range(args.depth)]
    R = int(st.primary_rays)
    log(f"[bench] timed render: {args.steps} passes, {rays} rays, {elapsed:.3f}s, "
        f"{value:.1f} Mrays/s; live per depth {live}")

    # The same render with the other arithmetic of the intersection test
    # (every product and sum rounded, as kernel.cl writes dot/cross), for
    # comparison only; the default is restored before anything else runs.
    literal = None
    lit_steps = args.steps if args.literal_steps is None else args.literal_steps
    if lit_steps > 0:
        runner.set_arithmetic(False)
        torch.cuda.synchronize()
        t_lit = time.perf_counter()
        runner.render(make_view(**view, sample=lit_steps, seed=1))
        torch.cuda.synchronize()
        lit_elapsed = time.perf_counter() - t_lit
        lit_rays = int(runner.last_stats.rays_cast)
        runner.set_arithmetic(True)
        literal = {"steps": lit_steps, "mrays_s": round(lit_rays / lit_elapsed / 1e6, 3),
                   "ms_per_step": round(lit_elapsed * 1e3 / lit_steps, 3)}
        log(f"[bench] literal arithmetic: {lit_elapsed:.3f}s, {literal['mrays_s']:.1f} Mrays/s")

    # Kernel time and algorithmic work of the same launches.
    pk = max(1, min(args.profile_steps, args.steps))
    runner.render(make_view(**view, sample=pk, seed=1), time_kernels=True)
    tst = runner.last_stats
    runner.render(make_view(**view, sample=pk, seed=1), count_work=True)
    cst = runner.last_stats
    cast_bytes = (32 * int(cst.rays_cast) + 40 * int(cst.node_visits) + 4 * int(cst.leaf_visits)
                  + 56 * int(cst.tri_tests))
    per_launch_ms = tst.cast_ms / max(1, tst.cast_launches)
    log(f"[bench] profile ({pk} passes): cast {tst.cast_ms:.1f} ms in {tst.cast_launches} "
        f"launches, shade {tst.shade_ms:.1f} ms in {tst.shade_launches}; "
        f"visits {cst.node_visits}, leaves {cst.leaf_visits}, tests {cst.tri_tests}")
    config = bench_config(args, runner)
    if args.profile_record:
        os.makedirs(args.profile_record, exist_ok=True)
        with open(os.path.join(args.profile_record, "bench_config.json"), "w") as f:
            json.dump(config, f, indent=1)
    rec, rec_source = load_counters(config)
    roofline = roofline_record(cast_bytes, per_launch_ms, tst, cst, rec, rec_source)
    log(f"[bench] roofline: {roofline}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(kd, view, args.steps, args.cpu_seconds, R)

    batching = None
    if args.batch_views > 0:
        batching = bench_batching(runner, args.batch_views, args.depth)

    serving = None
    if args.serve_requests > 0:
        del runner
        serving = bench_serving(args)

    cast_only = bench_cast_only(args.cast_tris) if args.cast_tris > 0 else None

    out = {
        "metric": "Mrays/s (all bounces)",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene, generated OBJ/MTL/PPM)",
        "config": {
            "workload": f"full path trace {args.width}x{args.height}"
                        f"{' -x' if ss == 2 else ''} depth {args.depth}, "
                        f"{kd.n_triangles}-tri scene, {args.steps} spp per step-set",
            "camera_rays": R,
            "spp": args.steps,
            "depth": args.depth,
            "triangles": kd.n_triangles,
            "kd_nodes": len(kd.nodes),
            "rays_cast": rays,
            "live_per_depth": live,
            "arithmetic": "rocm_opencl: dot and cross of the intersection test as ROCm's "
                          "OpenCL library evaluates them (hit triangles and distances equal "
                          "the reference kernel's bit for bit)",
            "parallelism": "single GPU",
        },
        "literal_arithmetic": literal,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batching": batching,
        "serving": serving,
        "cast_only": cast_only,
    }
    print(json.dumps(out), flush=True)


COUNTERS_FILE = os.path.join(ROOT, "profiles", "cast_counters_full1080p.json")
CAST_KERNEL = "k_cast_refill"
SQ_SET = ("SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_INSTS_VALU SQ_ACTIVE_INST_VALU "
          "SQ_THREAD_CYCLES_VALU SQ_INSTS_SALU SQ_WAIT_INST_ANY")
# The vector memory path of each CU: texture addresser (TA) and data (TD)
# busy and stalled on the L1 (TCP), the L1's stalls on data pending from
# the L2, and its L2 read requests with their summed latency, which give
# the mean number of L2 reads in flight per CU (Little's law).
MEM_SET = ("TA_TA_BUSY_sum TA_ADDR_STALLED_BY_TC_CYCLES_sum TD_TD_BUSY_sum TD_TC_STALL_sum "
           "TCP_PENDING_STALL_CYCLES_sum TCP_TCC_READ_REQ_LATENCY_sum TCP_TCC_READ_REQ_sum")
N_CU = 256


def device_source_sha():
    """sha256 over the kernel sources and the C header, in name order: a
    counter record applies only to the kernels it was measured on."""
    import glob
    import hashlib
    h = hashlib.sha256()
    files = sorted(glob.glob(os.path.join(ROOT, "hptracing_amd", "csrc", "device", "*")))
    for f in files + [os.path.join(ROOT, "include", "hptracing.h")]:
        h.update(os.path.basename(f).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def bench_config(args, runner):
    """What a counter record must match: kernel sources, tuning, workload."""
    return {"device_source_sha": device_source_sha(), "tuning": runner.get_tuning(),
            "width": args.width, "height": args.height, "depth": args.depth,
            "water": args.water, "supersample": not args.no_supersample,
            "profile_steps": max(1, min(args.profile_steps, args.steps))}


def load_counters(config):
    """The counter record of profiles/cast_counters_full1080p.json, or
    (None, reason) when it is missing or was measured on other kernel
    sources, another tuning or another workload."""
    if not os.path.exists(COUNTERS_FILE):
        return None, "no counter record"
    with open(COUNTERS_FILE) as f:
        rec = json.load(f)
    for k, v in config.items():
        if rec.get("config", {}).get(k) != v:
            return None, f"counter record stale ({k} differs)"
    return rec, "profiles/cast_counters_full1080p.json"


def calibrate_copy():
    """Plain device-to-device copies of known size for the byte counters."""
    import torch
    n = 1 << 28  # 256 Mi floats = 1 GiB
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("[bench] copied 3 x 1 GiB")


def _counter_rows(directory, kernel_substr):
    """Per-dispatch sums of every counter of the kernels whose name holds
    kernel_substr, from rocprofv3's counter_collection CSV under directory."""
    import csv
    import glob
    files = glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {directory}")
    per = {}
    for fn in files:
        with open(fn, newline="") as f:
            for row in csv.DictReader(f):
                if kernel_substr not in row["Kernel_Name"]:
                    continue
                key = (fn, row["Dispatch_Id"])
                d = per.setdefault(key, {"kernel": row["Kernel_Name"]})
                d[row["Counter_Name"]] = d.get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
    return list(per.values())


def _avg(rows, name):
    vals = [r[name] for r in rows if name in r]
    if not vals:
        raise SystemExit(f"counter {name} missing")
    return sum(vals) / len(vals)


def summarize_counters(directory):
    """profiles/cast_counters_full1080p.json from the counter runs under
    directory: bytes per launch (FETCH_SIZE x copy factor + WRITE_SIZE x
    copy factor), the SQ set per dispatch, the kernel-trace average and the
    derived issue figures."""
    import csv
    import glob
    with open(os.path.join(directory, "bench_config.json")) as f:
        cfg = json.load(f)
    fetch = _counter_rows(os.path.join(directory, "fetch"), CAST_KERNEL)
    write = _counter_rows(os.path.join(directory, "write"), CAST_KERNEL)
    sq = _counter_rows(os.path.join(directory, "sq"), CAST_KERNEL)
    cal_f = _counter_rows(os.path.join(directory, "cal_fetch"), "")
    cal_w = _counter_rows(os.path.join(directory, "cal_write"), "")
    copy_kib = float(1 << 20)  # one 1 GiB copy, in KiB
    f_factor = copy_kib / max(r["FETCH_SIZE"] for r in cal_f)
    w_factor = copy_kib / max(r["WRITE_SIZE"] for r in cal_w)
    fetch_b = _avg(fetch, "FETCH_SIZE") * 1024 * f_factor
    write_b = _avg(write, "WRITE_SIZE") * 1024 * w_factor
    stats = glob.glob(os.path.join(directory, "trace", "**", "*kernel_stats.csv"), recursive=True)
    avg_us = None
    for fn in stats:
        with open(fn, newline="") as f:
            for row in csv.DictReader(f):
                if CAST_KERNEL in row["Name"]:
                    avg_us = float(row["AverageNs"]) / 1e3
    counters = {k: _avg(sq, k) for k in SQ_SET.split()}
    l2_dir = os.path.join(directory, "l2")
    l2_hit = None
    if os.path.isdir(l2_dir):
        l2 = _counter_rows(l2_dir, CAST_KERNEL)
        hit, miss = _avg(l2, "TCC_HIT_sum"), _avg(l2, "TCC_MISS_sum")
        counters.update({"TCC_HIT_sum": hit, "TCC_MISS_sum": miss})
        l2_hit = hit / (hit + miss)
    n_simd, n_se = 1024, 32
    cycles = counters["SQ_BUSY_CYCLES"] / n_se
    derived = {
        "dispatch_cycles": cycles,
        "clock_ghz": cycles / (avg_us * 1e3) if avg_us else None,
        "valu_busy": counters["SQ_INSTS_VALU"] * 4 / (n_simd * cycles),
        "active_lanes_of_64": counters["SQ_THREAD_CYCLES_VALU"] / counters["SQ_ACTIVE_INST_VALU"],
        "waves_per_simd": counters["SQ_WAVE_CYCLES"] * 4 / (n_simd * cycles),
        "salu_per_valu": counters["SQ_INSTS_SALU"] / counters["SQ_INSTS_VALU"],
        "l2_hit": l2_hit,
    }
    mem_dir = os.path.join(directory, "mem")
    if os.path.isdir(mem_dir):
        mem = _counter_rows(mem_dir, CAST_KERNEL)
        m = {k: _avg(mem, k) for k in MEM_SET.split()}
        counters.update(m)
        cu_cycles = N_CU * cycles  # the _sum counters add one instance per CU
        derived.update({
            "ta_busy": m["TA_TA_BUSY_sum"] / cu_cycles,
            "ta_addr_stalled_by_l1": m["TA_ADDR_STALLED_BY_TC_CYCLES_sum"] / cu_cycles,
            "td_busy": m["TD_TD_BUSY_sum"] / cu_cycles,
            "td_stalled_on_l1": m["TD_TC_STALL_sum"] / cu_cycles,
            "l1_pending_stall": m["TCP_PENDING_STALL_CYCLES_sum"] / cu_cycles,
            "l2_reads_per_launch": m["TCP_TCC_READ_REQ_sum"],
            "l2_read_latency_cycles": m["TCP_TCC_READ_REQ_LATENCY_sum"] / m["TCP_TCC_READ_REQ_sum"],
            "l2_reads_in_flight_per_cu": m["TCP_TCC_READ_REQ_LATENCY_sum"] / cu_cycles,
        })
    rec = {
        "kernel": sq[0]["kernel"],
        "config": cfg,
        "commands": {
            "counters": "rocprofv3 --pmc <FETCH_SIZE | WRITE_SIZE | " + SQ_SET + " | "
                        "TCC_HIT_sum TCC_MISS_sum | " + MEM_SET + "> --output-format csv -d "
                        "<DIR>/<fetch|write|sq|l2|mem> -o run -- python bench.py "
                        "--steps 4 --warmup 0 --profile-steps 4 --no-cpu --batch-views 0 "
                        "--serve-requests 0 --cast-tris 0 --literal-steps 0 --profile-record <DIR>",
            "calibration": "rocprofv3 --pmc <FETCH_SIZE | WRITE_SIZE> --output-format csv "
                           "-d <DIR>/<cal_fetch|cal_write> -o run -- python bench.py "
                           "--calibrate-copy",
            "trace": "rocprofv3 --kernel-trace --stats --output-format csv -d <DIR>/trace -o run "
                     "-- python bench.py --steps 4 --warmup 0 --profile-steps 4 --no-cpu "
                     "--batch-views 0 --serve-requests 0 --cast-tris 0 --literal-steps 0",
        },
        "dispatches": len(sq),
        "avg_us": avg_us,
        "fetch_factor": round(f_factor, 4), "write_factor": round(w_factor, 4),
        "fetch_bytes_per_launch": int(fetch_b), "write_bytes_per_launch": int(write_b),
        "bytes_per_launch": int(fetch_b + write_b),
        "sq_per_dispatch": counters,
        "derived": derived,
    }
    with open(COUNTERS_FILE, "w") as f:
        json.dump(rec, f, indent=1)
    log(f"[bench] wrote {COUNTERS_FILE}: {rec['bytes_per_launch'] / 1e9:.2f} GB per launch, "
        f"VALU busy {derived['valu_busy']:.3f}, lanes {derived['active_lanes_of_64']:.1f}")
    write_stage_tables(directory, cfg, f_factor, w_factor, rec)


def _kernel_stats(directory):
    import csv
    import glob
    out = {}
    for fn in glob.glob(os.path.join(directory, "trace", "**", "*kernel_stats.csv"), recursive=True):
        with open(fn, newline="") as f:
            for row in csv.DictReader(f):
                out[row["Name"]] = (int(row["Calls"]), float(row["TotalDurationNs"]) / 1e6,
                                    float(row["AverageNs"]) / 1e3, float(row["Percentage"]))
    return out


def _short(name):
    return name.replace("void ", "").split("(")[0]


def write_stage_tables(directory, cfg, f_factor, w_factor, rec):
    """profiles/kernel_stats_full1080p.md and stage_counters_full1080p.md:
    every kernel of the profile workload, from the same rocprofv3 runs."""
    stats = _kernel_stats(directory)
    per = {}
    for key, sub in (("fetch", "fetch"), ("write", "write"), ("sq", "sq"), ("l2", "l2")):
        for r in _counter_rows(os.path.join(directory, sub), ""):
            d = per.setdefault(_short(r["kernel"]), {})
            for c, v in r.items():
                if c != "kernel":
                    d.setdefault(c, []).append(v)
    cmd_b = ("python bench.py --steps 4 --warmup 0 --profile-steps 4 --no-cpu --batch-views 0 "
             "--serve-requests 0 --cast-tris 0 --literal-steps 0")
    t = cfg["tuning"]
    lines = ["# Kernel time, full 1080p -x render (721 316 triangles, depth 8), 4 passes", "",
             "Command (MI355X, ROCm 7.2):", "",
             f"    rocprofv3 --kernel-trace --stats --output-format csv -d <DIR>/trace -o run -- {cmd_b}",
             "", f"Kernel sources {cfg['device_source_sha']} (bench.py device_source_sha); tuning "
             f"fetch {t['fetch']}, {t['cast_blocks_per_cu']} cast blocks per CU, refill "
             f"{t['refill_threshold']}, node steps {t['node_steps']}.  The run holds the 4-pass "
             "timed render and the 4-pass kernel-timed render (one launch sequence each: 1 primary "
             "+ 7 bounce casts), then the 4-pass work-counting render, whose cast is the counting "
             "variant `k_cast<true, ...>` (not the timed path).", "",
             "| kernel | calls | total ms | average us | % |", "|---|---|---|---|---|"]
    for name, (calls, tot, avg, pct) in sorted(stats.items(), key=lambda x: -x[1][1]):
        lines.append(f"| `{_short(name)}` | {calls} | {tot:.2f} | {avg:.1f} | {pct:.2f} |")
    with open(os.path.join(ROOT, "profiles", "kernel_stats_full1080p.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    n_simd, n_se = 1024, 32
    rows = ["# Per-stage hardware counters, full 1080p -x render (721 316 triangles, depth 8)", "",
            "Commands (MI355X, ROCm 7.2), one counter set per run, with the kernel trace of "
            "kernel_stats_full1080p.md for durations and copies of known size for the byte "
            "counters:", "", "    " + rec["commands"]["counters"], "    " + rec["commands"]["calibration"],
            "", f"Per dispatch, averaged over each kernel's dispatches.  HBM bytes are FETCH_SIZE x "
            f"{f_factor:.4f} + WRITE_SIZE x {w_factor:.4f} (what the counters read for a 1 GiB "
            "copy); HBM TB/s divides them by the kernel-trace average.  VALU busy = SQ_INSTS_VALU "
            "x 4 / (1024 SIMDs x SQ_BUSY_CYCLES / 32); lanes = SQ_THREAD_CYCLES_VALU / "
            "SQ_ACTIVE_INST_VALU (of 64); waves per SIMD = SQ_WAVE_CYCLES x 4 / (1024 x cycles); "
            "L2 hit = TCC_HIT / (TCC_HIT + TCC_MISS).  Kernel sources "
            f"{cfg['device_source_sha']}.", "",
            "| kernel | avg us | HBM read GB | HBM write GB | HBM TB/s | of 8 TB/s | L2 hit | "
            "VALU busy | lanes | waves/SIMD |", "|---|---|---|---|---|---|---|---|---|---|"]
    avg_of = {_short(k): v[2] for k, v in stats.items()}
    for name, d in sorted(per.items(), key=lambda x: -avg_of.get(x[0], 0)):
        if name not in avg_of or "SQ_BUSY_CYCLES" not in d or "FETCH_SIZE" not in d:
            continue
        a = lambda c: sum(d[c]) / len(d[c])  # noqa: E731
        us = avg_of[name]
        rd, wr = a("FETCH_SIZE") * 1024 * f_factor / 1e9, a("WRITE_SIZE") * 1024 * w_factor / 1e9
        cyc = a("SQ_BUSY_CYCLES") / n_se
        tbs = (rd + wr) / (us * 1e-6) / 1e3
        hit = a("TCC_HIT_sum") / max(1.0, a("TCC_HIT_sum") + a("TCC_MISS_sum")) if "TCC_HIT_sum" in d else float("nan")
        lanes = a("SQ_THREAD_CYCLES_VALU") / max(1.0, a("SQ_ACTIVE_INST_VALU"))
        rows.append(f"| `{name}` | {us:.1f} | {rd:.3f} | {wr:.3f} | {tbs:.2f} | {tbs / 8:.2f} | "
                    f"{hit:.3f} | {a('SQ_INSTS_VALU') * 4 / (n_simd * cyc):.2f} | {lanes:.1f} | "
                    f"{a('SQ_WAVE_CYCLES') * 4 / (n_simd * cyc):.2f} |")
    with open(os.path.join(ROOT, "profiles", "stage_counters_full1080p.md"), "w") as f:
        f.write("\n".join(rows) + "\n")
    log("[bench] wrote profiles/kernel_stats_full1080p.md and stage_counters_full1080p.md")


def roofline_record(cast_bytes, per_launch_ms, tst, cst, rec, rec_source):
    """The roofline object of the cast kernel: algorithmic bytes over the
    HIP-event launch time against the HBM peak (`frac`), and, from the
    counter record when it applies, the real traffic and the issue side."""
    launches = max(1, tst.cast_launches)
    algo = cast_bytes / launches
    launch_s = per_launch_ms / 1e3
    achieved = algo / launch_s / 1e9 if launch_s > 0 else 0.0
    out = {
        "bound": "hbm",
        "kernel": "k_cast_refill (KD traversal + triangle intersection)",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "avg_launch_ms": round(per_launch_ms, 3),
        "algorithmic_bytes_per_launch": int(algo),
        "counters_source": rec_source,
    }
    if rec is None:
        return out
    traffic = rec["bytes_per_launch"]
    tgbs = traffic / launch_s / 1e9
    d = rec["derived"]
    sq = rec["sq_per_dispatch"]
    lane_insts = sq["SQ_INSTS_VALU"] * d["active_lanes_of_64"]
    visits = int(cst.node_visits) / launches
    tests = int(cst.tri_tests) / launches
    memory = None
    if "l2_reads_in_flight_per_cu" in d:
        memory = {
            "l2_reads_in_flight_per_cu": round(d["l2_reads_in_flight_per_cu"], 1),
            "l2_read_latency_cycles": round(d["l2_read_latency_cycles"], 1),
            "l2_reads_per_launch": int(d["l2_reads_per_launch"]),
            "ta_busy": round(d["ta_busy"], 3),
            "ta_addr_stalled_by_l1": round(d["ta_addr_stalled_by_l1"], 3),
            "td_busy": round(d["td_busy"], 3),
            "td_stalled_on_l1": round(d["td_stalled_on_l1"], 3),
            "l1_pending_stall": round(d["l1_pending_stall"], 3),
        }
    if memory:
        binding = (f"the L1's reads in flight to the L2: {memory['l2_reads_in_flight_per_cu']:.0f} "
                   f"per CU on average (Little's law over {memory['l2_read_latency_cycles']:.0f}-cycle "
                   f"reads), the texture address path stalled by the L1 "
                   f"{memory['ta_addr_stalled_by_l1']:.2f} of cycles and its data path waiting "
                   f"on it {memory['td_stalled_on_l1']:.2f}, so throughput is that concurrency over "
                   f"the latency; VALU busy {d['valu_busy']:.2f} of cycles with "
                   f"{d['active_lanes_of_64']:.1f} of 64 lanes active, HBM traffic "
                   f"{tgbs / HBM_PEAK_GBS:.2f} of peak.  A third more node loads per trip (to "
                   f"lines already in flight) cost 17 %; 4.5 % fewer VALU instructions 1.6 %")
    else:
        binding = (f"VALU issue and the latency of dependent random fetches below the L2 "
                   f"together, neither alone: VALU busy {d['valu_busy']:.2f} of cycles with "
                   f"{d['active_lanes_of_64']:.1f} of 64 lanes active, HBM traffic "
                   f"{tgbs / HBM_PEAK_GBS:.2f} of peak in 64- and 128-byte random reads; "
                   f"4.5 % fewer VALU instructions (12 % fewer SALU, 17 % fewer LDS) "
                   f"shortened the launch 1.6 %, and 23 % fewer node-fetch trips per ray "
                   f"1.6 %")
    out.update({
        "traffic": traffic,
        "traffic_gbs": round(tgbs, 1),
        "traffic_frac": round(tgbs / HBM_PEAK_GBS, 4),
        "issue": {
            "valu_busy": round(d["valu_busy"], 3),
            "active_lanes_of_64": round(d["active_lanes_of_64"], 1),
            "useful_lane_frac": round(d["valu_busy"] * d["active_lanes_of_64"] / 64, 3),
            "waves_per_simd": round(d["waves_per_simd"], 2),
            "l2_hit": round(d["l2_hit"], 3) if d.get("l2_hit") is not None else None,
            "salu_per_valu": round(d["salu_per_valu"], 3),
            "lane_valu_insts_per_node_visit_or_test": round(lane_insts / (visits + tests), 1),
        },
        "memory": memory,
        "binding": binding,
    })
    return out


def bench_batching(runner, n_views, depth):
    """64 uneven views of the same scene in one batch (shared launches, one
    captured pass graph) against the same views rendered one at a time."""
    from hptracing_amd import serving

    views = serving.batch_views(n_views, seed=7, depth=depth)
    runner.render_batch(views[:2])  # warm the path
    t0 = time.perf_counter()
    runner.render_batch(views)
    batch_s = time.perf_counter() - t0
    rays = int(runner.last_stats.rays_cast)
    t0 = time.perf_counter()
    seq_rays = 0
    for v in views:
        runner.render(v)
        seq_rays += int(runner.last_stats.rays_cast)
    seq_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    for v in views:
        runner.render(v, pass_at_a_time=True)
    pat_s = time.perf_counter() - t0
    log(f"[bench] batching {n_views} views: batch {batch_s:.2f}s ({rays / batch_s / 1e6:.1f} "
        f"Mrays/s), one at a time {seq_s:.2f}s ({seq_rays / seq_s / 1e6:.1f} Mrays/s), one at a "
        f"time with one sample pass per launch sequence {pat_s:.2f}s")
    return {"views": n_views, "batch_s": round(batch_s, 3), "sequential_s": round(seq_s, 3),
            "batch_mrays_s": round(rays / batch_s / 1e6, 2),
            "sequential_mrays_s": round(seq_rays / seq_s / 1e6, 2),
            "speedup": round(seq_s / batch_s, 3),
            "sequential_pass_at_a_time_s": round(pat_s, 3),
            "speedup_vs_pass_at_a_time": round(pat_s / batch_s, 3)}


def bench_serving(args):
    """Serving: Poisson arrivals of the request mix on the 2M-triangle scene,
    p50/p99 latency per arrival rate, and served throughput against the same
    requests rendered one at a time (identical images, so equal error)."""
    from hptracing_amd import scenes, serving
    from hptracing_amd.runner import TraceRunner
    from hptracing_amd.scene import KDTree

    t0 = time.time()
    path = scenes.cached(args.scene_dir, scenes.write_full, 0, water_side=1000)
    kd = KDTree(path, os.path.dirname(path) + "/")
    log(f"[bench] serving scene: {kd.n_triangles} tris, KD {kd.build_seconds:.1f}s, "
        f"total {time.time() - t0:.1f}s")
    runner = TraceRunner(kd)
    scene_bytes = int(runner.stats["device_bytes"])
    log(f"[bench] serving scene on the device: {scene_bytes / 1e9:.2f} GB")
    views = serving.request_views(args.serve_requests, seed=3, depth=args.depth)
    runner.render(views[0])  # warm
    # The launch forms are tuned on this scene's requests (the library's
    # defaults are tuned on the 1080p scene): hpt_autotune on the largest
    # flat view of the mix (a 1280x720 one), then on a panorama.  Both the served and the
    # one-at-a-time runs below use the result.
    t1 = time.time()
    flat = [v for v in views if v.height > 0]
    pano = [v for v in views if v.height < 0]
    tuned = runner.get_tuning()
    if flat:
        tuned = runner.autotune(max(flat, key=lambda v: v.sample_x * v.sample_y))
    if pano:
        tuned = runner.autotune(pano[0])
    log(f"[bench] serving scene autotuned in {time.time() - t1:.1f}s: {tuned}")
    seq_s, seq_rays = serving.one_at_a_time(runner, views)
    seq_mrays = seq_rays / seq_s / 1e6
    log(f"[bench] one at a time: {len(views)} requests {seq_s:.2f}s {seq_mrays:.1f} Mrays/s")
    pat_s, _ = serving.one_at_a_time(runner, views, pass_at_a_time=True)
    log(f"[bench] one at a time, one sample pass per launch sequence: {pat_s:.2f}s")
    # The same requests as one hpt_batch_render: every request in the same
    # launches from the start, the best packing a server could reach on
    # them, so seq_s / packed_s bounds served against one at a time.
    t1 = time.perf_counter()
    runner.render_batch(views)
    packed_s = time.perf_counter() - t1
    packed_mrays = int(runner.last_stats.rays_cast) / packed_s / 1e6
    log(f"[bench] the mix as one batch: {packed_s:.2f}s {packed_mrays:.1f} Mrays/s "
        f"(ceiling of served vs one at a time {seq_s / packed_s:.3f})")
    rates = []
    for rate in args.serve_rates:
        r = serving.serve(runner, views, rate, seed=11)
        log(f"[bench] served at {rate}/s: p50 {r.p50_ms:.1f} ms p99 {r.p99_ms:.1f} ms, "
            f"{r.mrays_s:.1f} Mrays/s over {r.wall_s:.2f}s, {r.macro_passes} passes")
        rates.append({"arrivals_per_s": rate, "p50_ms": round(r.p50_ms, 2),
                      "p99_ms": round(r.p99_ms, 2), "mean_ms": round(r.mean_ms, 2),
                      "wall_s": round(r.wall_s, 3), "mrays_s": round(r.mrays_s, 2)})
    burst = serving.serve(runner, views, 0.0, seed=11)
    log(f"[bench] served all at once: {burst.wall_s:.2f}s {burst.mrays_s:.1f} Mrays/s, "
        f"p99 {burst.p99_ms:.1f} ms")
    return {"scene_triangles": kd.n_triangles, "scene_device_bytes": scene_bytes,
            "requests": len(views), "tuning": tuned,
            "one_at_a_time_s": round(seq_s, 3), "one_at_a_time_mrays_s": round(seq_mrays, 2),
            "served_burst_s": round(burst.wall_s, 3),
            "served_burst_mrays_s": round(burst.mrays_s, 2),
            "served_vs_one_at_a_time": round(seq_s / burst.wall_s, 3),
            "packed_batch_s": round(packed_s, 3), "packed_batch_mrays_s": round(packed_mrays, 2),
            "served_ceiling_vs_one_at_a_time": round(seq_s / packed_s, 3),
            "one_at_a_time_pass_at_a_time_s": round(pat_s, 3),
            "served_vs_one_at_a_time_pass_at_a_time": round(pat_s / burst.wall_s, 3),
            "burst_p50_ms": round(burst.p50_ms, 2), "burst_p99_ms": round(burst.p99_ms, 2),
            "poisson": rates}


def serve_profile(args):
    """The serving burst alone, for `rocprofv3 --kernel-trace`: the 40-request
    mix on the 2.2M-triangle scene submitted at once, twice, with a pause of
    one second before the measured second burst (the trace's last cluster of
    kernels)."""
    from hptracing_amd import scenes, serving
    from hptracing_amd.runner import TraceRunner
    from hptracing_amd.scene import KDTree
    path = scenes.cached(args.scene_dir, scenes.write_full, 0, water_side=1000)
    kd = KDTree(path, os.path.dirname(path) + "/")
    runner = TraceRunner(kd)
    views = serving.request_views(args.serve_requests or 40, seed=3, depth=args.depth)
    runner.render(views[0])
    serving.serve(runner, views, 0.0, seed=11)
    time.sleep(1.0)
    b = serving.serve(runner, views, 0.0, seed=11)
    log(f"[bench] serve profile: burst {b.wall_s:.3f}s, {b.mrays_s:.1f} Mrays/s, "
        f"p50 {b.p50_ms:.0f} ms, p99 {b.p99_ms:.0f} ms, {b.macro_passes} macro passes")


def bench_cast_only(n_tris):
    """hpt_cast alone on the ray-cast workload: Mrays/s over the median of
    five launches, and algorithmic bytes (the render's per-unit figures) from
    the counted walk of the same rays."""
    import numpy as np
    import torch

    from hptracing_amd import scenes
    from hptracing_amd.runner import TraceRunner
    from hptracing_amd.scene import KDTree, material

    t0 = time.time()
    pts, mesh = scenes.soup(n_tris, seed=1)
    kd = KDTree.from_arrays(pts, mesh, np.array([material(kd=(0.5, 0.5, 0.5))]))
    p = kd.points[:, :3]
    n = 1 << 22
    org, d = scenes.random_rays(n, p.min(0), p.max(0), seed=41)
    runner = TraceRunner(kd)
    org_t = torch.as_tensor(org).cuda()
    dir_t = torch.as_tensor(d).cuda()
    runner.cast(org_t, dir_t)
    torch.cuda.synchronize()
    times = []
    for _ in range(5):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        runner.cast(org_t, dir_t)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    ms = float(np.median(times))
    # The same batch after hpt_autotune_cast (the defaults are tuned for the
    # render's rays).
    chosen = runner.autotune_cast(org_t, dir_t)
    tuned = []
    for _ in range(5):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        runner.cast(org_t, dir_t)
        e1.record()
        torch.cuda.synchronize()
        tuned.append(e0.elapsed_time(e1))
    ms_tuned = float(np.median(tuned))
    _, _, cnt = runner.cast(org_t, dir_t, counted=True)
    torch.cuda.synchronize()
    c = cnt.sum(dim=0, dtype=torch.int64).cpu().numpy()
    algo = 32 * n + 40 * int(c[0]) + 4 * int(c[1]) + 56 * int(c[2])
    gbs = algo / (ms * 1e-3) / 1e9
    log(f"[bench] cast only: {kd.n_triangles} tris ({len(kd.leaf_data)} leaf ints), {n} rays, "
        f"{ms:.2f} ms, {n / ms / 1e3:.1f} Mrays/s, {algo / n:.0f} algorithmic B/ray, "
        f"{gbs:.0f} GB/s; autotuned {n / ms_tuned / 1e3:.1f} Mrays/s; setup {time.time() - t0:.1f}s")
    return {"workload": f"hpt_cast, {n} rays through a {kd.n_triangles}-triangle soup",
            "stand_in_for": ("the 1 000 000-triangle soup: its exact KD tree holds "
                             "9 321 406 973 leaf ints (hpt_host_kd_count), past the int32 "
                             "offsets of the reference's node format; 600 000 triangles "
                             "(1 664 622 455) is the largest size that fits"),
            "triangles": kd.n_triangles, "rays": n, "ms": round(ms, 3),
            "mrays_s": round(n / ms / 1e3, 2),
            "node_visits_per_ray": round(int(c[0]) / n, 1),
            "leaves_per_ray": round(int(c[1]) / n, 1),
            "tri_tests_per_ray": round(int(c[2]) / n, 1),
            "algorithmic_bytes": algo, "achieved_gbs": round(gbs, 1),
            "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
            "autotuned_ms": round(ms_tuned, 3),
            "autotuned_mrays_s": round(n / ms_tuned / 1e3, 2),
            "autotuned_tuning": chosen}


def cpu_baseline(kd, view, spp, target_s, R):
    """The oracle's float32 render (the reference kernels restated in C, one
    path per thread) on the host cores, on a stratified sample of the same
    view's camera rays: 32 bands of consecutive camera rays, one in the
    middle of each 32nd of the frame, every band at the full spp and depth,
    in one call of the oracle's render_rays.
    The sample grows until one run takes about target_s / 3; the last
    sizing run is the warm-up, then the median of three timed runs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    ov = O.make_view(view=view["view_p"], up=view["top_dir"], right=view["right_dir"],
                     width=view["width"], height=view["height"], sample_x=view["sample_x"],
                     sample_y=view["sample_y"], supersample_x=view["supersample_x"],
                     supersample_y=view["supersample_y"], background=view["background_color"],
                     sample=spp, depth=view["depth"], disable_diffuse=0, brightness=1.0, seed=1)
    import numpy as np
    s = O.Scene(kd)
    bands = 32
    stratum = R // bands
    accum = np.zeros(3 * R, np.float32)

    def run(band):
        first = np.arange(bands, dtype=np.int64) * stratum + (stratum - band) // 2
        picks = (first[:, None] + np.arange(band, dtype=np.int64)[None, :]).reshape(-1)
        t0 = time.perf_counter()
        _, rays = s.render_rays(ov, picks, cores, accum=accum)
        return rays, time.perf_counter() - t0

    per_run = target_s / 3.0
    band = max(8, cores)
    while True:
        rays, dt = run(band)              # sizing; the last one is the warm-up
        if dt >= per_run * 0.6 or band >= stratum:
            break
        band = min(stratum, int(band * max(2.0, per_run / max(dt, 1e-3))))
    runs = [run(band) for _ in range(3)]
    rates = sorted(r / t for r, t in runs)
    median = rates[1]
    n = band * bands
    log(f"[bench] cpu baseline: {bands} bands x {band} camera rays x {spp} spp on {cores} "
        f"threads, runs {[f'{r / t / 1e6:.3f}' for r, t in runs]} Mrays/s, median "
        f"{median / 1e6:.3f}")
    return {"value": round(median / 1e6, 4), "unit": "Mrays/s", "cores": cores,
            "kind": "port",
            "sample": f"{n} camera rays ({100.0 * n / R:.1f} % of the view) in {bands} bands "
                      f"spread evenly over the frame, {spp} spp, depth {view['depth']}: "
                      f"{runs[0][0]} rays per run; median of 3 runs after a warm-up "
                      f"({', '.join(f'{r / t / 1e6:.3f}' for r, t in runs)} Mrays/s)"}


if __name__ == "__main__":
    main()
