#!/usr/bin/env python3
"""Benchmark: D3Q19 TRT (Lambda = 1/4) body-force flow through a 256^3
periodic porous medium (overlapping spheres, radii U[6,12], 30 % solid,
seed 2024), tau = 0.9, g = 1e-6 along x, FP32 storage and compute.

One step = one fused stream-collide-boundary kernel over the whole lattice,
populations resident in HBM.  Every timed configuration first runs an untimed
settle phase (--settle device seconds, at least --warmup steps), then the
timed window of --steps launches, which write no FP64 macro fields.  Prints
one JSON line:
  value      MLUPS = all nodes x steps / time / 1e6 (the reference's convention,
             T/Resources/CGPerformance.xlsx sheet3 B6-B7)
  roofline   algorithmic bytes per launch / average launch time (HIP events on
             the solver's stream) against 8 TB/s
  median_launch_ms  median of a second window timed per launch
  f16_storage, bf16_storage  the same step with 16-bit deviation storage (FP32
             compute) and its ratio to the FP32 line
  mrt_f32, mrt_f16_storage   the workload's MRT collision
  cpu_baseline  the FP64 CPU oracle on the same lattice, a few steps on the
             host threads of one GPU and on one thread
  warm_start    wall time to a 1e-6 residual from rest vs from the learned start
                on held-out 512 x 128 channels (inference included)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_GBS = 8000.0


def algorithmic_bytes(n_nodes, n_fluid, q, sbytes):
    """Per launch: every fluid node reads and writes each population once,
    plus one byte of node class per node."""
    return n_fluid * 2 * q * sbytes + n_nodes * 1


def traffic_from_profiles(name):
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as fh:
            return json.load(fh).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cls, bb, n, seconds_target=12.0, seconds_one=8.0):
    """FP64 oracle (a C restatement of the reference step, OpenMP over the
    host cores) on the same lattice: steps until ~seconds_target elapse on the
    host threads the GPU box gives one GPU (at most 16), then on one thread."""
    from oracle import oracle as O
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = min(avail, 16)

    def rate(nthreads, target, max_steps):
        s = O.Sim(19, (n, n, n), cls, bb, collision=O.TRT, omega=1 / 0.9, trt_magic=0.25,
                  force_model=O.FORCE_GUO, g=(1e-6, 0, 0), periodic=(1, 1, 1), nthreads=nthreads)
        s.init(0)
        s.step(1)   # warm-up (first touch)
        steps, t0 = 0, time.perf_counter()
        while True:
            s.step(1)
            steps += 1
            el = time.perf_counter() - t0
            if el > target or steps >= max_steps:
                break
        return n ** 3 * steps / el / 1e6, steps, el

    mlups, steps, el = rate(threads, seconds_target, 50)
    one, steps1, el1 = rate(1, seconds_one, 10)
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "one_core_mlups": round(one, 3), "nproc": os.cpu_count(), "affinity_cpus": avail,
            "cpu_model": cpu_model(),
            "sample": f"FP64 oracle, D3Q19 TRT on the same {n}^3 lattice: {steps} steps after 1 warm-up step "
                      f"({el:.1f} s) on {threads} OpenMP threads (the host share of one GPU, of {avail} CPUs "
                      f"visible), and {steps1} steps ({el1:.1f} s) on 1 thread"}


def warm_start(n_channels=8, seed=9001):
    """Wall time to a 1e-6 residual (checked every 100 steps, FP64 D2Q9 TRT,
    512 x 128 channels with 1-6 obstacles) from rest and from the learned
    start, inference and injection included, on held-out channels; single
    lattice runs through the C ABI."""
    from thesis_amd import lattices as LT
    from thesis_amd import surrogate as S
    path = os.path.join(ROOT, "thesis_amd", "surrogate_weights.npz")
    if not os.path.exists(path):
        return None
    net = S.Net.load(path)
    test = LT.surrogate_channels(n_channels, seed)
    S.warm_start_single(net, test[0])[2].close()          # first-call set-up outside the timing
    tr, tw, nr, nw = [], [], [], []
    for d in test:
        n, t, s = S.rest_start_single(d)
        s.close()
        tr.append(t)
        nr.append(n)
        n, t, s = S.warm_start_single(net, d)
        s.close()
        tw.append(t)
        nw.append(n)
    return {"channels": n_channels, "seed": seed, "rest_s": round(sum(tr), 3), "warm_s": round(sum(tw), 3),
            "speedup": round(sum(tr) / sum(tw), 3), "steps_rest": nr, "steps_warm": nw,
            "note": "wall time to 1e-6, learned start includes inference and f_eq injection"}


def batch64():
    """The 64-channel sweep (128x32 .. 1024x256, FP64 TRT) to 1e-6 in one
    persistent launch."""
    from thesis_amd import lattices as LT
    from thesis_amd import lbm as L
    sweep = LT.channel_sweep()
    b = L.Batch(sweep, collision=L.TRT, storage="f64")
    steps, res, ms = b.run(1e-6, every=100, max_steps=400000)
    b.close()
    work = sum(int(steps[i]) * sweep[i]["nx"] * sweep[i]["ny"] for i in range(len(sweep)))
    fl = sum(int(steps[i]) * int((sweep[i]["node_class"] != L.SOLID).sum()) for i in range(len(sweep)))
    gbs = (fl * 2 * 9 * 8 + work * 1) / (ms / 1e3) / 1e9
    return {"ms": round(ms, 1), "glups": round(work / (ms / 1e3) / 1e9, 2), "gbs": round(gbs, 1),
            "steps_min": int(steps.min()), "steps_max": int(steps.max()), "max_residual": float(res.max())}


def colour_gradient_3d(n=256, steps=20, settle=0.3):
    """The D3Q19 colour-gradient step (one pass over z-columns of row tiles,
    phase field in LDS) on an n^3 periodic box with a red sphere of radius 40
    in blue (rest fractions 1/3, A = 0.01125, beta = 0.7, tau = 1, 18-point
    gradient), FP32 and FP16 storage.  Algorithmic bytes per step: one read and one write of
    the 38 stored values per node."""
    from thesis_amd import lattices as LT
    from thesis_amd import lbm as L
    r, b = LT.droplet3d(n, n, n, 40)
    res = {}
    for name, st, sb in (("f32", L.STORE_F32, 4), ("f16", L.STORE_F16, 2)):
        g = L.CgSolver(n, n, nz=n, storage=st, A=0.01125, r_alpha=1 / 3, b_alpha=1 / 3, high_order=0)
        g.init(r, b)
        dev = 0.0
        while dev < settle:
            dev += g.time_steps(20) / 1e3
        ms = g.time_steps(steps) / steps
        red = g.reduce()
        g.close()
        alg = n ** 3 * 2 * 38 * sb
        res[name] = {"ms_per_step": round(ms, 4), "mlups": round(n ** 3 / (ms / 1e3) / 1e6, 1),
                     "gbs": round(alg / (ms / 1e3) / 1e9, 1), "frac": round(alg / (ms / 1e3) / 1e9 / PEAK_GBS, 4),
                     "finite": bool(red["nonfinite"] == 0)}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--storage", default="f32", choices=["f64", "f32", "f16", "bf16"])
    ap.add_argument("--collision", default="trt", choices=["bgk", "trt", "mrt"])
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--variants", action="store_true", help="also time BGK and the 64-channel batch")
    ap.add_argument("--no-variants", action="store_true", help="skip the BF16 and MRT lines")
    ap.add_argument("--settle", type=float, default=0.5,
                    help="untimed device seconds before each timed window (steady clocks)")
    ap.add_argument("--no-warm-start", action="store_true")
    ap.add_argument("--no-f16", action="store_true", help="skip the FP16-storage line of the f32 run")
    ap.add_argument("--batch-only", action="store_true", help="only the 64-channel batch line (for profiling)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py: this build runs on one GPU; --gpus above 1 is not supported "
              "(the multi-GPU slab decomposition is a design note only).", file=sys.stderr)
        sys.exit(2)

    if args.batch_only:
        print(json.dumps({"batch64": batch64()}), flush=True)
        return

    from thesis_amd import lattices as LT
    from thesis_amd import lbm as L

    n = args.size
    cls, bb, info = LT.porous3d(n, seed=args.seed)
    n_nodes = n ** 3
    n_fluid = info["n_fluid"]
    coll = {"bgk": L.BGK, "trt": L.TRT, "mrt": L.MRT}[args.collision]
    sbytes = {"f64": 8, "f32": 4, "f16": 2, "bf16": 2}[args.storage]

    def make(storage, collision):
        s = L.Solver(19, (n, n, n), cls, bb, collision=collision, omega=1 / 0.9, trt_magic=0.25,
                     force_model=L.FORCE_GUO, g=(1e-6, 0.0, 0.0), periodic=(1, 1, 1), storage=storage)
        s.init(0)
        return s

    def settle(s):
        """Untimed steps until args.settle seconds of device time have run, so
        the timed window starts at steady clocks; at least args.warmup steps."""
        steps, dev = args.warmup, s.time_steps(args.warmup) / 1e3 if args.warmup else 0.0
        while dev < args.settle:
            dev += s.time_steps(100) / 1e3
            steps += 100
        return {"steps": steps, "device_s": round(dev, 3)}

    def line(storage, collision):
        """One timed configuration: settle, the timed window (mean per step),
        then a window of per-launch events (median per launch)."""
        s = make(storage, collision)
        st = settle(s)
        t0 = time.perf_counter()
        ms = s.time_steps(args.steps)          # HIP events on the solver's stream
        wall = time.perf_counter() - t0
        each = s.time_launches(min(max(args.steps, 20), 200))
        red = s.reduce()
        s.close()
        sb = {"f64": 8, "f32": 4, "f16": 2, "bf16": 2}[storage]
        avg = ms / args.steps
        return {"ms": ms, "avg_ms": avg, "wall": wall, "settle": st,
                "median_launch_ms": float(np.median(each)), "mean_launch_ms": float(np.mean(each)),
                "mlups": n_nodes / (avg / 1e3) / 1e6,
                "gbs": algorithmic_bytes(n_nodes, n_fluid, 19, sb) / (avg / 1e3) / 1e9, "mass": red["mass"]}

    main_line = line(args.storage, coll)
    ms, avg_ms, wall = main_line["ms"], main_line["avg_ms"], main_line["wall"]
    mlups = main_line["mlups"]
    bytes_launch = algorithmic_bytes(n_nodes, n_fluid, 19, sbytes)
    achieved = bytes_launch / (avg_ms / 1e3) / 1e9

    def brief(d, ref=None):
        o = {"mlups": round(d["mlups"], 1), "ms_per_step": round(d["avg_ms"], 5),
             "median_launch_ms": round(d["median_launch_ms"], 5), "gbs": round(d["gbs"], 1),
             "frac": round(d["gbs"] / PEAK_GBS, 4)}
        if ref is not None:
            o["vs_f32"] = round(d["mlups"] / ref, 3)
        return o

    variants = {}
    if args.storage == "f32" and args.collision == "trt" and not args.no_f16:
        # the 16-bit targets are stated against the FP32 line; MRT is the
        # second collision of the workload
        f16 = line("f16", L.TRT)
        variants["f16_storage"] = brief(f16, mlups)
        if not args.no_variants:
            variants["bf16_storage"] = brief(line("bf16", L.TRT), mlups)
            mrt32 = line("f32", L.MRT)
            variants["mrt_f32"] = brief(mrt32)
            variants["mrt_f16_storage"] = brief(line("f16", L.MRT), mrt32["mlups"])
            variants["colour_gradient_d3q19_256"] = colour_gradient_3d(n)
    if args.variants:
        variants["bgk_f32"] = brief(line("f32", L.BGK))
        variants["batch64"] = batch64()

    out = {
        "metric": "MLUPS & % HBM roofline (D3Q19 256^3); wall-time to 1e-6 residual vs plain",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(avg_ms, 5),
        "wall_s": round(wall, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32"}.get(args.storage, args.storage + "-storage/f32-compute"),
        "data": "synthetic",
        "config": {"workload": f"D3Q19 {args.collision.upper()} Lambda=1/4 body-force flow, {n}^3 periodic porous "
                               f"medium (spheres r~U[6,12], seed {args.seed}, solid {info['solid_fraction']:.4f}), "
                               f"tau=0.9, g=1e-6 x, {args.storage} storage",
                   "nodes": n_nodes, "fluid_nodes": n_fluid, "lattice": "D3Q19", "storage": args.storage},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(f"traffic_d3q19_{args.collision}_{args.storage}_{n}.json")},
        "mflups": round(n_fluid * args.steps / (ms / 1e3) / 1e6, 1),
        "mass": main_line["mass"],
        "settle": main_line["settle"],
        "median_launch_ms": round(main_line["median_launch_ms"], 5),
        "mean_launch_ms": round(main_line["mean_launch_ms"], 5),
        "timing_note": "value/ms_per_step: HIP events around the timed steps on the solver's stream after "
                       "an untimed settle phase; the timed launches write no FP64 macro fields (an untimed "
                       "moments pass follows); median/mean_launch_ms: a second window with one event pair "
                       "per launch",
    }
    out.update(variants)
    if not args.no_warm_start:
        out["warm_start"] = warm_start()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cls, bb, n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
