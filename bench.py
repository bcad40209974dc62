"""Benchmark of the D3Q19 collide-and-stream hot path on one MI355X.

A step is one fused pull-scheme collide+stream pass over every site of one
lattice resident in HBM.  The default workload is the synthetic symmetric
Y-bifurcation (parent r 24 L 200, daughters r 19 L 160 at +-30 deg, seeded
radius jitter), TRT with magic parameter 3/16 and tau 0.6, BFL walls, a Ladd
parabolic velocity inlet driven by a sinusoidal peak-speed table (peak 0.05,
period 4000) and two Nash pressure outlets at rho = 1, FP64 storage.

Prints one JSON line (rank 0, N = 1).  MFLUPS = sites x steps / seconds / 1e6.
The roofline object prices the step kernel against MI355X's 8 TB/s HBM with
the algorithmic byte count per site update (FP64: 19 reads + 19 writes of 8 B
plus 18 pull indices of 4 B = 376 B).  cpu_baseline times the CPU oracle (the
reference's push/AoS FP64 algorithm, OpenMP over site ranges) on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_SITE = {"f64": 19 * 8 * 2 + 18 * 4}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_case(workload: str, steps_total: int):
    import hemepure_amd as H
    from fixtures import bifurcation_case, cylinder_case

    if workload == "bifurcation":
        return bifurcation_case(steps=steps_total), (
            "Y-bifurcation: parent r24 L200, daughters r19 L160 at +-30deg, jitter 3% seed 2026; TRT "
            "lambda 3/16 tau 0.6; BFL walls; Ladd parabolic inlet, sinusoidal peak 0.05 period 4000; "
            "2 Nash outlets rho 1.0")
    if workload == "cylinder":
        return cylinder_case(), ("straight cylinder r10 L100 in 24x24 box; LBGK tau 0.8; SBB; Nash "
                                 "cosine rho 1.001 / 0.999")
    raise SystemExit(f"unknown workload {workload}")


def cpu_baseline(case, seconds_target: float = 15.0):
    import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    m = case.oracle()
    m.init_equilibrium(1.0)
    m.run(1, threads=threads)
    t0 = time.perf_counter()
    m.run(2, threads=threads)
    per_step = (time.perf_counter() - t0) / 2
    steps = max(2, min(200, int(seconds_target / max(per_step, 1e-6))))
    t0 = time.perf_counter()
    m.run(steps, threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": m.n * steps / dt / 1e6,
        "unit": "MFLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} steps of the same workload ({m.n} sites), FP64 push/AoS with an int64 "
                  f"neighbour table, {dt:.1f} s",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="bifurcation", choices=["bifurcation", "cylinder"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    import hemepure_amd as H

    total = args.steps + args.warmup + 8
    case, desc = make_case(args.workload, total)
    log(f"workload {args.workload}: {case.geom.n_sites} sites")
    lat = case.product()
    lat.init_equilibrium(1.0)
    lat.step(max(1, args.warmup))
    lat.synchronize()
    log(f"warmup {args.warmup} steps done")

    # Timed region: HIP events recorded on the stream the kernels run on.
    ms_total = 0.0
    left = args.steps
    wall0 = time.perf_counter()
    while left > 0:
        chunk = min(left, 1024)
        ms_total += lat.step_timed(chunk)
        left -= chunk
    wall = time.perf_counter() - wall0
    lat.synchronize()
    if not lat.stable():
        log("warning: the state went unstable")
    n = lat.n
    sec = ms_total / 1e3
    mflups = n * args.steps / sec / 1e6
    per_launch_s = sec / args.steps
    bytes_per_launch = n * BYTES_PER_SITE["f64"]
    achieved = bytes_per_launch / per_launch_s / 1e9
    log(f"{args.steps} steps in {ms_total:.2f} ms (host wall {wall*1e3:.2f} ms): {mflups:.1f} MFLUPS")

    out = {
        "metric": "MFLUPS",
        "value": round(mflups, 2),
        "unit": "MFLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (procedurally generated seeded geometry)",
        "config": {"workload": desc, "sites": n, "storage": "fp64 SoA, 32-bit pull table",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_site_update": BYTES_PER_SITE["f64"], "kernel": "step_fp64"},
    }
    if not args.no_cpu_baseline:
        log("timing the CPU oracle on a bounded sample")
        out["cpu_baseline"] = cpu_baseline(case)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
