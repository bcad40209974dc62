"""Benchmark of the D3Q19 collide-and-stream hot path on one MI355X.

A step is one fused pull-scheme collide+stream pass over every site of one
lattice resident in HBM.  The default workload is the synthetic symmetric
Y-bifurcation (parent r 24 L 200, daughters r 19 L 160 at +-30 deg, seeded
radius jitter), TRT with magic parameter 3/16 and tau 0.6, BFL walls, a Ladd
parabolic velocity inlet driven by a sinusoidal peak-speed table (peak 0.05,
period 4000) and two Nash pressure outlets at rho = 1, FP64 storage.

The same workload is then timed in the FP32-shifted storage form (float
g = f - w rho0), reported under "forms".

Prints one JSON line (rank 0, N = 1).  MFLUPS = sites x steps / seconds / 1e6.
The roofline object prices the step kernel against MI355X's 8 TB/s HBM with
the algorithmic byte count of a pull step per site update: 19 reads + 19
writes of the distributions plus an 18 x 4 B neighbour index, 376 B in FP64
and 224 B in FP32.  This layout finds neighbours arithmetically and reads no
index, so it moves 304 B / 152 B of distributions per update
("layout_bytes_per_site_update"); "traffic" is the measured bytes per
launch.  cpu_baseline times the CPU oracle (the
reference's push/AoS FP64 algorithm, OpenMP over site ranges) on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import re
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

HBM_PEAK_GBS = 8000.0
# algorithmic bytes of a pull step per site update: distributions read and
# written plus an 18 x 4 B neighbour index
BYTES_PER_SITE = {"f64": 19 * 8 * 2 + 18 * 4, "f32": 19 * 4 * 2 + 18 * 4}
# what this layout itself moves: the distributions only (no index table)
LAYOUT_BYTES_PER_SITE = {"f64": 19 * 8 * 2, "f32": 19 * 4 * 2}


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


CAL_BYTES = 1 << 31  # 2 GiB copied per calibration launch


def calibrate_copy() -> int:
    """A plain streaming kernel over 2 GiB (read 2 GiB, write 2 GiB), five
    times: the reference whose FETCH_SIZE / WRITE_SIZE readings scale the
    counters.  An elementwise multiply, not copy_, so it is a kernel and not a
    DMA transfer."""
    import torch

    src = torch.zeros(CAL_BYTES // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    return 0


def _counter_means(path: Path) -> dict:
    sums, counts = {}, {}
    with open(path) as fh:
        for r in csv.DictReader(fh):
            k = r["Kernel_Name"]
            sums[k] = sums.get(k, 0.0) + float(r["Counter_Value"])
            counts[k] = counts.get(k, 0) + 1
    return {k: sums[k] / counts[k] for k in sums}


def measure_traffic(workload: str, variant: int = 1) -> dict:
    """HBM bytes per step launch from two rocprofv3 --pmc passes, each scaled
    by the same counter's reading for the 2 GiB calibration copy.  Runs the
    profiled programs as child processes; this process never touches the GPU."""
    out = {}
    tmp = Path(os.environ.get("TMPDIR", "/tmp")) / f"hlb_pmc_{os.getpid()}"
    me = [sys.executable, "-u", str(Path(__file__).resolve())]
    runs = {"bench": me + ["--steps", "10", "--warmup", "2", "--no-cpu-baseline", "--no-ensemble", "--no-serving",
                           "--no-vessel-tree", "--workload", workload, "--variant", str(variant)],
            "cal": me + ["--calibrate"]}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        for tag, cmd in runs.items():
            d = tmp / f"{tag}_{counter}"
            full = ["rocprofv3", "--pmc", counter, "-f", "csv", "-d", str(d), "-o", tag, "--"] + cmd
            log(f"pmc pass: {' '.join(full)}")
            subprocess.run(["timeout", "-s", "KILL", "180"] + full, check=True, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL)
            out[(tag, counter)] = _counter_means(next(d.rglob("*counter_collection.csv")))
    res = {}
    for kernel in ("step_fp64", "step_fp32"):
        r = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            cal = out[("cal", counter)]
            copy_kib = max(v for k, v in cal.items() if "copy" in k.lower() or "elementwise" in k.lower())
            factor = (CAL_BYTES / 1024.0) / copy_kib
            # the steady-state launches: not the first step (reads F), not a property step
            # FP64: the template without the first-step / property flags; FP32's
            # flags are run-time, its template arguments are the launch shape
            def steady(name):
                args = [x.strip() for x in name.split("<", 1)[1].split(">", 1)[0].split(",")]
                return kernel == "step_fp32" or (args[1] == "false" and args[2] == "false")

            step = [v for k, v in out[("bench", counter)].items() if kernel in k and steady(k)]
            r[counter] = {"raw_kib": max(step), "factor": factor, "bytes": max(step) * 1024.0 * factor}
        r["bytes_per_launch"] = r["FETCH_SIZE"]["bytes"] + r["WRITE_SIZE"]["bytes"]
        res[kernel] = r
    res["bytes_per_launch"] = res["step_fp64"]["bytes_per_launch"]
    return res


def _committed_traffic(workload: str, n: int):
    """The bytes per launch recorded in profiles/ for this workload and site
    count by an earlier --measure-traffic run, or None."""
    path = ROOT / "profiles" / f"step_fp64_{workload}_traffic.md"
    if not path.exists():
        return None
    text = path.read_text()
    m_sites = re.search(r"sites:\s*(\d+)", text)
    m_bytes = re.search(r"bytes_per_launch:\s*(\d+)", text)
    if not (m_sites and m_bytes) or int(m_sites.group(1)) != n:
        return None
    return int(m_bytes.group(1))


def time_steps(lat, steps: int):
    """Milliseconds of `steps` steps, from HIP events on the lattice's stream."""
    ms_total = 0.0
    left = steps
    while left > 0:
        chunk = min(left, 1024)
        ms_total += lat.step_timed(chunk)
        left -= chunk
    return ms_total


def time_form(case, precision, steps: int, warmup: int, variant=None):
    lat = case.product(precision=precision)
    lat.init_equilibrium(1.0)
    lat.step(max(1, warmup))
    if variant is None:
        variant, _ = lat.autotune(10)
    else:
        lat.set_variant(variant)
    lat.synchronize()
    ms = time_steps(lat, steps)
    lat.synchronize()
    stable = lat.stable()
    n = lat.n
    lat.close()
    return n, ms, stable, variant


def percentile(xs, q):
    import numpy as np

    return float(np.percentile(np.asarray(xs, dtype=np.float64), q))


def bench_ensemble(members: int, steps: int, solve_steps: int, workers: int):
    """The stenosed-pipe ensemble: `members` pipes (r U(6,24), L U(60,400),
    stenosis 0-70%, tau U(0.55,0.9), inlet amplitude U(1e-4,1e-3), seed 2026),
    BFL + TRT.  Batched form: every member FP32-shifted in one batch (one
    launch per step, graph-replayed, autotuned launch shape), timed over
    `steps` steps.  Plain form: each member alone in FP64, one after
    another, each timed over a short run.  A solve is `solve_steps` steps of
    one member; every member arrives at t = 0, so the completion times are
    all equal for the batch and cumulative for the plain loop."""
    from concurrent.futures import ThreadPoolExecutor

    import hemepure_amd as H
    from fixtures import ensemble_cases

    t0 = time.perf_counter()
    cases = ensemble_cases(members, seed=2026, workers=workers)
    with ThreadPoolExecutor(workers) as ex:
        lats = list(ex.map(lambda c: c.product(precision=H.FP32_SHIFTED), cases))
    sites = sum(l.n for l in lats)
    log(f"ensemble: {members} members, {sites} sites, built in {time.perf_counter() - t0:.1f} s")
    for l in lats:
        l.init_equilibrium(1.0)
    batch = H.Batch(lats)
    variant, _ = batch.autotune(10)
    batch.step_timed(min(steps, 1024))  # captures the timed sequence's graph
    ms = 0.0
    left = steps
    while left > 0:
        n = min(left, 1024)
        ms += batch.step_timed(n)
        left -= n
    stable = all(l.stable() for l in lats)
    ms_step = ms / steps
    batch.close()
    for l in lats:
        l.close()
    mflups_b = sites * steps / (ms / 1e3) / 1e6
    t_solve_b = solve_steps * ms_step / 1e3
    log(f"ensemble batched fp32: {ms_step:.3f} ms/step, {mflups_b:.1f} MFLUPS, variant {variant}"
        + ("" if stable else " (unstable member)"))

    with ThreadPoolExecutor(workers) as ex:
        plain = list(ex.map(lambda c: c.product(precision=H.FP64), cases))
    per_member = []
    for l in plain:
        l.init_equilibrium(1.0)
        l.step(2)
        per_member.append(l.step_timed(20) / 20)
        l.close()
    ms_plain = sum(per_member)
    done = [solve_steps * x / 1e3 for x in per_member]
    completion = [sum(done[:i + 1]) for i in range(len(done))]
    mflups_p = sites / (ms_plain / 1e3) / 1e6
    log(f"ensemble plain fp64 loop: {ms_plain:.3f} ms per step of all members, {mflups_p:.1f} MFLUPS")
    return {
        "workload": f"{members} stenosed pipes, r U(6,24) L U(60,400), stenosis 0-70% width 2R, tau "
                    "U(0.55,0.9), inlet amplitude U(1e-4,1e-3), seed 2026; TRT lambda 3/16, BFL, Nash cosine",
        "members": members, "sites": sites, "steps_timed": steps, "steps_per_solve": solve_steps,
        "batched_fp32": {"mflups": round(mflups_b, 2), "ms_per_step": ms_step,
                         "hbm_frac": round(sites * BYTES_PER_SITE["f32"] / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                         "solves_per_s": round(members / t_solve_b, 3), "p50_s": round(t_solve_b, 3),
                         "p99_s": round(t_solve_b, 3), "launch_variant": variant, "hip_graph": True},
        "plain_fp64_loop": {"mflups": round(mflups_p, 2), "ms_per_step_all_members": ms_plain,
                            "solves_per_s": round(members / completion[-1], 3),
                            "p50_s": round(percentile(completion, 50), 3),
                            "p99_s": round(percentile(completion, 99), 3)},
        "speedup_solves_per_s": round((members / t_solve_b) / (members / completion[-1]), 3),
    }


def bench_vessel_tree(steps: int, warmup: int, threads: int):
    """The sparse vessel tree (63 tubes, root r 40 tapering 0.79, branch
    angles 20-45 deg, box 1024 x 1024 x 768, LES around tau 0.505, WK3
    outlets, cosine pressure inlet), one instance
    timed over `steps` steps in the FP32-shifted form and the FP64 form.
    Voxelising the geometry is not timed."""
    import hemepure_amd as H
    from fixtures import vessel_tree_case

    t0 = time.perf_counter()
    case = vessel_tree_case(threads=threads)
    n = case.geom.n_sites
    log(f"vessel tree: {n} sites, {len(case.outlets)} WK3 outlets, voxelised in {time.perf_counter() - t0:.1f} s")
    out = {"workload": "binary vessel tree of 63 tubes, root r40 x0.79 per generation, length 6r, angles "
                       "U(20,45) deg, seed 2026, box 1024x1024x768; LBGK+LES (C_s 0.1) tau 0.505; SBB walls; "
                       "cosine pressure inlet rho 1.0001 +- 1e-4 period 1000; 32 WK3 outlets R1=R2=1e-3 C=1e9 "
                       "(C=10 empties the outlets within steps under the absolute-pressure WK3 update)",
           "sites": n, "box_fill": round(n / (1024 * 1024 * 768), 5), "steps_timed": steps}
    for key, prec, kind in (("fp32_shifted", H.FP32_SHIFTED, "f32"), ("fp64", H.FP64, "f64")):
        n_, ms, stable, variant = time_form(case, prec, steps, warmup)
        ms_step = ms / steps
        out[key] = {"mflups": round(n_ * steps / (ms / 1e3) / 1e6, 2), "ms_per_step": ms_step,
                    "hbm_frac": round(n_ * BYTES_PER_SITE[kind] / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                    "launch_variant": variant, "stable": bool(stable)}
        log(f"vessel tree {key}: {out[key]['mflups']} MFLUPS" + ("" if stable else " (unstable)"))
    return out


def bench_serving(count: int, rate: float, max_sites: float, workers: int, quantum: int = 500):
    """The serving stream: `count` requests (pipes, stenoses, bifurcations,
    20k..max_sites sites, LBGK or TRT) arriving as a Poisson process of
    `rate` per second (seeded), each run to periodic steady state
    (cycle-to-cycle max |du| < 1e-6, cosine period 2000, cap 40k steps) by the
    request loop under a 200 GB admission budget, FP32-shifted.  Geometry is
    voxelised before the clock starts (it is the client's input); table
    building, admission, stepping, convergence tests and retirement are timed.
    Plain: each request alone in FP64, one at a time in arrival order; its
    service time is the request's step count times its FP64 ms/step measured
    on the GPU, replayed against the same arrival times."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    import hemepure_amd as H
    from fixtures import serving_requests

    period, tol, cap = 2000, 1e-6, 40000
    cases = serving_requests(count, seed=7, max_sites=max_sites, workers=workers)
    rng = np.random.default_rng(8)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
    sites = [c.geom.n_sites for c in cases]
    log(f"serving: {count} requests, {sum(sites)} sites, arrivals over {arrivals[-1]:.1f} s")
    srv = H.Server(budget_bytes=200 << 30, quantum_steps=quantum)
    ids = [None] * count
    nxt = 0
    t0 = time.perf_counter()
    finished = 0
    while finished < count:
        now = time.perf_counter() - t0
        while nxt < count and arrivals[nxt] <= now:
            ids[nxt] = cases[nxt].submit(srv, max_steps=cap, period=period, tolerance=tol)
            nxt += 1
        run, q = srv.advance()
        if run == 0 and q == 0 and nxt < count:
            time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
        finished = sum(1 for i in ids if i is not None and srv.status(i).state >= H.REQ_CONVERGED)
    makespan = time.perf_counter() - t0
    st = [srv.status(i) for i in ids]
    tts = [s.t_done - s.t_submit for s in st]
    steps = [s.steps for s in st]
    converged = sum(1 for s in st if s.state == H.REQ_CONVERGED)
    site_steps = srv.site_steps()
    srv.close()
    log(f"serving: makespan {makespan:.1f} s, {converged}/{count} converged, p50 {percentile(tts, 50):.2f} s, "
        f"p99 {percentile(tts, 99):.2f} s")

    # Plain: FP64 alone, FIFO, from measured ms/step.
    with ThreadPoolExecutor(workers) as ex:
        plain = list(ex.map(lambda c: c.product(precision=H.FP64), cases))
    service = []
    for l, n in zip(plain, steps):
        l.init_equilibrium(1.0)
        l.step(2)
        service.append(n * l.step_timed(20) / 20 / 1e3)
        l.close()
    done, t = [], 0.0
    for a, s_ in zip(arrivals, service):
        t = max(t, a) + s_
        done.append(t - a)
    plain_makespan = max(a + d for a, d in zip(arrivals, done))
    return {
        "workload": f"{count} requests (straight pipes, stenoses, Y-bifurcations; 2e4..{max_sites:.0e} sites "
                    f"log-uniform; LBGK or TRT; BFL; cosine inlet period {period}), Poisson {rate}/s seed 8, "
                    f"to cycle-to-cycle max|du| < {tol} (cap {cap} steps), budget 200 GB, quantum {quantum}",
        "requests": count, "converged": converged, "site_steps": site_steps,
        "served_fp32": {"mflups": round(site_steps / makespan / 1e6, 2), "makespan_s": round(makespan, 3),
                        "solves_per_s": round(count / makespan, 4),
                        "p50_s": round(percentile(tts, 50), 3), "p99_s": round(percentile(tts, 99), 3)},
        "plain_fp64_fifo": {"mflups": round(sum(n * s for n, s in zip(sites, steps)) / plain_makespan / 1e6, 2),
                            "busy_mflups": round(sum(n * s for n, s in zip(sites, steps)) / sum(service) / 1e6, 2),
                            "makespan_s": round(plain_makespan, 3),
                            "solves_per_s": round(count / plain_makespan, 4),
                            "p50_s": round(percentile(done, 50), 3), "p99_s": round(percentile(done, 99), 3)},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="bifurcation", choices=["bifurcation", "cylinder"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--ensemble-members", type=int, default=128)
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-stream measurement")
    ap.add_argument("--no-vessel-tree", action="store_true", help="skip the sparse vessel-tree measurement")
    ap.add_argument("--vessel-tree-steps", type=int, default=1000)
    ap.add_argument("--serving-requests", type=int, default=24)
    ap.add_argument("--serving-rate", type=float, default=8.0, help="Poisson arrivals per second")
    ap.add_argument("--serving-max-sites", type=float, default=5e6)
    ap.add_argument("--calibrate", action="store_true", help="run the 2 GiB calibration copy and exit")
    ap.add_argument("--variant", type=int, default=None,
                    help="fixed launch shape instead of autotuning (used by the counter passes)")
    ap.add_argument("--measure-traffic", action="store_true",
                    help="measure HBM bytes per launch with rocprofv3 counters before the timed run")
    args = ap.parse_args()
    if args.calibrate:
        return calibrate_copy()
    traffic = None
    if args.measure_traffic:
        traffic = measure_traffic(args.workload, 1 if args.variant is None else args.variant)
        log(f"traffic: {json.dumps(traffic)}")
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    import hemepure_amd as H

    total = args.steps + args.warmup + 8 + 36 * 2 * 10 + 64  # + the launch-variant trials
    case, desc = make_case(args.workload, total)
    log(f"workload {args.workload}: {case.geom.n_sites} sites")
    lat = case.product()
    lat.init_equilibrium(1.0)
    lat.step(max(1, args.warmup))
    if args.variant is None:
        variant64, _ = lat.autotune(10)  # launch shape; the trial steps are real steps
    else:
        variant64 = args.variant
        lat.set_variant(variant64)
    lat.synchronize()
    log(f"warmup {args.warmup} steps done, launch shape {variant64}")

    # Timed region: HIP events recorded on the stream the kernels run on.
    wall0 = time.perf_counter()
    ms_total = time_steps(lat, args.steps)
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
    lat.close()

    n32, ms32, stable32, variant32 = time_form(case, H.FP32_SHIFTED, args.steps, args.warmup, args.variant)
    mflups32 = n32 * args.steps / (ms32 / 1e3) / 1e6
    achieved32 = n32 * BYTES_PER_SITE["f32"] / (ms32 / 1e3 / args.steps) / 1e9
    log(f"fp32-shifted: {args.steps} steps in {ms32:.2f} ms: {mflups32:.1f} MFLUPS"
        + ("" if stable32 else " (unstable)"))

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
        "config": {"workload": desc, "sites": n, "storage": "fp64 SoA, compact 8x8 slabs of 8^3 bricks, arithmetic neighbour addressing",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic["bytes_per_launch"]) if traffic else _committed_traffic(args.workload, n),
                     "bytes_per_site_update": BYTES_PER_SITE["f64"],
                     "layout_bytes_per_site_update": LAYOUT_BYTES_PER_SITE["f64"], "kernel": "step_fp64"},
    }
    out["forms"] = {
        "fp64": {"mflups": round(mflups, 2), "ms_per_step": ms_total / args.steps, "launch_variant": variant64,
                 "bytes_per_site_update": BYTES_PER_SITE["f64"], "hbm_frac": round(achieved / HBM_PEAK_GBS, 4)},
        "fp32_shifted": {"mflups": round(mflups32, 2), "ms_per_step": ms32 / args.steps,
                         "bytes_per_site_update": BYTES_PER_SITE["f32"],
                         "hbm_frac": round(achieved32 / HBM_PEAK_GBS, 4),
                         "vs_fp64": round(mflups32 / mflups, 3), "kernel": "step_fp32",
                         "launch_variant": variant32},
    }
    if not args.no_ensemble:
        out["ensemble"] = bench_ensemble(args.ensemble_members, args.ensemble_steps, 5000, args.workers)
    if not args.no_vessel_tree:
        out["vessel_tree"] = bench_vessel_tree(args.vessel_tree_steps, args.warmup, args.workers)
    if not args.no_serving:
        out["serving"] = bench_serving(args.serving_requests, args.serving_rate, args.serving_max_sites,
                                       args.workers)
    if not args.no_cpu_baseline:
        log("timing the CPU oracle on a bounded sample")
        out["cpu_baseline"] = cpu_baseline(case)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
