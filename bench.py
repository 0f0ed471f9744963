#!/usr/bin/env python3
"""Benchmark of the vortex-particle interaction + RK4 hot path on one MI355X.

Workload: the frozen synthetic 4-blade rotor wake at N = 262,144 particles
(core 0.02, nu = 1e-4, freestream 0.05 along -z).  A step is one RK4 step:
four all-pairs RHS evaluations (N(N-1) ordered pair-interactions each) plus
the fused stage combinations, on device-resident state.

Prints one JSON line:
  value       binary64 pair-interactions/s over the K timed steps
  fp32c       the compensated binary32 path on the same workload: rate,
              speedup over binary64, relative RMS rate error per term
  roofline    the binary64 pair kernel's algorithmic FLOP rate (95 FLOP per
              ordered pair) over its mean launch time from HIP events on its
              stream, against the FP64 vector peak
  cpu_baseline  the OpenMP CPU oracle (pawan's __parallel loop restated, with
              its per-pair heap temporaries) on a stratified sample of source
              rows of the same workload, on this host's cores
  serving     a short continuous-batching run: p50/p99 step and job latency
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

FLOP_PER_PAIR = 95.0          # algorithmic FLOP per ordered pair (one-sided form)
FP64_VECTOR_PEAK = 78.6e12    # MI355X FP64 vector FLOP/s
FP32_VECTOR_PEAK = 157.3e12   # MI355X FP32 vector FLOP/s (published, packed)
# HBM bytes per binary64 pair-kernel launch at N = 262,144, from rocprofv3
# FETCH_SIZE + WRITE_SIZE passes (KiB per dispatch, write factor 1.00 from a
# copy kernel of known size); profiles/pair_kernels_rotor_wake_262144.md
PAIR64_TRAFFIC_262144 = (229_000 + 426_528) * 1024.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


STEP_DT = 1e-3  # stable for the tip-vortex swirl (~20 velocity units at 0.016 spacing)


def run_path(pa, w, precision, steps, warmup, variant=None):
    v = variant or {}
    S = pa.Interaction(pa.Wake.from_dict(w), nu=w["nu"], precision=precision,
                       uinf=w.get("uinf", (0, 0, 0)), block=v.get("block", 0),
                       targets_per_thread=v.get("targets_per_thread", 0),
                       unroll=v.get("unroll", 0))
    dt = STEP_DT
    S.solve()  # builds the device job
    for _ in range(warmup):
        S.rk4(dt, 1)
    t0 = time.perf_counter()
    S.rk4(dt, steps)
    wall = time.perf_counter() - t0
    tm = S.last_timing()
    n = w["n"]
    pairs = 4.0 * steps * n * (n - 1)
    launches = 4 * steps
    out = {
        "wall_s": wall, "pairs": pairs, "device_ms": tm["total_ms"],
        "pair_kernel_ms_mean": tm["pair_ms"] / launches,
        "pairs_per_launch": n * (n - 1.0),
        "pairs_per_s": pairs / wall,
    }
    return S, out


def rates_split_error(S64, S32, w):
    """Relative RMS error of the binary32 path's rates vs the binary64 path at
    the initial state, per term: velocity, stretching (nu = 0) and PSE
    diffusion (the difference of the nu and nu = 0 rates)."""
    n = w["n"]
    x0 = np.concatenate([w["pos"].ravel(), w["vort"].ravel()])
    out = {}
    for S in (S64, S32):
        S.setStates(x0)
        S.solve()
        r = S.getRates()
        S.setNu(0.0)
        S.solve()
        r0 = S.getRates()
        S.setNu(w["nu"])
        out[S] = (r[:3 * n], r0[3 * n:], r[3 * n:] - r0[3 * n:], r[3 * n:])

    def rr(a, b):
        return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))

    a, b = out[S32], out[S64]
    return {"velocity": rr(a[0], b[0]), "stretching": rr(a[1], b[1]), "diffusion": rr(a[2], b[2]),
            "vorticity_rate": rr(a[3], b[3])}


def cpu_baseline(w, budget_s=15.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc  # CPU baseline leg only

    n = w["n"]
    # probe, then a stratified sample of source rows sized to the budget
    probe_rows = 8
    stride = n // probe_rows
    pairs, secs = orc.omp_rows(w["pos"], w["vort"], w["radius"], w["volume"], w["nu"],
                               stride // 2, stride, probe_rows, faithful=True)
    rate = pairs / max(secs, 1e-9)
    nrows = int(max(probe_rows, min(n, budget_s * rate / (n / 2.0))))
    stride = max(1, n // nrows)
    nrows = n // stride
    pairs, secs = orc.omp_rows(w["pos"], w["vort"], w["radius"], w["volume"], w["nu"],
                               stride // 2, stride, nrows, faithful=True)
    ordered = 2.0 * pairs  # each unordered pair updates both particles
    return {
        "value": ordered / secs, "unit": "pair-interactions/s", "cores": orc.max_threads(),
        "kind": "port",
        "sample": (f"{nrows} of {n} source rows (stride {stride}) of one RHS of the same "
                   f"rotor wake, OpenMP loop of parallel.cpp:13-48 with per-pair heap "
                   f"temporaries; {pairs:.3e} unordered pairs in {secs:.1f} s"),
    }


def serving_run(pa, gen, duration_s=20.0, rate_hz=4.0, seed=1, budget=200 << 30):
    """Poisson arrivals for duration_s; ring or rotor-wake jobs with N
    log-uniform in [2048, 32768], steps uniform in [50, 200], precision
    50/50; runs until all arrived jobs finish or 3 x duration_s passes."""
    u = gen.uniform(seed, 4096)
    srv = pa.Server(budget)
    t_next, k, ids = 0.0, 0, []
    t_end = duration_s * 1000.0
    while True:
        now = srv.now_ms()
        while now >= t_next and t_next < t_end:
            a = u[(6 * k) % 4090: (6 * k) % 4090 + 6]
            n = int(2048 * (32768 / 2048) ** a[0])
            steps = int(50 + 150 * a[1])
            prec = pa.FP64 if a[2] < 0.5 else pa.FP32C
            if a[3] < 0.5:
                n_cs = gen.ring_cs_for(n)
                w = gen.vortex_ring(max(8, n // n_cs), n_cs, R=1.0, core=0.1, seed=seed + k)
                ids.append(srv.submit(w, 1e-3, 0.005, steps, prec, arrival_ms=t_next))
            else:
                w = gen.rotor_wake(max(1, n // 256))
                ids.append(srv.submit(w, 1e-4, STEP_DT, steps, prec, uinf=(0, 0, -0.05),
                                      arrival_ms=t_next))
            k += 1
            t_next += -math.log(1.0 - a[4]) / rate_hz * 1000.0
        running = srv.step()
        st = srv.stats()
        if t_next >= t_end and running == 0 and st["queued"] == 0:
            break
        if srv.now_ms() > 3 * t_end:
            break
        if running == 0 and st["queued"] == 0:
            time.sleep(max(0.0, min(t_next - srv.now_ms(), 50.0)) / 1000.0)
    st = srv.stats()
    srv.close()
    return {
        "arrival_rate_hz": rate_hz, "duration_s": duration_s, "jobs": st["submitted"],
        "finished": st["finished"], "failed": st["failed"],
        "step_p50_ms": st["step_p50_ms"], "step_p99_ms": st["step_p99_ms"],
        "job_p50_ms": st["job_p50_ms"], "job_p99_ms": st["job_p99_ms"],
        "pairs_per_s": st["pairs"] / max(st["busy_ms"], 1e-9) * 1000.0,
        "steps": st["steps"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-shed", type=int, default=1024, help="rotor shedding steps (N = 256 x)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32c", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner; default variant")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import pawan_amd as pa
    from pawan_amd import generators as gen

    w = gen.config_rotor_wake(args.n_shed)
    n = w["n"]
    tuned = {}
    if not args.no_tune:
        for name, prec in (("fp64", pa.FP64), ("fp32c", pa.FP32C)):
            if name == "fp32c" and args.no_fp32c:
                continue
            best, allv = pa.autotune(w, prec, reps=1)
            tuned[name] = best
            log(f"autotune {name}: best {best} of {len(allv)} variants")
    log(f"workload rotor wake N={n}; FP64 warmup {args.warmup} steps {args.steps}")
    S64, r64 = run_path(pa, w, pa.FP64, args.steps, args.warmup, tuned.get("fp64"))
    log(f"fp64: {r64['pairs_per_s']:.4e} pairs/s, pair kernel {r64['pair_kernel_ms_mean']:.2f} ms")
    achieved = FLOP_PER_PAIR * r64["pairs_per_launch"] / (r64["pair_kernel_ms_mean"] * 1e-3)
    line = {
        "metric": "pair-interactions/s at N=262k, FP64 vs FP32-compensated; p50/p99 job latency",
        "value": r64["pairs_per_s"],
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["wall_s"] * 1000.0 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"rotor_wake_{n}", "particles": n, "blades": 4, "stations": 64,
                   "shed_steps": args.n_shed, "core": 0.02, "nu": 1e-4, "uinf_z": -0.05,
                   "step": "one RK4 step = 4 all-pairs RHS", "dt": STEP_DT, "pairs_per_step": 4.0 * n * (n - 1)},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": FP64_VECTOR_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / FP64_VECTOR_PEAK,
                     "traffic": PAIR64_TRAFFIC_262144 if n == 262144 else None,
                     "kernel": "pair64_kernel", "flop_per_pair": FLOP_PER_PAIR,
                     "launch_ms": r64["pair_kernel_ms_mean"]},
        "tuned_variants": tuned,
    }
    if not args.no_fp32c:
        log("fp32c path")
        S32, r32 = run_path(pa, w, pa.FP32C, args.steps, args.warmup, tuned.get("fp32c"))
        err = rates_split_error(S64, S32, w)
        line["fp32c"] = {
            "value": r32["pairs_per_s"], "ms_per_step": r32["wall_s"] * 1000.0 / args.steps,
            "speedup_vs_fp64": r32["pairs_per_s"] / r64["pairs_per_s"],
            "pair_kernel_ms": r32["pair_kernel_ms_mean"],
            "frac_fp32_peak": FLOP_PER_PAIR * r32["pairs_per_launch"] /
            (r32["pair_kernel_ms_mean"] * 1e-3) / FP32_VECTOR_PEAK,
            "rel_rms_err_vs_fp64_gpu": err,
        }
        S32.close()
    S64.close()
    if not args.no_serving:
        log("serving run")
        line["serving"] = serving_run(pa, gen, duration_s=args.serve_seconds)
    if not args.no_cpu:
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
