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
  drift       total vorticity, linear impulse and kinetic energy before the
              warm-up steps and after the timed steps, per path
  boundary    the binary64 rate across the host boundary: state uploaded, one
              RHS, rates downloaded (as pawan's setStates / solve / getRates)
  cpu_baseline  the OpenMP CPU oracle (pawan's __parallel loop restated, with
              its per-pair heap temporaries) on a stratified sample of source
              rows of the same workload, on this host's cores; and the
              oracle's rates of a stratified sample of targets (against all
              262,144 sources), which each path's rates are checked against:
              fp32c.rel_rms_err_vs_fp64_oracle (per term) and
              max_rel_err_vs_fp64_oracle (binary64, per particle)
  serving     the serving stream (N log-uniform in [2,048, 262,144], 50-500
              steps, binary64 / binary32 50/50, 200 GB budget) for a window of
              Poisson arrivals: p50/p99 step and job latency, unfinished jobs
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
FLOP_PER_PAIR_SYM = 55.0      # the symmetric form: ~110 per unordered pair, shared by both
FP64_VECTOR_PEAK = 78.6e12    # MI355X FP64 vector FLOP/s
FP32_VECTOR_PEAK = 157.3e12   # MI355X FP32 vector FLOP/s (published, packed)
# HBM bytes per binary64 pair-kernel launch at N = 262,144 with 256-thread
# blocks, by (targets per thread, source chunks) -- every chunk count the
# autotuner tries (the policy's 16 and 1/2, 3/4, 3/2, 2x of it): rocprofv3
# FETCH_SIZE and WRITE_SIZE in KiB per dispatch (mean over 7 dispatches), FETCH
# scaled by what the same counter reads for a cold 2 GiB copy of known size
# (x2), WRITE as read (x1; the kernel writes exactly its partial sums,
# chunks x 10 x npad x 8 bytes); profiles/pair_kernels_rotor_wake_262144.md
PAIR64_KIB_262144 = {  # (targets per thread, chunks): (FETCH_SIZE, WRITE_SIZE)
    (1, 8): (166_713, 164_096), (1, 12): (177_171, 246_144), (1, 16): (204_471, 328_192),
    (1, 24): (265_307, 492_288), (1, 32): (329_679, 656_384),
    (2, 8): (160_770, 163_968), (2, 12): (184_528, 245_952), (2, 16): (208_962, 327_936),
    (2, 24): (265_057, 491_904), (2, 32): (322_809, 655_872),
}


def pair64_traffic(n, block, targets_per_thread, chunks):
    """HBM bytes of one binary64 pair launch at N = 262,144 for a profiled
    variant (None otherwise)."""
    key = (targets_per_thread, chunks)
    if n != 262144 or block != 256 or key not in PAIR64_KIB_262144:
        return None
    fetch, write = PAIR64_KIB_262144[key]
    return (2.0 * fetch + write) * 1024.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# The synthetic rotor wake (tip vortex of strength 1 in cores of 0.02) is an
# unstable particle set: vorticity grows with an e-folding time of about 1 ms
# of simulated time whatever the step (the CPU oracle shows the same: NaN by
# t = 0.08 at dt 1e-3, 1e17 by t = 0.03 at dt 2.5e-4).  Its runs use
# STEP_DT = 1e-5, so that over a timed run of tens of steps (and the growth
# and serving runs of hundreds) the simulated time stays well under 1 ms, the
# state stays finite and the drift of the invariants means something.  A
# step costs the same for any dt.
STEP_DT = 1e-5
LONG_DT = 1e-5

# The variants the autotuner picks for the N = 262,144 rotor wake on MI355X
# (used with --no-tune, e.g. under a profiler): binary64 one-sided 256 x 1
# target, unroll 4, 24 source chunks (103.6 ms per RHS; 103.7 at 32 chunks,
# 105.7 for 256 x 2, unroll 2, at the policy's 16); binary64 symmetric 256 x 2,
# 256 chunks and strips (73.5 against 75.3 at the policy's 128); compensated
# binary32 symmetric 256 x 1, 256 chunks (46.7 against 48.7 at 128, 49.8 for
# 512 x 1).
MEASURED_BEST = {
    "fp64": {"block": 256, "targets_per_thread": 1, "unroll": 4, "scheme": 1, "chunks": 24},
    "fp64_symmetric": {"block": 256, "targets_per_thread": 2, "unroll": 0, "scheme": 2,
                       "chunks": 256},
    "fp32c": {"block": 256, "targets_per_thread": 1, "unroll": 0, "scheme": 2, "chunks": 256},
}


def run_path(pa, w, precision, steps, warmup, variant=None):
    v = variant or {}
    S = pa.Interaction(pa.Wake.from_dict(w), nu=w["nu"], precision=precision,
                       uinf=w.get("uinf", (0, 0, 0)), block=v.get("block", 0),
                       targets_per_thread=v.get("targets_per_thread", 0),
                       unroll=v.get("unroll", 0), scheme=v.get("scheme", 0),
                       chunks=v.get("chunks", 0))
    dt = STEP_DT
    S.solve()  # builds the device job
    d0 = S.diagnose()  # before the warm-up steps (outside the timed region)
    for _ in range(warmup):
        S.rk4(dt, 1)
    t0 = time.perf_counter()
    S.rk4(dt, steps)
    wall = time.perf_counter() - t0
    tm = S.last_timing()
    d1 = S.diagnose()
    n = w["n"]
    pairs = 4.0 * steps * n * (n - 1)
    launches = 4 * steps
    out = {
        "wall_s": wall, "pairs": pairs, "device_ms": tm["total_ms"],
        "pair_kernel_ms_mean": tm["pair_ms"] / launches,
        "pairs_per_launch": n * (n - 1.0),
        "pairs_per_s": pairs / wall,
        "drift": dict(_drift(d0, d1, float(np.linalg.norm(w["vort"], axis=1).sum())),
                      steps=warmup + steps, dt=dt),
    }
    return S, out


def boundary_rate(S, w, reps=3):
    """One RHS with the state handed over from host memory and the rates
    handed back, as pawan's setStates / solve / getRates do (interaction.cpp:
    204-229): the pair rate including both PCIe copies (never the bench's
    value, which keeps the state resident)."""
    n = w["n"]
    x = np.concatenate([w["pos"].ravel(), w["vort"].ravel()])
    r = np.empty(6 * n)
    S.setStates(x)
    S.solve()
    S.getRates(r)
    wall = dev = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        S.setStates(x)
        S.solve()
        S.getRates(r)
        wall += time.perf_counter() - t0
        dev += S.last_timing()["total_ms"]
    ms = wall * 1e3 / reps
    return {"ms_per_rhs": ms, "device_ms_per_rhs": dev / reps,
            "pairs_per_s": n * (n - 1.0) / (ms * 1e-3), "state_bytes_each_way": 48 * n,
            "note": "host state in, one RHS, host rates out (PCIe both ways); not the value"}


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


def split_error_for(pa, w, nu, uinf=(0.0, 0.0, 0.0)):
    """rates_split_error for the particles of w (one wake) at their state."""
    ww = {k: w[k] for k in ("pos", "vort", "radius", "volume")}
    ww["n"], ww["nu"] = int(w["pos"].shape[0]), nu
    S64 = pa.Interaction(pa.Wake.from_dict(ww), nu=nu, uinf=uinf)
    S32 = pa.Interaction(pa.Wake.from_dict(ww), nu=nu, uinf=uinf, precision=pa.FP32C)
    e = rates_split_error(S64, S32, ww)
    S64.close()
    S32.close()
    return e


def gpu_terms(S, w):
    """The path's rates at the workload's initial state, per term: velocity,
    vorticity rate, stretching (nu = 0) and diffusion (their difference)."""
    n = w["n"]
    S.setStates(np.concatenate([w["pos"].ravel(), w["vort"].ravel()]))
    S.solve()
    r = S.getRates()
    S.setNu(0.0)
    S.solve()
    r0 = S.getRates()
    S.setNu(w["nu"])
    v, d = r[:3 * n].reshape(n, 3), r[3 * n:].reshape(n, 3)
    st = r0[3 * n:].reshape(n, 3)
    return v, d, st, d - st


def sample_targets(n, k, seed=0):
    """About k target indices: one per equal stratum (seeded), plus both ends
    and the indices either side of every 16,384th (where source chunks split)."""
    rng = np.random.default_rng(seed)
    b = np.linspace(0, n, k + 1).astype(np.int64)
    idx = [int(rng.integers(lo, hi)) for lo, hi in zip(b[:-1], b[1:]) if hi > lo]
    idx += [0, n - 1] + [m + e for m in range(16384, n, 16384) for e in (-1, 0)]
    return np.unique(np.clip(idx, 0, n - 1))


def oracle_check(orc, w, gpu, targets):
    """Rates of the sampled targets from the oracle (the reference's binary64
    pair arithmetic, every source) against each GPU path's: binary64 max
    per-particle relative error (floor 1e-3 RMS), binary32 relative RMS per term."""
    t0 = time.perf_counter()
    rv, rs, rf = orc.rates_sample(w["pos"], w["vort"], w["radius"], w["volume"], w["nu"],
                                  targets, uinf=np.asarray(w.get("uinf", (0.0, 0.0, 0.0))))
    secs = time.perf_counter() - t0
    rd = rs + rf

    def per_particle(a, b):
        nb = np.linalg.norm(b, axis=1)
        fl = 1e-3 * np.sqrt(np.mean(nb * nb))
        return float(np.max(np.linalg.norm(a - b, axis=1) / np.maximum(nb, fl)))

    def rr(a, b):
        return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))

    out = {"targets": int(len(targets)), "oracle_seconds": secs}
    for name, (v, d, st, df) in gpu.items():
        v, d, st, df = v[targets], d[targets], st[targets], df[targets]
        if name.startswith("fp64"):
            out[name] = {"max_rel_err_velocity": per_particle(v, rv),
                         "max_rel_err_vorticity_rate": per_particle(d, rd)}
        else:
            out[name] = {"velocity": rr(v, rv), "stretching": rr(st, rs), "diffusion": rr(df, rf),
                         "vorticity_rate": rr(d, rd)}
    return out


def cpu_baseline(w, budget_s=15.0, gpu=None, n_sample=512):
    """The CPU leg: the reference-faithful OpenMP loop timed on a sample of
    source rows, and (gpu given) the oracle's rates of a sample of targets,
    which every GPU path's rates at the same state are checked against."""
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
    out = {
        "value": ordered / secs, "unit": "pair-interactions/s", "cores": orc.max_threads(),
        "kind": "port",
        "sample": (f"{nrows} of {n} source rows (stride {stride}) of one RHS of the same "
                   f"rotor wake, OpenMP loop of parallel.cpp:13-48 with per-pair heap "
                   f"temporaries; {pairs:.3e} unordered pairs in {secs:.1f} s"),
    }
    if gpu:
        out["oracle_check"] = oracle_check(orc, w, gpu, sample_targets(n, n_sample))
    return out


def serving_run(pa, gen, duration_s=120.0, rate_hz=0.1, seed=1, budget=200 * 10 ** 9, n_lo=2048,
                n_hi=262144, s_lo=50, s_hi=500, drain_factor=1.5, check_failed=False):
    """The serving stream (generators.serving_stream): Poisson arrivals at
    rate_hz for duration_s; ring or rotor-wake jobs with N log-uniform in
    [n_lo, n_hi], steps uniform in [s_lo, s_hi], precision 50/50, admitted
    against an HBM budget of `budget` bytes.  Runs until every arrived job
    has finished or drain_factor x duration_s has passed; jobs still running
    or queued then are reported as unfinished (their latency is not in the
    job percentiles).
    check_failed: rerun every failed binary32 job alone in binary64 for the
    steps it completed and report the largest |state| it reaches (a job
    that diverges there too diverges under the method itself)."""
    count = int(rate_hz * duration_s * 3) + 16
    stream = [j for j in gen.serving_stream(count, seed=seed, rate_hz=rate_hz, n_lo=n_lo,
                                            n_hi=n_hi, s_lo=s_lo, s_hi=s_hi, rotor_dt=LONG_DT)
              if j["arrival_ms"] < duration_s * 1000.0]
    srv = pa.Server(budget)
    k, ids, meta = 0, [], {}
    t_end = duration_s * 1000.0
    t_log = 0.0
    in_step = 0.0  # wall seconds spent inside pw_server_step
    w0 = time.perf_counter()
    while True:
        if srv.now_ms() >= t_log:  # progress line every 20 s
            st = srv.stats()
            log(f"  serving t={srv.now_ms() / 1000.0:.0f} s: submitted {st['submitted']} "
                f"running {st['running']} queued {st['queued']} finished {st['finished']}")
            t_log += 20000.0
        now = srv.now_ms()
        while k < len(stream) and now >= stream[k]["arrival_ms"]:
            j = stream[k]
            ids.append(srv.submit(j["w"], j["nu"], j["dt"], j["steps"], j["precision"],
                                  uinf=j["uinf"], arrival_ms=j["arrival_ms"]))
            meta[ids[-1]] = (j["kind"], j["n"], "fp64" if j["precision"] == 0 else "fp32c",
                             j["steps"], k)
            k += 1
        t0 = time.perf_counter()
        running = srv.step()
        in_step += time.perf_counter() - t0
        st = srv.stats()
        if k == len(stream) and running == 0 and st["queued"] == 0:
            break
        if srv.now_ms() > drain_factor * t_end:
            break
        if running == 0 and st["queued"] == 0 and k < len(stream):
            time.sleep(max(0.0, min(stream[k]["arrival_ms"] - srv.now_ms(), 50.0)) / 1000.0)
    wall = time.perf_counter() - w0
    st = srv.stats()
    failed, unfinished = [], []
    for i in ids:
        state, done = srv.state(i)
        if state in (srv.QUEUED, srv.RUNNING):
            unfinished.append(list(meta[i][:4]) + [done])
        if state == srv.FAILED:
            failed.append(list(meta[i][:4]) + [done])
            if check_failed and meta[i][2] == "fp32c":
                j = stream[meta[i][4]]
                S = pa.Interaction(pa.Wake.from_dict(j["w"]), nu=j["nu"], uinf=j["uinf"])
                try:
                    S.rk4(j["dt"], done)
                    fin = True
                except pa.PawanError as e:
                    if e.code != 4:
                        raise
                    fin = False
                x = S.getStates()
                S.close()
                ok = np.isfinite(x)
                failed[-1].append({"fp64_finite": fin, "fp64_max_abs_state": float(np.abs(x[ok]).max()),
                                   "max_abs_state_t0": float(np.abs(j["w"]["pos"]).max())})
                log(f"  failed {meta[i][:4]} at step {done}: alone in binary64 {failed[-1][-1]}")
    srv.close()
    pairs_per_s = st["pairs"] / max(st["busy_ms"], 1e-9) * 1000.0
    return {
        "failed_jobs": failed,  # [kind, n, precision, steps asked, steps done(, binary64 check)]
        "unfinished_jobs": unfinished,  # [kind, n, precision, steps asked, steps done]
        "arrival_rate_hz": rate_hz, "duration_s": duration_s, "stop_after_s": drain_factor * duration_s,
        "jobs": st["submitted"], "finished": st["finished"], "failed": st["failed"],
        "step_p50_ms": st["step_p50_ms"], "step_p99_ms": st["step_p99_ms"],
        "job_p50_ms": st["job_p50_ms"], "job_p99_ms": st["job_p99_ms"],
        # pair interactions over the device time of the passes, and over the
        # wall time of the whole run (idle gaps between arrivals included)
        "pairs_per_s": pairs_per_s, "pairs_per_s_wall": st["pairs"] / wall,
        # device time of the passes over the wall time spent in pw_server_step:
        # the rest is the host's share of a pass (planning, flags, retirement)
        "device_busy_in_passes": st["busy_ms"] / 1000.0 / max(in_step, 1e-9),
        # the served mix at the binary64 method's 95 algorithmic FLOP per ordered
        # pair against the FP64 vector peak (binary32 jobs count at that rate too)
        "frac_fp64_peak_95flop": pairs_per_s * FLOP_PER_PAIR / FP64_VECTOR_PEAK,
        "steps": st["steps"], "n_range": [n_lo, n_hi], "steps_range": [s_lo, s_hi],
        "hbm_budget_gb": budget / 1e9, "still_running": st["running"] + st["queued"],
    }


def copy_calibration(mib):
    """Two device-to-device copies of mib MiB between fresh buffers larger
    than the caches: under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, the known
    byte count of this kernel gives the scale of the counters (torch owns the
    memory and the copy kernel)."""
    import torch
    n = mib << 20
    out = []
    for _ in range(2):
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty(n, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.copy_(a)
        torch.cuda.synchronize()
        out.append(time.perf_counter() - t0)
        del a, b
    return {"copy_mib": mib, "bytes_read": n, "bytes_written": n, "seconds": out}


# ------------------------------------------------------------------ workloads
def _drift(d0, d1, strength=None):
    """Change of total vorticity and linear impulse (interaction.cpp:240-259):
    |change|, and |change| relative to |initial| (total vorticity: relative to
    the summed strength magnitudes `strength` when given, since a wake's total
    vorticity can start at zero); kinetic energy (:351-366): relative change."""
    out = {}
    for k in ("total_vorticity", "linear_impulse"):
        a, b = np.asarray(d0[k]), np.asarray(d1[k])
        out[k] = float(np.linalg.norm(b - a))
        ref = strength if (k == "total_vorticity" and strength) else np.linalg.norm(a)
        out[k + "_rel"] = float(np.linalg.norm(b - a) / max(ref, 1e-300))
    out["kinetic_energy_rel"] = float((d1["kinetic_energy"] - d0["kinetic_energy"]) /
                                      d0["kinetic_energy"])
    return out


def bench_ring1(pa, gen, args):
    """Single ring, 2,048 particles, binary64, RK4 dt 0.01 x 100 steps, and the
    same run on the CPU oracle (reference-faithful OpenMP loop) for parity and
    the CPU time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc  # CPU baseline / checker
    w = gen.config_single_ring()
    S = pa.Interaction(pa.Wake.from_dict(w), nu=w["nu"])
    d0 = S.diagnose()
    x0 = S.getStates()
    t0 = time.perf_counter()
    S.rk4(w["dt"], w["steps"])
    gpu_s = time.perf_counter() - t0
    d1 = S.diagnose()
    xg = S.getStates()
    S.close()
    t0 = time.perf_counter()
    xc = orc.rk4(x0, w["radius"], w["volume"], w["nu"], w["dt"], w["steps"], mode=orc.OMP_FAITHFUL)
    cpu_s = time.perf_counter() - t0
    n = w["n"]
    pairs = 4.0 * w["steps"] * n * (n - 1)
    err = float(np.max(np.abs(xg - xc)) / np.max(np.abs(xc)))
    return {"metric": "single ring 2,048 particles, RK4 dt 0.01 x 100: pair-interactions/s",
            "value": pairs / gpu_s, "unit": "pair-interactions/s", "n_gpus": 1,
            "steps": w["steps"], "ms_per_step": gpu_s * 1e3 / w["steps"], "dtype": "f64",
            "config": {"workload": "single_ring_2048", "dt": w["dt"], "nu": w["nu"]},
            "trajectory_max_rel_diff_vs_cpu": err, "drift": _drift(d0, d1),
            "fp32c_rel_rms_err_vs_fp64_t0": split_error_for(pa, w, w["nu"]),
            "cpu_baseline": {"value": pairs / cpu_s, "seconds": cpu_s, "cores": orc.max_threads(),
                             "kind": "port", "sample": "the full 100-step run"}}


def bench_frames(pa, gen, args):
    """A .wake frame after every RK4 step, as integrate() writes them
    (integration.cpp:16-35): per-step time with no frames, with synchronous
    frames (copy to the host, then fwrite, before the next step) and with
    asynchronous frames (snapshot on the device; copy and file write overlap
    the next step).  Files go to a temporary directory and are checked equal."""
    import tempfile
    cases = []
    ring = gen.config_single_ring()
    cases.append(("single_ring_2048", [ring], pa.FP64, ring["dt"], 100))
    lf = gen.config_leapfrog()
    n1 = lf["wakes"][0]
    halves = [{k: lf[k][:n1] for k in ("pos", "vort", "radius", "volume")},
              {k: lf[k][n1:] for k in ("pos", "vort", "radius", "volume")}]
    cases.append(("leapfrog_65536", halves, pa.FP64, lf["dt"], 20))
    rot = gen.config_rotor_wake(args.n_shed)
    cases.append((f"rotor_wake_{rot['n']}_fp32c", [rot], pa.FP32C, STEP_DT, 6))
    out = {"metric": "RK4 step time with a .wake frame per step: none / synchronous / asynchronous",
           "unit": "ms/step", "n_gpus": 1, "cases": {}}
    with tempfile.TemporaryDirectory() as tmp:
        for name, wakes, prec, dt, steps in cases:
            res = {}
            files = {}
            for mode in ("none", "sync", "async"):
                S = pa.Interaction(*[pa.Wake.from_dict(w) for w in wakes], nu=1e-4, precision=prec)
                S.rk4(dt, 1)  # warm: plan, graph, first launches
                path = os.path.join(tmp, f"{name}_{mode}.wake")
                f = pa.WakeFile(path)
                t0 = time.perf_counter()
                if mode == "sync":
                    S.write(f, 0.0)
                elif mode == "async":
                    S.write_async(f, 0.0)
                for i in range(1, steps + 1):
                    S.rk4(dt, 1)
                    if mode == "sync":
                        S.write(f, i * dt)
                    elif mode == "async":
                        S.write_async(f, i * dt)
                f.close()  # waits for queued frames
                sec = time.perf_counter() - t0
                S.close()
                res[mode] = sec * 1e3 / steps
                if mode != "none":
                    with open(path, "rb") as fh:
                        files[mode] = fh.read()
                os.remove(path)
            n = sum(len(w["pos"]) for w in wakes)
            res["frame_bytes"] = len(files["sync"]) // (steps + 1)
            res["files_identical"] = files["sync"] == files["async"]
            res["sync_overhead"] = res["sync"] / res["none"] - 1.0
            res["async_overhead"] = res["async"] / res["none"] - 1.0
            res["particles"] = n
            log(f"  frames {name}: none {res['none']:.3f} sync {res['sync']:.3f} "
                f"async {res['async']:.3f} ms/step, identical {res['files_identical']}")
            out["cases"][name] = res
    return out


def bench_leapfrog(pa, gen, args):
    """Two coaxial rings, 2 x 32,768 particles, as two wakes; binary64 and
    compensated binary32, RK4 dt 0.005 x steps (400 for the full config)."""
    w = gen.config_leapfrog()
    steps = args.steps
    out = {"metric": "leapfrogging rings 65,536 particles: pair-interactions/s",
           "unit": "pair-interactions/s", "n_gpus": 1, "steps": steps,
           "config": {"workload": "leapfrog_65536", "wakes": w["wakes"], "dt": w["dt"],
                      "nu": w["nu"]}}
    n, n1 = w["n"], w["wakes"][0]
    x0 = np.concatenate([w["pos"][:n1].ravel(), w["vort"][:n1].ravel(),
                         w["pos"][n1:].ravel(), w["vort"][n1:].ravel()])
    rates, final = {}, {}
    for name, prec in (("fp64", pa.FP64), ("fp32c", pa.FP32C)):
        a = {k: w[k][:n1] for k in ("pos", "vort", "radius", "volume")}
        b = {k: w[k][n1:] for k in ("pos", "vort", "radius", "volume")}
        S = pa.Interaction(pa.Wake.from_dict(a), pa.Wake.from_dict(b), nu=w["nu"], precision=prec)
        S.setStates(x0)
        S.solve()
        rates[name] = S.getRates()
        d0 = S.diagnose()
        t0 = time.perf_counter()
        nonfinite = False
        try:
            S.rk4(w["dt"], steps)  # every step runs; a non-finite state is reported at the end
        except pa.PawanError as e:
            if e.code != 4:
                raise
            nonfinite = True
        sec = time.perf_counter() - t0
        log(f"  leapfrog {name}: {steps} steps in {sec:.1f} s")
        final[name] = S.getStates()
        res = {"value": 4.0 * steps * n * (n - 1) / sec, "ms_per_step": sec * 1e3 / steps,
               "non_finite": nonfinite}
        fin = np.isfinite(final[name])
        res["max_abs_finite_state_end"] = float(np.abs(final[name][fin]).max()) if fin.any() else None
        if not nonfinite:
            res["drift"] = _drift(d0, S.diagnose())
        S.close()
        out[name] = res
    out["value"] = out["fp64"]["value"]
    out["dtype"] = "f64"
    r64, r32 = rates["fp64"], rates["fp32c"]
    rr = lambda a, b: float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))  # noqa: E731
    v = np.r_[r64[:3 * n1], r64[6 * n1:6 * n1 + 3 * (n - n1)]]
    v32 = np.r_[r32[:3 * n1], r32[6 * n1:6 * n1 + 3 * (n - n1)]]
    out["fp32c"]["rel_rms_rate_err_vs_fp64_t0"] = {"all": rr(r32, r64), "velocity": rr(v32, v)}
    out["fp32c"]["rel_rms_err_vs_fp64_t0_per_term"] = split_error_for(pa, w, w["nu"])
    out["note"] = ("the reference HIGHORDER kernel has q ~ 1/rho^3 at small separation (the phi term "
                   "keeps t^-1/2 while Z decays as (t+1)^-7.5); over long runs the two rings' "
                   "particles meet and the state diverges in binary64 as in binary32")
    # state after the run: binary32 pairs vs binary64, positions and strengths
    f64, f32 = final["fp64"], final["fp32c"]
    pos = np.r_[np.arange(3 * n1), 6 * n1 + np.arange(3 * (n - n1))]
    vor = np.r_[3 * n1 + np.arange(3 * n1), 6 * n1 + 3 * (n - n1) + np.arange(3 * (n - n1))]
    if np.isfinite(f64).all() and np.isfinite(f32).all():
        out["fp32c"]["rel_rms_state_diff_vs_fp64_end"] = {"position": rr(f32[pos], f64[pos]),
                                                         "vorticity": rr(f32[vor], f64[vor])}
    return out


def bench_rotor_growth(pa, gen, args):
    """Rotor wake grown by shedding: each step emits 256 particles, then one
    RK4 step (dt LONG_DT); compensated binary32 pairs, binary64 state.  Rows
    are placed kinematically as in the frozen wake: in the wake's frame the
    rotor climbs by inflow x shed interval (0.05 x 0.1) per shed step, so row k
    is emitted at z = 0.005 k and the grown wake has the frozen wake's
    geometry (rows emitted at one height would pile into each other's cores)."""
    shed = args.shed_steps
    climb = 0.05 * 0.1

    def row(k):
        e = gen.rotor_emission(k)
        e["pos"][:, 2] += climb * k
        return e

    e0 = row(0)
    S = pa.Interaction(pa.Wake.from_dict(e0), nu=1e-4, precision=pa.FP32C, uinf=(0, 0, -0.05))
    pairs, t0 = 0.0, time.perf_counter()
    n = e0["n"]
    for k in range(shed):
        if k > 0:
            e = row(k)
            S.appendParticles(0, e["pos"], e["vort"], e["radius"], e["volume"])
            n += e["n"]
        S.rk4(LONG_DT, 1)
        pairs += 4.0 * n * (n - 1)
        if k % 128 == 0:
            log(f"  rotor growth step {k} N={n} t={time.perf_counter() - t0:.1f} s")
    sec = time.perf_counter() - t0
    d = S.diagnose(ke=False)
    # per-term error of the binary32 pairs at the grown state
    st = S.getStates()
    W = S._W[0]
    fin = {"pos": st[:3 * n].reshape(n, 3), "vort": st[3 * n:].reshape(n, 3),
           "radius": W._radius, "volume": W._volume}
    S.close()
    err = split_error_for(pa, fin, 1e-4, uinf=(0.0, 0.0, -0.05))
    # both paths' rates at the grown state against the oracle's on 512
    # sampled targets (each against all sources)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc  # the checker
    ww = dict(fin, n=n, nu=1e-4, uinf=(0.0, 0.0, -0.05))
    gpu = {}
    for name, prec in (("fp64", pa.FP64), ("fp32c", pa.FP32C)):
        Sx = pa.Interaction(pa.Wake.from_dict(ww), nu=1e-4, uinf=ww["uinf"], precision=prec)
        gpu[name] = gpu_terms(Sx, ww)
        Sx.close()
    chk = oracle_check(orc, ww, gpu, sample_targets(n, 512))
    return {"metric": "rotor wake grown to N particles, FP32c pairs, FP64 state: pair-interactions/s",
            "value": pairs / sec, "unit": "pair-interactions/s", "n_gpus": 1, "steps": shed,
            "dtype": "f32c", "config": {"workload": f"rotor_growth_{n}", "final_n": n,
                                        "dt": LONG_DT}, "seconds": sec,
            "total_vorticity": d["total_vorticity"].tolist(),
            "fp32c_rel_rms_err_vs_fp64_end": err,
            "oracle_check_end": chk}


def bench_batch(pa, gen, args):
    """256 independent rings (N log-uniform in [1,024, 65,536]) stepped
    together: one persistent launch per stage over the job table, HIP-graph
    replay of the step, variant picked by timing candidates on the batch.

    Over the full 200 steps at dt 0.005 about 100 of the 256 rings diverge
    under the reference method itself (the CPU oracle takes job 196 from
    |x| = 0.57 at step 100 to 5e19 at step 200): binary64 stays finite and
    huge, binary32 overflows and is flagged.  Both counts are reported, with
    the overlap."""
    jobs = gen.batch_jobs(args.jobs, seed=0)
    pairs_rhs = sum(float(j["n"]) * (j["n"] - 1) for j in jobs)
    out = {"metric": f"{len(jobs)} batched ring simulations: pair-interactions/s",
           "unit": "pair-interactions/s", "n_gpus": 1, "steps": args.steps,
           "config": {"workload": f"batch_{len(jobs)}_rings", "particles": sum(j["n"] for j in jobs),
                      "pairs_per_rhs": pairs_rhs, "dt": 0.005}}
    sets = {}
    for name, prec in (("fp64", pa.FP64), ("fp32c", pa.FP32C)):
        best = None
        for block in (64, 128, 256):
            for tpt in (1, 2):
                B = pa.Batch(block=block, targets_per_thread=tpt)
                for j in jobs:
                    B.add(j, j["nu"], j["dt"], prec)
                B.step(1, graph=True)
                t0 = time.perf_counter()
                B.step(1, graph=True)
                sec = time.perf_counter() - t0
                B.close()
                log(f"  batch {name} block {block} tpt {tpt}: {sec * 1e3:.1f} ms/step")
                if best is None or sec < best[0]:
                    best = (sec, block, tpt)
        B = pa.Batch(block=best[1], targets_per_thread=best[2])
        ids = [B.add(j, j["nu"], j["dt"], prec) for j in jobs]
        B.step(args.warmup, graph=True)
        t0 = time.perf_counter()
        B.step(args.steps, graph=True)
        sec = time.perf_counter() - t0
        bad = [k for k, i in enumerate(ids) if B.status(i) != 0]
        div = []
        for k, i in enumerate(ids):
            p = B.get(i)["pos"]
            if not np.all(np.isfinite(p)) or np.abs(p).max() > 100.0:
                div.append(k)
        B.close()
        out[name] = {"value": 4.0 * args.steps * pairs_rhs / sec, "ms_per_step": sec * 1e3 / args.steps,
                     "variant": {"block": best[1], "targets_per_thread": best[2]},
                     "failed_jobs": len(bad), "diverged_jobs": len(div)}
        sets[name] = (set(bad), set(div))
    out["value"] = out["fp64"]["value"]
    out["dtype"] = "f64"
    out["fp32c"]["speedup_vs_fp64"] = out["fp32c"]["value"] / out["fp64"]["value"]
    # per-term binary32 errors at t = 0 on the 4 largest and 4 other jobs
    order = sorted(range(len(jobs)), key=lambda k: -jobs[k]["n"])
    sample = order[:4] + order[4::max(1, (len(order) - 4) // 4)][:4]
    errs = [split_error_for(pa, jobs[k], jobs[k]["nu"]) for k in sample]
    out["fp32c"]["rel_rms_err_vs_fp64_t0_sample"] = {
        "jobs": sample, "max": {t: max(e[t] for e in errs) for t in errs[0]},
        "median": {t: float(np.median([e[t] for e in errs])) for t in errs[0]}}
    # jobs the binary32 path flags that did not also diverge in binary64
    out["fp32c"]["failed_not_diverged_in_fp64"] = len(sets["fp32c"][0] - sets["fp64"][1])
    return out


def bench_serve(pa, gen, args):
    """Serving stream: Poisson arrivals at --rate per second for
    --serve-seconds; each job a ring or rotor wake with N log-uniform in
    [2,048, 262,144], steps uniform in [50, 500], FP64 or FP32c 50/50, 200 GB
    HBM budget; stops after --drain x --serve-seconds."""
    r = serving_run(pa, gen, duration_s=args.serve_seconds, rate_hz=args.rate,
                    drain_factor=args.drain, check_failed=True)
    return {"metric": "serving: p50/p99 step and job latency, aggregate pair-interactions/s",
            "value": r["pairs_per_s"], "unit": "pair-interactions/s", "n_gpus": 1, "dtype": "mixed",
            "config": {"workload": "serving_stream", "n_range": [2048, 262144],
                       "steps_range": [50, 500], "hbm_budget_gb": 200}, "serving": r}


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
    ap.add_argument("--serve-seconds", type=float, default=150.0,
                    help="serving: arrival window in seconds")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device (twice, fresh buffers): the "
                         "known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--variant64", default="", metavar="B,T,U,C",
                    help="with --no-tune: the binary64 one-sided variant (block, targets per "
                         "thread, unroll, chunks) instead of the measured best")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner; run the variants it picks on MI355X at N=262,144")
    ap.add_argument("--workload", default="rotor262k",
                    choices=["rotor262k", "ring1", "leapfrog", "rotor_growth", "batch", "serve",
                             "frames"])
    ap.add_argument("--shed-steps", type=int, default=1024)
    ap.add_argument("--jobs", type=int, default=256)
    ap.add_argument("--rate", type=float, default=0.2, help="serving arrivals per second")
    ap.add_argument("--drain", type=float, default=1.6, help="serving: stop after drain x duration")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        print(json.dumps(copy_calibration(args.copy_calibration)), flush=True)
        return

    import pawan_amd as pa
    from pawan_amd import generators as gen

    other = {"ring1": bench_ring1, "leapfrog": bench_leapfrog, "rotor_growth": bench_rotor_growth,
             "batch": bench_batch, "serve": bench_serve, "frames": bench_frames}
    if args.workload in other:
        print(json.dumps(other[args.workload](pa, gen, args)), flush=True)
        return

    w = gen.config_rotor_wake(args.n_shed)
    n = w["n"]
    tuned = dict(MEASURED_BEST) if args.no_tune else {}
    if args.no_tune and args.variant64:
        b, t, u, c = (int(x) for x in args.variant64.split(","))
        tuned["fp64"] = {"block": b, "targets_per_thread": t, "unroll": u, "scheme": 1, "chunks": c}
    if not args.no_tune:
        for name, prec in (("fp64", pa.FP64), ("fp32c", pa.FP32C)):
            if name == "fp32c" and args.no_fp32c:
                continue
            best, allv = pa.autotune(w, prec, reps=1)
            if prec == pa.FP64:
                # the plain binary64 method is one-sided; the symmetric pass
                # is reported beside it
                one = [e for e in allv if e["scheme"] == pa.SCHEME_ONESIDED]
                sym = [e for e in allv if e["scheme"] == pa.SCHEME_SYMMETRIC]
                best = min(one, key=lambda e: e["pair_ms"])
                if sym:
                    tuned["fp64_symmetric"] = min(sym, key=lambda e: e["pair_ms"])
            tuned[name] = best
            log(f"autotune {name}: best {best} of {len(allv)} variants")
            for sc in (pa.SCHEME_ONESIDED, pa.SCHEME_SYMMETRIC):
                for e in sorted((e for e in allv if e["scheme"] == sc),
                                key=lambda e: e["pair_ms"])[:5]:
                    log(f"    {e['block']:4d} x{e['targets_per_thread']} unroll {e['unroll']} "
                        f"scheme {e['scheme']} chunks {e['chunks'] or 'policy'}: "
                        f"{e['pair_ms']:.2f} ms")
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
                     "traffic": pair64_traffic(n, **{k: S64.variant()[k] for k in
                                                     ("block", "targets_per_thread", "chunks")}),
                     "kernel": "pair64_kernel", "flop_per_pair": FLOP_PER_PAIR,
                     "launch_ms": r64["pair_kernel_ms_mean"]},
        "variants": tuned, "autotuned": not args.no_tune,
        # total vorticity, linear impulse and kinetic energy before the warm-up
        # steps and after the timed ones (outside the timed region)
        "drift": r64["drift"],
    }
    line["boundary"] = boundary_rate(S64, w)
    log(f"fp64 across the host boundary: {line['boundary']['ms_per_rhs']:.2f} ms per RHS "
        f"({line['boundary']['device_ms_per_rhs']:.2f} on the device)")
    # each path's rates at the initial state, checked against the oracle in
    # the CPU leg
    gpu = {"fp64": gpu_terms(S64, w)}

    def rel_pp(a, b):
        a, b = a.reshape(-1, 3), b.reshape(-1, 3)
        nb = np.linalg.norm(b, axis=1)
        return np.linalg.norm(a - b, axis=1) / np.maximum(nb, 1e-3 * np.sqrt(np.mean(nb * nb)))

    def rr(a, b):
        return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))

    if "fp64_symmetric" in tuned:
        log("fp64 symmetric path")
        Ssym, rsym = run_path(pa, w, pa.FP64, args.steps, args.warmup, tuned["fp64_symmetric"])
        gpu["fp64_symmetric"] = gpu_terms(Ssym, w)
        rel = np.maximum(rel_pp(gpu["fp64_symmetric"][0], gpu["fp64"][0]),
                         rel_pp(gpu["fp64_symmetric"][1], gpu["fp64"][1]))
        line["fp64_symmetric"] = {
            "value": rsym["pairs_per_s"], "ms_per_step": rsym["wall_s"] * 1000.0 / args.steps,
            "pair_kernel_ms": rsym["pair_kernel_ms_mean"],
            "speedup_vs_plain_fp64": rsym["pairs_per_s"] / r64["pairs_per_s"],
            "flop_per_pair": FLOP_PER_PAIR_SYM,
            "frac_fp64_peak": FLOP_PER_PAIR_SYM * rsym["pairs_per_launch"] /
            (rsym["pair_kernel_ms_mean"] * 1e-3) / FP64_VECTOR_PEAK,
            "max_rel_rate_diff_vs_plain": float(rel.max()),
            "drift": rsym["drift"],
        }
        Ssym.close()
    if not args.no_fp32c:
        log("fp32c path")
        S32, r32 = run_path(pa, w, pa.FP32C, args.steps, args.warmup, tuned.get("fp32c"))
        gpu["fp32c"] = g32 = gpu_terms(S32, w)
        g64 = gpu["fp64"]
        err = {"velocity": rr(g32[0], g64[0]), "stretching": rr(g32[2], g64[2]),
               "diffusion": rr(g32[3], g64[3]), "vorticity_rate": rr(g32[1], g64[1])}
        sym = tuned.get("fp32c", {}).get("scheme", pa.SCHEME_DEFAULT) != pa.SCHEME_ONESIDED
        flop32 = FLOP_PER_PAIR_SYM if sym else FLOP_PER_PAIR
        line["fp32c"] = {
            "value": r32["pairs_per_s"], "ms_per_step": r32["wall_s"] * 1000.0 / args.steps,
            "speedup_vs_fp64": r32["pairs_per_s"] / r64["pairs_per_s"],
            "speedup_vs_fp64_symmetric": (r32["pairs_per_s"] / line["fp64_symmetric"]["value"]
                                          if "fp64_symmetric" in line else None),
            "pair_kernel_ms": r32["pair_kernel_ms_mean"],
            # per RHS, pair kernels only (the step adds the reduction and RK4 combine)
            "pair_kernel_speedup_vs_fp64": r64["pair_kernel_ms_mean"] / r32["pair_kernel_ms_mean"],
            "scheme": "symmetric" if sym else "one-sided", "flop_per_pair": flop32,
            "frac_fp32_peak": flop32 * r32["pairs_per_launch"] /
            (r32["pair_kernel_ms_mean"] * 1e-3) / FP32_VECTOR_PEAK,
            "rel_rms_err_vs_fp64_gpu": err,
            "drift": r32["drift"],
        }
        S32.close()
    S64.close()
    if not args.no_serving:
        log("serving run")
        line["serving"] = serving_run(pa, gen, duration_s=args.serve_seconds, rate_hz=args.rate,
                                      drain_factor=args.drain)
    if not args.no_cpu:
        log("cpu baseline")
        line["cpu_baseline"] = cb = cpu_baseline(w, args.cpu_budget, gpu=gpu)
        chk = cb.get("oracle_check", {})
        if "fp32c" in chk:
            line["fp32c"]["rel_rms_err_vs_fp64_oracle"] = chk["fp32c"]
        if "fp64" in chk:
            line["max_rel_err_vs_fp64_oracle"] = chk["fp64"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
