"""Writes the golden fixtures in this directory.

The reference ships no fixtures, so these are produced by the CPU oracle (the
restatement of pawan's method pinned by tests/test_oracle.py) and the seeded
generators.  They freeze the generators' output and the oracle's rates and
one RK4 step for the single-ring workload, so later changes to either show
up as a diff.  Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as orc  # noqa: E402
from pawan_amd import generators as gen  # noqa: E402


def fingerprint(w):
    out = {}
    for k in ("pos", "vort", "radius", "volume"):
        a = np.asarray(w[k], dtype=np.float64).reshape(len(w["radius"]), -1)
        out[k] = {"fsum": math.fsum(a.ravel()), "fsum_abs": math.fsum(np.abs(a).ravel()),
                  "head": a[:8].tolist(), "tail": a[-8:].tolist()}
    out["n"] = int(len(w["radius"]))
    return out


def generators():
    cases = {
        "single_ring": gen.config_single_ring(),
        "leapfrog_small": gen.config_leapfrog(n_azi=32, n_cs=16),
        "rotor_wake_262144": gen.config_rotor_wake(1024),
        "ring_262144": gen.config_ring_large(),
        "rotor_emission_17": gen.rotor_emission(17),
    }
    jobs = gen.batch_jobs(8, seed=0)
    for k, j in enumerate(jobs[:4]):
        cases[f"batch_job_{k}"] = j
    return {name: fingerprint(w) for name, w in cases.items()}


def main():
    with open(os.path.join(HERE, "generators.json"), "w") as f:
        json.dump(generators(), f, indent=1)
    w = gen.config_single_ring()
    vel, dv = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], w["nu"], orc.SERIAL)
    x0 = orc.pack_state(w["pos"], w["vort"])
    x1 = orc.rk4(x0, w["radius"], w["volume"], w["nu"], w["dt"], 1, mode=orc.SERIAL)
    dg = orc.diagnostics(w["pos"], w["vort"], w["radius"])
    np.savez_compressed(os.path.join(HERE, "single_ring.npz"), vel=vel, dvort=dv, state1=x1,
                        omega=dg["total_vorticity"], impulse=dg["linear_impulse"],
                        ke=np.array([dg["kinetic_energy"]]))


if __name__ == "__main__":
    main()
