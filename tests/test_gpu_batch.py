"""Batched and served simulations: every job's result is the same alone and
in a mix (bitwise: each particle's sums are formed in an order fixed by its
own job), graph replay equals eager launches, and failures stay isolated."""
import numpy as np
import pytest

import oracle as orc
import pawan_amd as pa
from pawan_amd import generators as gen
from helpers import per_particle_rel

pytestmark = pytest.mark.gpu


def _jobs():
    a = gen.vortex_ring(40, 16, R=1.0, core=0.1, seed=1)              # 640
    b = gen.vortex_ring(100, 32, R=0.7, core=0.08, seed=2)            # 3200
    c = gen.rotor_wake(6)                                              # 1536
    d = gen.vortex_ring(17, 9, R=1.3, core=0.12, seed=3, center=(0.1, 0.2, 0.3))  # 153
    return [(a, 1e-3, 0.005, pa.FP64, (0, 0, 0)), (b, 5e-4, 0.005, pa.FP32C, (0, 0, 0)),
            (c, 1e-4, 0.05, pa.FP64, (0, 0, -0.05)), (d, 2e-3, 0.01, pa.FP32C, (0, 0, 0))]


def _run_batch(jobs, steps, graph=True):
    B = pa.Batch(block=128)
    ids = [B.add(w, nu, dt, prec, uinf) for (w, nu, dt, prec, uinf) in jobs]
    B.step(steps, graph=graph)
    out = [B.get(i) for i in ids]
    st = [B.status(i) for i in ids]
    B.close()
    return out, st


def test_alone_equals_mixed_bitwise():
    jobs = _jobs()
    print("case mixed batch of 4 jobs, 3 steps", flush=True)
    mixed, st = _run_batch(jobs, 3)
    assert all(s == 0 for s in st)
    for k, job in enumerate(jobs):
        print(f"case job {k} alone", flush=True)
        alone, _ = _run_batch([job], 3)
        for key in ("pos", "vort", "vel", "dvort"):
            assert np.array_equal(alone[0][key], mixed[k][key]), (k, key)


def test_graph_equals_eager():
    jobs = _jobs()[:2]
    print("case graph vs eager", flush=True)
    g, _ = _run_batch(jobs, 2, graph=True)
    e, _ = _run_batch(jobs, 2, graph=False)
    for k in range(len(jobs)):
        for key in ("pos", "vort"):
            assert np.array_equal(g[k][key], e[k][key])


def test_batch_matches_oracle():
    jobs = _jobs()
    out, _ = _run_batch(jobs, 2)
    for k, (w, nu, dt, prec, uinf) in enumerate(jobs):
        x0 = orc.pack_state(w["pos"], w["vort"])
        xo = orc.rk4(x0, w["radius"], w["volume"], nu, dt, 2, mode=orc.OMP_CLEAN,
                     uinf=np.array(uinf, dtype=float))
        po, vo = orc.unpack_state(xo)
        ep, ea = per_particle_rel(out[k]["pos"], po), per_particle_rel(out[k]["vort"], vo)
        print(f"  job {k} prec {prec}: pos {ep:.3e} vort {ea:.3e}", flush=True)
        tol = 1e-11 if prec == pa.FP64 else 1e-6
        assert ep < tol and ea < (1e-9 if prec == pa.FP64 else 1e-5)


def test_server_alone_vs_mixed_and_failures():
    jobs = _jobs()
    alone = []
    for k, (w, nu, dt, prec, uinf) in enumerate(jobs):
        print(f"case server job {k} alone", flush=True)
        srv = pa.Server(8 << 30)
        jid = srv.submit(w, nu, dt, 3 + k, prec, uinf)
        while srv.step() > 0:
            pass
        assert srv.state(jid)[0] == pa.Server.DONE
        alone.append(srv.fetch(jid))
        srv.close()
    print("case server mixed with staggered arrivals, a failing job and a cancel", flush=True)
    srv = pa.Server(8 << 30)
    bad = gen.vortex_ring(8, 1, seed=5)
    bad["pos"][3] = bad["pos"][2]  # coincident pair -> non-finite rates
    ids = {}
    ids[0] = srv.submit(*_args(jobs[0], 3))
    srv.step()
    ids["bad"] = srv.submit(bad, 1e-3, 0.01, 5)
    ids[1] = srv.submit(*_args(jobs[1], 4))
    victim = srv.submit(*_args(jobs[2], 50))
    srv.step()
    ids[2] = srv.submit(*_args(jobs[2], 5))
    srv.cancel(victim)
    srv.step()
    ids[3] = srv.submit(*_args(jobs[3], 6))
    while srv.step() > 0:
        pass
    assert srv.state(ids["bad"])[0] == pa.Server.FAILED
    assert srv.state(victim)[0] == pa.Server.CANCELLED
    for k in range(4):
        assert srv.state(ids[k])[0] == pa.Server.DONE
        pos, vort = srv.fetch(ids[k])
        assert np.array_equal(pos, alone[k][0]) and np.array_equal(vort, alone[k][1]), k
    st = srv.stats()
    print(f"  stats {st}", flush=True)
    assert st["finished"] == 4 and st["failed"] == 1 and st["cancelled"] == 1
    assert st["hbm_in_use"] == 0
    srv.close()


def test_server_budget():
    srv = pa.Server(1 << 20)  # 1 MiB: a 3,200-particle job does not fit
    w = gen.vortex_ring(100, 32, seed=2)
    with pytest.raises(pa.PawanError):
        srv.submit(w, 1e-3, 0.01, 2)
    srv.close()


def _args(job, steps):
    w, nu, dt, prec, uinf = job
    return (w, nu, dt, steps, prec, uinf)


def test_gaussian_batch_default_variant():
    """A GAUSSIAN batch with the default variant (one target per thread for
    both precisions: the GAUSSIAN kernels are compiled so) matches the oracle."""
    jobs = _jobs()[:2]
    print("case GAUSSIAN batch, default variant", flush=True)
    B = pa.Batch(kernel=pa.KERNEL_GAUSSIAN)
    ids = [B.add(w, nu, dt, prec, uinf) for (w, nu, dt, prec, uinf) in jobs]
    B.step(2, graph=True)
    out = [B.get(i) for i in ids]
    B.close()
    for k, (w, nu, dt, prec, uinf) in enumerate(jobs):
        x0 = orc.pack_state(w["pos"], w["vort"])
        xo = orc.rk4(x0, w["radius"], w["volume"], nu, dt, 2, mode=orc.OMP_CLEAN, kind=orc.GAUSSIAN,
                     uinf=np.array(uinf, dtype=float))
        po, vo = orc.unpack_state(xo)
        ep, ea = per_particle_rel(out[k]["pos"], po), per_particle_rel(out[k]["vort"], vo)
        print(f"  job {k} prec {prec}: pos {ep:.3e} vort {ea:.3e}", flush=True)
        assert ep < (1e-11 if prec == pa.FP64 else 1e-6)
        assert ea < (1e-9 if prec == pa.FP64 else 1e-5)


def test_server_passes_by_stage():
    """A pass runs 4 RK4 stages of a job whose stage costs at most a quarter
    of the largest resident stage, and 1 stage of the largest: with a
    65,536-particle rotor wake resident, a 2,048-particle ring completes a
    step per pass while the wake completes one every 4 passes.  A cancel in
    the middle of the wake's step, and both results, are as when alone."""
    big = gen.rotor_wake(256)                          # 65,536
    small = gen.vortex_ring(128, 16, seed=4)           # 2,048
    mid = gen.vortex_ring(160, 64, R=1.0, core=0.1, seed=6)  # 10,240
    print("case server passes by stage: 65,536 + 2,048 + 10,240 (cancelled)", flush=True)
    alone = {}
    for name, w, nu, dt, steps in (("big", big, 1e-4, 1e-5, 2), ("small", small, 1e-3, 0.005, 3)):
        srv = pa.Server(64 << 30)
        jid = srv.submit(w, nu, dt, steps)
        while srv.step() > 0:
            pass
        alone[name] = srv.fetch(jid)
        srv.close()
    srv = pa.Server(64 << 30)
    jb = srv.submit(big, 1e-4, 1e-5, 2)
    js = srv.submit(small, 1e-3, 0.005, 3)
    jm = srv.submit(mid, 1e-3, 0.005, 50)
    for p in range(3):
        srv.step()
        assert srv.state(js)[1] == p + 1  # a whole step per pass
        assert srv.state(jb)[1] == 0      # one stage per pass
    assert srv.state(js)[0] == pa.Server.DONE
    srv.cancel(jm)  # in the middle of its run
    passes = 3
    while True:
        passes += 1
        if srv.step() == 0:
            break
    assert passes == 8  # 2 steps x 4 stages of the wake
    for name, jid in (("big", jb), ("small", js)):
        pos, vort = srv.fetch(jid)
        assert np.array_equal(pos, alone[name][0]) and np.array_equal(vort, alone[name][1]), name
    st = srv.stats()
    print(f"  passes {st['steps']}, step p50/p99 {st['step_p50_ms']:.2f}/{st['step_p99_ms']:.2f} ms, "
          f"job step p50 {st['job_step_p50_ms']:.2f} ms", flush=True)
    assert st["hbm_in_use"] == 0 and st["cancelled"] == 1
    srv.close()


@pytest.mark.parametrize("precision,n_shed", [(pa.FP64, 640), (pa.FP32C, 720)])
def test_server_slices_a_stage_larger_than_a_pass(precision, n_shed):
    """A stage costing more than a pass's cap (1.6e10 pair interactions, a
    binary32 pair priced at half) runs in slices: a 163,840-particle binary64
    or 184,320-particle binary32 rotor wake takes 2 passes per stage, 8 for
    its step, while a 2,048-particle ring takes a whole step every pass.  The
    wake's state equals a batch's, which runs each stage in one launch, bit
    for bit, and the ring's equals its state served alone."""
    big = gen.rotor_wake(n_shed)
    small = gen.vortex_ring(128, 16, seed=4)
    n = big["n"]
    print(f"case server slices: rotor wake {n} precision {precision} + ring 2,048", flush=True)
    B = pa.Batch()
    jb0 = B.add(big, 1e-4, 1e-5, precision)
    B.step(1)
    ref_big = B.get(jb0)
    B.close()
    srv = pa.Server(64 << 30)
    js0 = srv.submit(small, 1e-3, 0.005, 3)
    while srv.step() > 0:
        pass
    ref_small = srv.fetch(js0)
    srv.close()
    srv = pa.Server(64 << 30)
    jb = srv.submit(big, 1e-4, 1e-5, 1, precision)
    js = srv.submit(small, 1e-3, 0.005, 3)
    passes = 0
    while True:
        passes += 1
        left = srv.step()
        if passes <= 3:
            assert srv.state(js)[1] == passes  # a whole step per pass
        if passes < 8:
            assert srv.state(jb) == (pa.Server.RUNNING, 0)
        if left == 0:
            break
    assert passes == 8  # 4 stages x 2 slices
    pos, vort = srv.fetch(jb)
    assert np.array_equal(pos, ref_big["pos"]) and np.array_equal(vort, ref_big["vort"])
    pos, vort = srv.fetch(js)
    assert np.array_equal(pos, ref_small[0]) and np.array_equal(vort, ref_small[1])
    st = srv.stats()
    pairs = 4.0 * n * (n - 1) + 3 * 4.0 * 2048 * 2047
    print(f"  passes {st['steps']}, step p50/p99 {st['step_p50_ms']:.2f}/{st['step_p99_ms']:.2f} ms, "
          f"pairs {st['pairs']:.4e} (expected {pairs:.4e})", flush=True)
    assert abs(st["pairs"] - pairs) < 1e-9 * pairs
    srv.close()


def test_server_large_jobs_take_turns():
    """Two large jobs (65,536-particle rotor wakes, one stage each the whole
    reference cost) take turns pass by pass while a small ring takes a whole
    step every pass: the ring is done after 2 passes, each wake after 4 of
    its own stages (passes 7 and 8), and all three results equal their
    results alone."""
    big = gen.rotor_wake(256)
    small = gen.vortex_ring(128, 16, seed=4)
    jobs = [(big, 1e-4, 1e-5, 1), (dict(big, pos=big["pos"] + np.array([0.0, 0.0, 3.0])), 1e-4,
                                   1e-5, 1), (small, 1e-3, 0.005, 2)]
    print("case server: two 65,536-particle wakes taking turns, one 2,048-particle ring", flush=True)
    alone = []
    for w, nu, dt, steps in jobs:
        srv = pa.Server(64 << 30)
        jid = srv.submit(w, nu, dt, steps)
        while srv.step() > 0:
            pass
        alone.append(srv.fetch(jid))
        srv.close()
    srv = pa.Server(64 << 30)
    ids = [srv.submit(w, nu, dt, steps) for w, nu, dt, steps in jobs]
    done_at = {}
    for p in range(1, 9):
        running = srv.step()
        for k, i in enumerate(ids):
            if k not in done_at and srv.state(i)[0] == pa.Server.DONE:
                done_at[k] = p
        if p == 2:
            assert srv.state(ids[0])[1] == 0 and srv.state(ids[1])[1] == 0
    assert running == 0
    assert done_at == {2: 2, 0: 7, 1: 8}, done_at
    for k, i in enumerate(ids):
        pos, vort = srv.fetch(i)
        assert np.array_equal(pos, alone[k][0]) and np.array_equal(vort, alone[k][1]), k
    srv.close()
