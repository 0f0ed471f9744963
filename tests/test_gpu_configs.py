"""The workload configs at their stated sizes, checked against the CPU oracle.

* rotor wake, N = 262,144 (the metric's size): rates of a stratified sample of
  targets against all 262,144 sources from the oracle (the reference's
  binary64 pair arithmetic, oracle/vpm_oracle.c orc_rates_sample, restating
  interaction_utils.h:429-481 summed as parallel.cpp:13-85 does), binary64
  path <= 1e-10 per particle, compensated binary32 <= 1e-6 relative RMS per
  term (velocity, stretching, diffusion);
* two leapfrogging rings, 2 x 32,768 particles as two wakes
  (interaction.cpp:24-27,100-107): the same, and the state after RK4 steps;
* one ring, 2,048 particles, RK4 dt 0.01 for the full 100 steps
  (rk4.cpp:12-77, integration.cpp:16-35) against the oracle's RK4;
* 256 batched rings, N log-uniform in [1,024, 65,536]: rates of sampled jobs
  against the oracle, graph-replayed steps, a job alone == in the batch;
* a serving stream drawn from the serving distribution (N log-uniform in
  [2,048, 262,144], both precisions, 200 GB budget) including two jobs of
  262,144 particles: every job's result alone == in the mix, bit for bit.

Bars: binary64 rates 1e-10 per particle (relative to max(|r_i|, 1e-3 RMS));
binary32 1e-6 relative RMS for velocity, stretching, diffusion and the
vorticity rate.  (In smooth ring cores the two PSE sums v_i sum k a_j and
a_i sum k v_j agree to a digit or two; close pairs therefore add the
differences a_j - a_i, v_j - v_i formed in binary64, which keeps the
diffusion term there at 3-8e-7 instead of 1.5-2e-6.)
"""
import numpy as np
import pytest

import oracle as orc
import pawan_amd as pa
from pawan_amd import generators as gen
from helpers import per_particle_rel, rel_rms, sample_targets

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 1e-6


def _wakes(w):
    offs = np.cumsum([0] + list(w.get("wakes", [w["n"]])))
    return [pa.Wake(w["pos"][a:b], w["vort"][a:b], w["radius"][a:b], w["volume"][a:b])
            for a, b in zip(offs[:-1], offs[1:])]


def _split(r, wakes):
    """[pos 3n | vort 3n] per wake -> (velocities, vorticity rates), N x 3."""
    v, d, o = [], [], 0
    for m in wakes:
        v.append(r[o:o + 3 * m].reshape(m, 3))
        d.append(r[o + 3 * m:o + 6 * m].reshape(m, 3))
        o += 6 * m
    return np.concatenate(v), np.concatenate(d)


def _gpu_terms(w, precision, **kw):
    """GPU velocity, vorticity rate, stretching (nu = 0) and diffusion
    (difference of the two) at the particles' state."""
    wakes = w.get("wakes", [w["n"]])
    S = pa.Interaction(*_wakes(w), nu=w["nu"], uinf=w.get("uinf", (0, 0, 0)),
                       precision=precision, **kw)
    S.solve()
    v, d = _split(S.getRates(), wakes)
    S.setNu(0.0)
    S.solve()
    _, s = _split(S.getRates(), wakes)
    S.close()
    return v, d, s, d - s


def _oracle_terms(w, targets):
    v, s, f = orc.rates_sample(w["pos"], w["vort"], w["radius"], w["volume"], w["nu"], targets,
                               uinf=w.get("uinf"))
    return v, s + f, s, f


def _check(w, name, precision, targets, ref, **kw):
    v, d, s, f = _gpu_terms(w, precision, **kw)
    rv, rd, rs, rf = ref
    t = targets
    if precision == pa.FP64:
        ev, ed = per_particle_rel(v[t], rv), per_particle_rel(d[t], rd)
        print(f"  {name} fp64 vs oracle, {len(t)} targets: velocity {ev:.3e} "
              f"vorticity rate {ed:.3e} (max per particle)", flush=True)
        assert ev < TOL64 and ed < TOL64
    else:
        ev, ed = rel_rms(v[t], rv), rel_rms(d[t], rd)
        es, ef = rel_rms(s[t], rs), rel_rms(f[t], rf)
        print(f"  {name} fp32c vs oracle, {len(t)} targets (rel RMS): velocity {ev:.3e} "
              f"stretching {es:.3e} diffusion {ef:.3e} vorticity rate {ed:.3e}", flush=True)
        assert ev < TOL32 and es < TOL32 and ed < TOL32 and ef < TOL32


# ----------------------------------------------------------- N = 262,144
@pytest.fixture(scope="module")
def rotor():
    w = gen.config_rotor_wake(1024)
    t = sample_targets(w["n"], 768, seed=5)
    print(f"case rotor wake N={w['n']}: oracle on {len(t)} sampled targets", flush=True)
    return w, t, _oracle_terms(w, t)


def test_rotor262k_fp64_vs_oracle(rotor):
    w, t, ref = rotor
    _check(w, "rotor 262k one-sided (plain)", pa.FP64, t, ref)
    _check(w, "rotor 262k symmetric option", pa.FP64, t, ref, scheme=pa.SCHEME_SYMMETRIC)


def test_rotor262k_fp32c_vs_oracle(rotor):
    w, t, ref = rotor
    _check(w, "rotor 262k default (256 x 1)", pa.FP32C, t, ref)
    _check(w, "rotor 262k 512 x 1", pa.FP32C, t, ref, block=512)
    _check(w, "rotor 262k one-sided", pa.FP32C, t, ref, scheme=pa.SCHEME_ONESIDED)


# ------------------------------------------------- leapfrog, 2 x 32,768
@pytest.fixture(scope="module")
def leapfrog():
    w = gen.config_leapfrog()
    t = sample_targets(w["n"], 1024, seed=6)
    print(f"case leapfrog 2 x {w['wakes'][0]} (two wakes): oracle on {len(t)} targets",
          flush=True)
    return w, t, _oracle_terms(w, t)


def test_leapfrog_rates_vs_oracle(leapfrog):
    w, t, ref = leapfrog
    _check(w, "leapfrog", pa.FP64, t, ref)
    _check(w, "leapfrog", pa.FP32C, t, ref)
    _check(w, "leapfrog one-sided", pa.FP32C, t, ref, scheme=pa.SCHEME_ONESIDED)


def test_leapfrog_state_after_steps_vs_oracle(leapfrog):
    """Four RK4 steps (dt 0.005) on the device, then the rates at the new
    state against the oracle's rates of the GPU's own state (so the check is
    of the step's stage combinations and the pass at a moved state), and the
    state against the binary32 run."""
    w, t, _ = leapfrog
    states = {}
    for prec in (pa.FP64, pa.FP32C):
        S = pa.Interaction(*_wakes(w), nu=w["nu"], precision=prec)
        S.rk4(w["dt"], 4)
        x = S.getStates()
        S.solve()
        r = S.getRates()
        S.close()
        states[prec] = (x, r)
    x, r = states[pa.FP64]
    m = w["wakes"][0]
    pos = np.concatenate([x[:3 * m], x[6 * m:9 * m]]).reshape(-1, 3)
    vort = np.concatenate([x[3 * m:6 * m], x[9 * m:]]).reshape(-1, 3)
    rv, rd = orc.rates_sample(pos, vort, w["radius"], w["volume"], w["nu"], t, split=False)
    v, d = _split(r, w["wakes"])
    ev, ed = per_particle_rel(v[t], rv), per_particle_rel(d[t], rd)
    print(f"  leapfrog after 4 steps fp64 vs oracle: velocity {ev:.3e} vorticity rate {ed:.3e}",
          flush=True)
    assert ev < TOL64 and ed < TOL64
    x32 = states[pa.FP32C][0]
    ep, ea = rel_rms(np.concatenate([x32[:3 * m], x32[6 * m:9 * m]]), pos.ravel()), \
        rel_rms(np.concatenate([x32[3 * m:6 * m], x32[9 * m:]]), vort.ravel())
    print(f"  leapfrog after 4 steps fp32c vs fp64 state: positions {ep:.3e} "
          f"vorticities {ea:.3e}", flush=True)
    assert ep < 1e-9 and ea < 1e-7


# ------------------------------------------------- single ring, 100 steps
def test_single_ring_100_steps_vs_oracle():
    """The single-ring config at its stated step count: 2,048 particles, RK4
    dt 0.01 x 100 on the device against the oracle's rk4.cpp restatement
    (OpenMP sweep), with the drift of total vorticity, linear impulse and
    kinetic energy reported by both."""
    w = gen.config_single_ring()
    print(f"case single ring N={w['n']}, {w['steps']} RK4 steps of {w['dt']}", flush=True)
    S = pa.Interaction(pa.Wake.from_dict(w), nu=w["nu"])
    d0 = S.diagnose()
    x0 = S.getStates()
    S.rk4(w["dt"], w["steps"])
    x = S.getStates()
    d1 = S.diagnose()
    S.close()
    xo = orc.rk4(x0, w["radius"], w["volume"], w["nu"], w["dt"], w["steps"], mode=orc.OMP_CLEAN)
    n = w["n"]
    ep = per_particle_rel(x[:3 * n], xo[:3 * n])
    ea = per_particle_rel(x[3 * n:], xo[3 * n:])
    po, vo = orc.unpack_state(xo)
    do = orc.diagnostics(po, vo, w["radius"])
    print(f"  after 100 steps: positions {ep:.3e} vorticities {ea:.3e} (max per particle)",
          flush=True)
    for k in ("total_vorticity", "linear_impulse"):
        print(f"  drift {k}: gpu {np.linalg.norm(np.subtract(d1[k], d0[k])):.3e} "
              f"oracle {np.linalg.norm(np.subtract(do[k], d0[k])):.3e}", flush=True)
    ke_g = (d1["kinetic_energy"] - d0["kinetic_energy"]) / d0["kinetic_energy"]
    ke_o = (do["kinetic_energy"] - d0["kinetic_energy"]) / d0["kinetic_energy"]
    print(f"  drift kinetic energy (relative): gpu {ke_g:.6e} oracle {ke_o:.6e}", flush=True)
    assert ep < 1e-10 and ea < 1e-9
    assert np.allclose(d1["total_vorticity"], do["total_vorticity"], rtol=0, atol=1e-13)
    assert np.allclose(d1["linear_impulse"], do["linear_impulse"], rtol=1e-10, atol=1e-14)
    assert abs(ke_g - ke_o) < 1e-9


# ----------------------------------------------------- 256 batched rings
@pytest.fixture(scope="module")
def batch256():
    return gen.batch_jobs(256, seed=0)


def test_batch_rings_fp32c_per_term(batch256):
    """Compensated binary32 per term (velocity, stretching, diffusion) on
    rings of the batch: the two smallest in full, the two largest on sampled
    targets, each as its own context (the batch runs the same pass)."""
    order = sorted(range(len(batch256)), key=lambda k: batch256[k]["n"])
    for k in order[:2] + order[-2:]:
        j = batch256[k]
        t = np.arange(j["n"]) if j["n"] < 4096 else sample_targets(j["n"], 256, seed=k)
        print(f"case batch ring {k} N={j['n']} ({len(t)} targets)", flush=True)
        _check(j, f"batch ring {k}", pa.FP32C, t, _oracle_terms(j, t))


@pytest.mark.parametrize("precision", [pa.FP64, pa.FP32C])
def test_batch_256_rings(batch256, precision):
    jobs = batch256
    tot = sum(j["n"] for j in jobs)
    print(f"case batch of {len(jobs)} rings ({tot} particles), precision {precision}",
          flush=True)
    B = pa.Batch()
    ids = [B.add(j, j["nu"], j["dt"], precision) for j in jobs]
    B.solve()
    order = sorted(range(len(jobs)), key=lambda k: jobs[k]["n"])
    # the 3 smallest jobs fully, the 2 largest and 3 others on sampled targets
    full = order[:3]
    sampled = order[-2:] + order[len(order) // 4::len(order) // 4][:3]
    for k in full + sampled:
        j = jobs[k]
        g = B.get(ids[k])
        if k in full:
            t = np.arange(j["n"])
            rv, rd = orc.rates(j["pos"], j["vort"], j["radius"], j["volume"], j["nu"],
                               mode=orc.OMP_CLEAN)
        else:
            t = sample_targets(j["n"], 256, seed=k)
            rv, rd = orc.rates_sample(j["pos"], j["vort"], j["radius"], j["volume"], j["nu"], t,
                                      split=False)
        v, d = g["vel"][t], g["dvort"][t]
        if precision == pa.FP64:
            ev, ed = per_particle_rel(v, rv), per_particle_rel(d, rd)
            assert ev < TOL64 and ed < TOL64, (k, ev, ed)
        else:
            ev, ed = rel_rms(v, rv), rel_rms(d, rd)
            assert ev < TOL32 and ed < TOL32, (k, ev, ed)
        print(f"  job {k} N={j['n']} ({len(t)} targets): velocity {ev:.3e} "
              f"vorticity rate {ed:.3e}", flush=True)
    # two graph-replayed RK4 steps of the whole batch
    B.step(2, graph=True)
    assert all(B.status(i) == 0 for i in ids)
    mixed = {k: B.get(ids[k]) for k in order[:2] + order[-1:]}
    B.close()
    # the smallest job's state against the oracle's two RK4 steps
    k = order[0]
    j = jobs[k]
    xo = orc.rk4(orc.pack_state(j["pos"], j["vort"]), j["radius"], j["volume"], j["nu"], j["dt"],
                 2, mode=orc.OMP_CLEAN)
    po, vo = orc.unpack_state(xo)
    ep, ea = per_particle_rel(mixed[k]["pos"], po), per_particle_rel(mixed[k]["vort"], vo)
    print(f"  job {k} after 2 steps vs oracle: positions {ep:.3e} vorticities {ea:.3e}",
          flush=True)
    if precision == pa.FP64:
        assert ep < 1e-11 and ea < 1e-9
    else:
        assert ep < 1e-6 and ea < 1e-5
    # the same three jobs alone in their own batch: bitwise the same
    B = pa.Batch()
    alone_ids = {k: B.add(jobs[k], jobs[k]["nu"], jobs[k]["dt"], precision) for k in mixed}
    B.step(2, graph=True)
    for k, i in alone_ids.items():
        a = B.get(i)
        for key in ("pos", "vort", "vel", "dvort"):
            assert np.array_equal(a[key], mixed[k][key]), (k, key)
    B.close()


# ------------------------------------------------------- serving stream
def test_serving_stream_alone_equals_mixed():
    """Jobs drawn from the serving distribution (seed 3), plus a 262,144-
    particle rotor wake in binary64 and a 262,144-particle ring in compensated
    binary32, submitted while the server runs (three arrive after it has
    started stepping); steps are capped at 3 per job so that the test stays
    short.  Every job's final state equals its state when served alone."""
    stream = gen.serving_stream(8, seed=3)
    big = [dict(kind="rotor", w=gen.config_rotor_wake(1024), nu=1e-4, dt=1e-5,
                uinf=(0.0, 0.0, -0.05), precision=pa.FP64),
           dict(kind="ring", w=gen.vortex_ring(2048, 128, R=1.0, core=0.1, seed=9), nu=1e-3,
                dt=0.005, uinf=(0.0, 0.0, 0.0), precision=pa.FP32C)]
    jobs = big + stream
    for k, j in enumerate(jobs):
        j["steps"] = 1 + k % 3
    print("case serving stream: " + ", ".join(
        f"{j['kind']} {j['w']['n']} {'fp64' if j['precision'] == 0 else 'fp32c'} x{j['steps']}"
        for j in jobs), flush=True)
    budget = 200 << 30

    def submit(srv, j):
        return srv.submit(j["w"], j["nu"], j["dt"], j["steps"], j["precision"], j["uinf"])

    srv = pa.Server(budget)
    ids = [submit(srv, j) for j in jobs[:-3]]
    srv.step()
    ids += [submit(srv, j) for j in jobs[-3:]]
    while srv.step() > 0:
        pass
    st = srv.stats()
    print(f"  mixed: {st['finished']} finished, step p50/p99 {st['step_p50_ms']:.1f}/"
          f"{st['step_p99_ms']:.1f} ms", flush=True)
    mixed = []
    for i in ids:
        assert srv.state(i)[0] == pa.Server.DONE
        mixed.append(srv.fetch(i))
    assert srv.stats()["hbm_in_use"] == 0
    srv.close()
    for k, j in enumerate(jobs):
        srv = pa.Server(budget)
        i = submit(srv, j)
        while srv.step() > 0:
            pass
        pos, vort = srv.fetch(i)
        srv.close()
        assert np.array_equal(pos, mixed[k][0]) and np.array_equal(vort, mixed[k][1]), k
    print(f"  {len(jobs)} jobs: alone == mixed bitwise", flush=True)
