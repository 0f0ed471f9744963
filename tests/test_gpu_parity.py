"""Parity of the HIP path with the CPU oracle (restatement of pawan's
reference method), through the C ABI.  Tolerances: binary64 rates 1e-10 per
particle relative (floor 1e-3 x RMS); compensated binary32 1e-6 relative RMS."""
import numpy as np
import pytest

import oracle as orc
import pawan_amd as pa
from pawan_amd import generators as gen
from helpers import per_particle_rel, rel_rms

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 1e-6


def _gpu_rates(w, nu, wake_n=None, precision=pa.FP64, kernel=pa.KERNEL_HIGHORDER, uinf=(0, 0, 0),
               block=0, chunk=0):
    wake_n = wake_n or [w["pos"].shape[0]]
    offs = np.cumsum([0] + list(wake_n))
    wakes = [pa.Wake(w["pos"][a:b], w["vort"][a:b], w["radius"][a:b], w["volume"][a:b])
             for a, b in zip(offs[:-1], offs[1:])]
    S = pa.Interaction(*wakes, nu=nu, precision=precision, kernel=kernel, uinf=uinf, block=block,
                       chunk=chunk)
    S.solve()
    r = S.getRates()
    S.close()
    pos, _ = orc.unpack_state(np.zeros(6 * offs[-1]), wake_n)  # shape helper
    vel, dv = orc.unpack_state(r, wake_n)
    return vel, dv


@pytest.mark.parametrize("block", [64, 128, 256])
def test_single_ring_rates_fp64(block):
    w = gen.config_single_ring()
    print(f"case single ring N={w['n']} block={block}", flush=True)
    vel, dv = _gpu_rates(w, w["nu"], block=block)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], w["nu"], mode=orc.OMP_CLEAN)
    ev, ed = per_particle_rel(vel, rv), per_particle_rel(dv, rd)
    print(f"  vel {ev:.3e} dvort {ed:.3e}", flush=True)
    assert ev < TOL64 and ed < TOL64


def test_two_wakes_match_single_concatenation():
    """__interaction(W1, W2) (interaction.cpp:24-27) == one wake holding both."""
    a = gen.vortex_ring(32, 16, R=1.0, core=0.1, seed=3)
    b = gen.vortex_ring(24, 16, R=0.8, core=0.1, seed=4, center=(0, 0, 0.3))
    w = {k: np.concatenate([a[k], b[k]]) for k in ("pos", "vort", "radius", "volume")}
    print(f"case two wakes {a['n']}+{b['n']}", flush=True)
    nu = 1e-3
    v2, d2 = _gpu_rates(w, nu, wake_n=[a["n"], b["n"]])
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL,
                       wake_n=[a["n"], b["n"]])
    assert per_particle_rel(v2, rv) < TOL64 and per_particle_rel(d2, rd) < TOL64


def test_nonuniform_core_and_gaussian():
    w = gen.vortex_ring(48, 16, R=1.0, core=0.1, seed=5)
    w["radius"] = w["radius"] * (1.0 + 0.2 * gen.uniform(99, w["n"]))
    nu = 2.5e-3
    for kernel, okind in ((pa.KERNEL_HIGHORDER, orc.HIGHORDER), (pa.KERNEL_GAUSSIAN, orc.GAUSSIAN)):
        print(f"case nonuniform core kernel={kernel}", flush=True)
        vel, dv = _gpu_rates(w, nu, kernel=kernel)
        rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL,
                           kind=okind)
        ev, ed = per_particle_rel(vel, rv), per_particle_rel(dv, rd)
        print(f"  vel {ev:.3e} dvort {ed:.3e}", flush=True)
        assert ev < TOL64 and ed < TOL64


def test_freestream_and_chunked_sources():
    w = gen.rotor_wake(16)  # N = 4096
    nu = 1e-4
    uinf = (0.0, 0.0, -0.05)
    print(f"case rotor wake N={w['n']} chunk 512", flush=True)
    vel, dv = _gpu_rates(w, nu, uinf=uinf, block=128, chunk=512)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.OMP_CLEAN,
                       uinf=np.array(uinf))
    ev, ed = per_particle_rel(vel, rv), per_particle_rel(dv, rd)
    print(f"  vel {ev:.3e} dvort {ed:.3e}", flush=True)
    assert ev < TOL64 and ed < TOL64


def test_ragged_sizes():
    """N not a multiple of any tile; tiny systems (1, 2, 3 particles)."""
    for n_azi, n_cs in ((1, 1), (2, 1), (3, 1), (7, 9), (33, 7)):
        w = gen.vortex_ring(n_azi, n_cs, R=1.0, core=0.1, seed=n_azi)
        print(f"case ragged N={w['n']}", flush=True)
        vel, dv = _gpu_rates(w, 1e-3)
        rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], 1e-3, mode=orc.SERIAL)
        if w["n"] == 1:
            assert np.all(vel == 0) and np.all(dv == 0)
        else:
            assert per_particle_rel(vel, rv) < TOL64 and per_particle_rel(dv, rd) < TOL64


def test_empty_inputs():
    """Empty input: the reference cannot hold an empty wake (wake.cpp:10-21
    allocates its matrices with gsl_matrix_calloc, which rejects a zero
    dimension), so a wake or batch job of no particles is refused with
    PW_EINVAL and the context stays usable; a solve before any wake is
    refused the same way; appending no particles changes nothing."""
    from pawan_amd import _lib
    print("case empty inputs", flush=True)
    z3, z1 = np.zeros((0, 3)), np.zeros(0)
    S = pa.Interaction(nu=1e-3)
    with pytest.raises(pa.PawanError) as e:
        S.solve()
    assert e.value.code == _lib.PW_EINVAL
    with pytest.raises(pa.PawanError) as e:
        S.addWake(pa.Wake(z3, z3, z1, z1))
    assert e.value.code == _lib.PW_EINVAL
    w = gen.vortex_ring(33, 7, R=1.0, core=0.1, seed=5)
    S.addWake(pa.Wake.from_dict(w))
    assert S._size == 6 * w["n"]
    S.solve()
    r0 = S.getRates().copy()
    S.appendParticles(0, z3, z3, z1, z1)
    assert S._size == 6 * w["n"]
    S.solve()
    r1 = S.getRates()
    S.close()
    assert np.array_equal(r0, r1)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], 1e-3, mode=orc.SERIAL)
    n = w["n"]
    assert per_particle_rel(r0[:3 * n].reshape(n, 3), rv) < TOL64
    assert per_particle_rel(r0[3 * n:].reshape(n, 3), rd) < TOL64
    B = pa.Batch()
    with pytest.raises(pa.PawanError) as e:
        B.add({"pos": z3, "vort": z3, "radius": z1, "volume": z1}, 1e-3, 0.01)
    assert e.value.code == _lib.PW_EINVAL
    B.close()


def test_rk4_trajectory_single_ring():
    """Config 1 for 10 RK4 steps (dt = 0.01) against the oracle's rk4.cpp."""
    w = gen.config_single_ring()
    print(f"case rk4 trajectory N={w['n']} 10 steps", flush=True)
    S = pa.Interaction(pa.Wake.from_dict(w), nu=w["nu"])
    x0 = S.getStates()
    S.rk4(w["dt"], 10)
    x = S.getStates()
    d0 = S.diagnose()
    S.close()
    xo = orc.rk4(x0, w["radius"], w["volume"], w["nu"], w["dt"], 10, mode=orc.OMP_CLEAN)
    n = w["n"]
    ep = per_particle_rel(x[:3 * n], xo[:3 * n])
    ea = per_particle_rel(x[3 * n:], xo[3 * n:])
    print(f"  pos {ep:.3e} vort {ea:.3e}", flush=True)
    assert ep < 1e-11 and ea < 1e-9
    po, vo = orc.unpack_state(xo)
    do = orc.diagnostics(po, vo, w["radius"])
    assert np.allclose(d0["total_vorticity"], do["total_vorticity"], atol=1e-14)
    assert np.allclose(d0["linear_impulse"], do["linear_impulse"], rtol=1e-10, atol=1e-14)
    assert abs(d0["kinetic_energy"] - do["kinetic_energy"]) < 1e-10 * abs(do["kinetic_energy"])


def test_euler_step():
    w = gen.vortex_ring(32, 16, seed=11)
    S = pa.Interaction(pa.Wake.from_dict(w), nu=1e-3)
    x0 = S.getStates()
    S.euler(0.01, 2)
    x = S.getStates()
    S.close()
    xo = orc.rk4(x0, w["radius"], w["volume"], 1e-3, 0.01, 2, mode=orc.SERIAL, euler=True)
    assert np.max(np.abs(x - xo)) <= 1e-13 * np.max(np.abs(xo))


def test_set_state_roundtrip_and_rates_layout():
    """setStates/getStates/getRates use the wake.cpp:87-120 layout exactly."""
    a = gen.vortex_ring(8, 16, seed=1)
    b = gen.vortex_ring(5, 9, seed=2, center=(0, 0, 1))
    S = pa.Interaction(pa.Wake.from_dict(a), pa.Wake.from_dict(b), nu=1e-3)
    st = np.arange(S._size, dtype=np.float64) * 1e-3 + 0.25
    S.setStates(st)
    assert np.array_equal(S.getStates(), st)
    S.close()


def test_fp32c_error_per_term():
    """Compensated binary32: relative RMS error vs the binary64 oracle,
    stated per term (velocity, stretching, diffusion), bound 1e-6."""
    w = gen.config_single_ring()
    nu = w["nu"]
    vel, dv = _gpu_rates(w, nu, precision=pa.FP32C)
    rv, rs, rdf = orc.rates_split(w["pos"], w["vort"], w["radius"], w["volume"], nu)
    ev = rel_rms(vel, rv)
    ed = rel_rms(dv, rs + rdf)
    # per-term: run with nu = 0 (stretching only) and compare the difference
    v0, d0 = _gpu_rates(w, 0.0, precision=pa.FP32C)
    es = rel_rms(d0, rs)
    edf = rel_rms(dv - d0, rdf)
    print(f"  fp32c rel rms: velocity {ev:.3e} stretching {es:.3e} diffusion {edf:.3e} "
          f"total dvort {ed:.3e}", flush=True)
    assert ev < TOL32 and ed < TOL32 and es < TOL32 and edf < TOL32


def test_fp32c_error_rotor_wake():
    """Compensated binary32 on the rotor wake (16,384 particles), whose
    stretching sums cancel strongly: relative RMS error per term <= 1e-6."""
    w = gen.rotor_wake(64)
    nu = 1e-4
    print(f"case fp32c rotor wake N={w['n']}", flush=True)
    vel, dv = _gpu_rates(w, nu, precision=pa.FP32C)
    v0, d0 = _gpu_rates(w, 0.0, precision=pa.FP32C)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.OMP_CLEAN)
    _, rs = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], 0.0, mode=orc.OMP_CLEAN)
    ev, es, edf, ed = rel_rms(vel, rv), rel_rms(d0, rs), rel_rms(dv - d0, rd - rs), rel_rms(dv, rd)
    print(f"  fp32c rel rms: velocity {ev:.3e} stretching {es:.3e} diffusion {edf:.3e} "
          f"vorticity rate {ed:.3e}", flush=True)
    assert ev < TOL32 and es < TOL32 and ed < TOL32 and edf < TOL32
