"""Next rows of the path on the GPU: influence()/resolve(), the diagnostics
diagnose() prints, the .wake frame writer and particle emission."""
import numpy as np
import pytest

import oracle as orc
import pawan_amd as pa
from pawan_amd import generators as gen
from helpers import per_particle_rel, rel_rms

pytestmark = pytest.mark.gpu


def _two_wakes():
    a = gen.vortex_ring(40, 16, R=1.0, core=0.1, seed=21)
    b = gen.vortex_ring(30, 9, R=0.7, core=0.12, seed=22, center=(0.1, 0.0, 0.4))
    w = {k: np.concatenate([a[k], b[k]]) for k in ("pos", "vort", "radius", "volume")}
    return a, b, w


@pytest.mark.parametrize("kernel,okind", [(pa.KERNEL_HIGHORDER, orc.HIGHORDER),
                                          (pa.KERNEL_GAUSSIAN, orc.GAUSSIAN)])
def test_resolve_influence(kernel, okind):
    a, b, w = _two_wakes()
    print(f"case influence kernel={kernel} N={a['n']}+{b['n']}", flush=True)
    S = pa.Interaction(pa.Wake.from_dict(a), pa.Wake.from_dict(b), nu=1e-3, kernel=kernel)
    f = S.resolve()
    S.close()
    ref = orc.influence(w["pos"], w["vort"], w["radius"], kind=okind, wake_n=[a["n"], b["n"]])
    err = per_particle_rel(f, ref)
    print(f"  field {err:.3e}", flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("kernel,okind", [(pa.KERNEL_HIGHORDER, orc.HIGHORDER),
                                          (pa.KERNEL_GAUSSIAN, orc.GAUSSIAN)])
def test_diagnose_all(kernel, okind):
    a, b, w = _two_wakes()
    print(f"case diagnose kernel={kernel}", flush=True)
    S = pa.Interaction(pa.Wake.from_dict(a), pa.Wake.from_dict(b), nu=1e-3, kernel=kernel)
    d = S.diagnose_all()
    S.close()
    wn = [a["n"], b["n"]]
    do = orc.diagnostics(w["pos"], w["vort"], w["radius"], wake_n=wn)
    # the sums cancel (a ring's total vorticity is ~0), so compare against the
    # scale of the summands: sum |a|, sum |x||a|, sum |x|^2 |a|
    na = np.linalg.norm(w["vort"], axis=1)
    nx = np.linalg.norm(w["pos"], axis=1)
    for k, scale in (("total_vorticity", na.sum()), ("linear_impulse", (nx * na).sum()),
                     ("angular_impulse", (nx * nx * na).sum())):
        assert np.abs(d[k] - do[k]).max() <= 1e-14 * scale, (k, d[k], do[k])
    assert abs(d["kinetic_energy"] - do["kinetic_energy"]) <= 1e-12 * abs(do["kinetic_energy"])
    H = orc.helicity(w["pos"], w["vort"], w["radius"], wake_n=wn) if okind == orc.HIGHORDER else None
    if H is not None:
        assert abs(d["helicity"] - H) <= 1e-10 * max(abs(H), 1e-12)
        E = orc.enstrophy(w["pos"], w["vort"], w["radius"], wake_n=wn)
        assert abs(d["enstrophy"] - E) <= 1e-12 * abs(E)
    else:
        assert d["enstrophy"] == 0.0
    print(f"  {d}", flush=True)


def test_write_frame_matches_reference_format(tmp_path):
    a, b, w = _two_wakes()
    S = pa.Interaction(pa.Wake.from_dict(a), pa.Wake.from_dict(b), nu=1e-3)
    path = tmp_path / "run.wake"
    S.write(path, t=0.0, append=False)
    S.rk4(0.01, 2)
    S.write(path, t=0.02)
    st = S.getStates()
    S.close()
    raw = path.read_bytes()
    # parse as scripts/readWake.py does: f64 t, u64 nWake, per wake u64 N, N x 3, N x 3, N
    off, frames = 0, []
    while off < len(raw):
        t = np.frombuffer(raw, np.float64, 1, off)[0]
        off += 8
        nw = int(np.frombuffer(raw, np.uint64, 1, off)[0])
        off += 8
        wakes = []
        for _ in range(nw):
            n = int(np.frombuffer(raw, np.uint64, 1, off)[0])
            off += 8
            p = np.frombuffer(raw, np.float64, 3 * n, off).reshape(n, 3)
            off += 24 * n
            v = np.frombuffer(raw, np.float64, 3 * n, off).reshape(n, 3)
            off += 24 * n
            r = np.frombuffer(raw, np.float64, n, off)
            off += 8 * n
            wakes.append((p, v, r))
        frames.append((t, wakes))
    assert [f[0] for f in frames] == [0.0, 0.02]
    (p0, v0, r0), (p1, v1, r1) = frames[0][1]
    assert np.array_equal(p0, a["pos"]) and np.array_equal(v1, b["vort"])
    assert np.array_equal(r0, a["radius"]) and np.array_equal(r1, b["radius"])
    last = frames[1][1]
    na = a["n"]
    assert np.array_equal(last[0][0].ravel(), st[:3 * na])
    assert np.array_equal(last[1][1].ravel(), st[6 * na + 3 * b["n"]:])


@pytest.mark.parametrize("precision,tol_pos,tol_vort", [(pa.FP64, 1e-11, 1e-9),
                                                         (pa.FP32C, 1e-6, 1e-5)])
def test_emission_growing_rotor_wake(precision, tol_pos, tol_vort):
    """Rotor shedding: every step appends the new row of particles, then one
    RK4 step; the oracle repeats the same sequence on growing arrays.  The
    stride grows 256 -> 2,048 on the way, and so do the binary32 symmetric
    pass's source-side slots."""
    nu, dt, steps = 1e-4, 1e-3, 5
    uinf = np.array([0.0, 0.0, -0.05])
    e0 = gen.rotor_emission(0)
    S = pa.Interaction(pa.Wake.from_dict(e0), nu=nu, uinf=tuple(uinf), precision=precision)
    rad, vol = e0["radius"].copy(), e0["volume"].copy()
    x = orc.pack_state(e0["pos"], e0["vort"])
    for k in range(steps):
        if k > 0:
            e = gen.rotor_emission(k)
            S.appendParticles(0, e["pos"], e["vort"], e["radius"], e["volume"])
            p, v = orc.unpack_state(x)
            x = orc.pack_state(np.concatenate([p, e["pos"]]), np.concatenate([v, e["vort"]]))
            rad = np.concatenate([rad, e["radius"]])
            vol = np.concatenate([vol, e["volume"]])
        S.rk4(dt, 1)
        x = orc.rk4(x, rad, vol, nu, dt, 1, mode=orc.OMP_CLEAN, uinf=uinf)
    g = S.getStates()
    n = rad.size
    print(f"case emission precision={precision}: N = {n} after {steps} steps", flush=True)
    assert S._size == 6 * n
    ep = per_particle_rel(g[:3 * n], x[:3 * n])
    ea = per_particle_rel(g[3 * n:], x[3 * n:])
    print(f"  pos {ep:.3e} vort {ea:.3e}", flush=True)
    assert ep < tol_pos and ea < tol_vort
    # growth past the first allocation (1,280 > 256-particle start) kept the state
    S.close()


def test_emission_into_first_of_two_wakes():
    a = gen.vortex_ring(16, 16, seed=31)
    b = gen.vortex_ring(8, 9, seed=32, center=(0, 0, 0.5))
    e = gen.vortex_ring(4, 16, seed=33, center=(0, 0, -0.3))
    S = pa.Interaction(pa.Wake.from_dict(a), pa.Wake.from_dict(b), nu=1e-3)
    S.rk4(0.005, 1)
    S.appendParticles(0, e["pos"], e["vort"], e["radius"], e["volume"])
    S.solve()
    r = S.getRates()
    st = S.getStates()
    S.close()
    n1, n2 = a["n"] + e["n"], b["n"]
    p, v = orc.unpack_state(st, [n1, n2])
    rad = np.concatenate([a["radius"], e["radius"], b["radius"]])
    vol = np.concatenate([a["volume"], e["volume"], b["volume"]])
    vel, dv = orc.rates(p, v, rad, vol, 1e-3, orc.SERIAL, wake_n=[n1, n2])
    ref = orc.pack_state(vel, dv, [n1, n2])
    assert np.max(np.abs(r - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_add_wake_after_emission_into_first_wake():
    """A wake added while emission into the first wake is still only on the
    host (no solve in between) keeps the emitted particles and the stepped
    state of the others."""
    a = gen.vortex_ring(16, 16, seed=41)
    b = gen.vortex_ring(8, 9, seed=42, center=(0, 0, 0.5))
    e = gen.vortex_ring(4, 16, seed=43, center=(0, 0, -0.3))
    c = gen.vortex_ring(8, 8, seed=44, center=(0, 0, 1.0))
    S = pa.Interaction(pa.Wake.from_dict(a), pa.Wake.from_dict(b), nu=1e-3)
    S.rk4(0.005, 1)
    stepped = S.getStates()
    S.appendParticles(0, e["pos"], e["vort"], e["radius"], e["volume"])
    S.addWake(pa.Wake.from_dict(c))
    st = S.getStates()
    S.close()
    pa0, va0 = orc.unpack_state(stepped, [a["n"], b["n"]])
    want_p = np.concatenate([pa0[: a["n"]], e["pos"], pa0[a["n"]:], c["pos"]])
    want_v = np.concatenate([va0[: a["n"]], e["vort"], va0[a["n"]:], c["vort"]])
    ns = [a["n"] + e["n"], b["n"], c["n"]]
    assert np.array_equal(st, orc.pack_state(want_p, want_v, ns))


@pytest.mark.parametrize("precision", [pa.FP64, pa.FP32C])
def test_async_frames_equal_sync_frames(tmp_path, precision):
    """Frames queued with write_async (snapshot on the device, copy stream,
    writer thread) are byte-identical to synchronous write() frames of the
    same run, including frames after emission into the first wake (a
    rebuild of the device arrays while frames are queued) and into the last."""
    a, b, _ = _two_wakes()
    e = gen.vortex_ring(8, 16, seed=34, center=(0.0, 0.0, -0.4))
    print(f"case async frames precision={precision}", flush=True)
    S1 = pa.Interaction(pa.Wake.from_dict(a), pa.Wake.from_dict(b), nu=1e-3, precision=precision)
    S2 = pa.Interaction(pa.Wake.from_dict(a), pa.Wake.from_dict(b), nu=1e-3, precision=precision)
    ps, pa_ = tmp_path / "sync.wake", tmp_path / "async.wake"
    S1.write(ps, t=0.0, append=False)
    with pa.WakeFile(pa_) as f:
        S2.write_async(f, 0.0)
        for k in range(1, 7):
            if k == 2:  # into the first wake: the device arrays are rebuilt
                for S in (S1, S2):
                    S.appendParticles(0, e["pos"], e["vort"], e["radius"], e["volume"])
            if k == 4:  # into the last wake: appended in place
                for S in (S1, S2):
                    S.appendParticles(1, e["pos"], e["vort"], e["radius"], e["volume"])
            S1.rk4(0.01, 1)
            S2.rk4(0.01, 1)
            S1.write(ps, t=0.01 * k)
            S2.write_async(f, 0.01 * k)
        S2.write_async(f)  # a frame without the time, as write() without t
    S1.write(ps)
    st1, st2 = S1.getStates(), S2.getStates()
    S1.close()
    S2.close()
    assert np.array_equal(st1, st2)
    raw1, raw2 = ps.read_bytes(), pa_.read_bytes()
    print(f"  {len(raw1)} bytes sync, {len(raw2)} async", flush=True)
    assert raw1 == raw2


def test_integrate_to_file_async(tmp_path):
    """Integration.integrate(S, path=...) (asynchronous frames) writes what
    integrate() with a synchronous frame writer writes."""
    w = gen.vortex_ring(64, 16, seed=35)
    p1, p2 = tmp_path / "a.wake", tmp_path / "b.wake"
    S1 = pa.Interaction(pa.Wake.from_dict(w), nu=1e-3)
    S2 = pa.Interaction(pa.Wake.from_dict(w), nu=1e-3)
    first = [True]

    def sync_writer(t, S):
        S.write(p1, t=t, append=not first[0])
        first[0] = False

    pa.RK4(0.05, 5).integrate(S1, writer=sync_writer)
    pa.RK4(0.05, 5).integrate(S2, path=p2)
    x1, x2 = S1.getStates(), S2.getStates()
    S1.close()
    S2.close()
    assert np.array_equal(x1, x2)
    assert p1.read_bytes() == p2.read_bytes()
    assert len(p2.read_bytes()) == 6 * (8 + 8 + 8 + 56 * w["n"])


def test_async_frame_write_failure_reported(tmp_path):
    """A frame whose file write fails (a stream opened read-only) is reported
    by the following wait, once; the context keeps working."""
    import ctypes as C
    from pawan_amd import _lib
    from pawan_amd.interaction import lib
    w = gen.vortex_ring(16, 9, seed=36)
    S = pa.Interaction(pa.Wake.from_dict(w), nu=1e-3)
    p = tmp_path / "ro.wake"
    p.write_bytes(b"")
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    f = libc.fopen(str(p).encode(), b"rb")
    assert f
    try:
        assert lib().pw_write_async(S._h, f, 0.0, 1) == _lib.PW_OK
        assert lib().pw_write_wait(S._h) == _lib.PW_EINVAL
        assert b"write" in lib().pw_last_error(S._h)
        assert lib().pw_write_wait(S._h) == _lib.PW_OK
    finally:
        libc.fclose(f)
    good = tmp_path / "ok.wake"
    with pa.WakeFile(good) as g:
        S.write_async(g, 0.0)
    S.close()
    assert len(good.read_bytes()) == 8 + 8 + 8 + 56 * w["n"]


# ------------------------------------------ binary32 variant past 524,288
def _fp32c_rates(w, **kw):
    S = pa.Interaction(pa.Wake.from_dict(w), nu=w["nu"], uinf=w["uinf"], precision=pa.FP32C, **kw)
    S.solve()
    r, v = S.getRates(), S.variant()
    S.close()
    return r, v


def test_fp32c_large_context_explicit_tpt_and_chunk():
    """Past 524,288 particles a binary32 symmetric context defaults to 512 x 1;
    an explicit two targets per thread keeps 256 threads (256 x 2), and so
    does a chunk length that 512 does not divide (here 129 x 256); each
    solves, and the three agree to binary32 rounding."""
    w = gen.config_rotor_wake(2112)  # 540,672 particles
    print(f"case fp32c N={w['n']}: default / targets_per_thread 2 / chunk 33,024", flush=True)
    r0, v0 = _fp32c_rates(w)
    assert (v0["block"], v0["targets_per_thread"], v0["scheme"]) == (512, 1, pa.SCHEME_SYMMETRIC)
    r2, v2 = _fp32c_rates(w, targets_per_thread=2)
    assert (v2["block"], v2["targets_per_thread"]) == (256, 2), v2
    rc, vc = _fp32c_rates(w, chunk=129 * 256)
    assert vc["block"] == 256, vc
    for r in (r2, rc):
        e = rel_rms(r, r0)
        print(f"  vs default: {e:.3e}", flush=True)
        assert e < 1e-6


def test_fp32c_append_across_524288_rebuilds_at_512():
    """Particle emission that takes a binary32 context past 524,288 particles
    moves it to the 512 x 1 pass (half the source-side slots) instead of
    keeping the 256 x 1 pass chosen at its first build; appends below the
    threshold stay in place.  The grown context's state is the particles
    given, and its rates equal a context built at the final size, bit for bit."""
    full = gen.config_rotor_wake(2050)  # 524,800 particles
    n0, n1 = 522240, 522752
    part = {k: full[k][:n0] for k in ("pos", "vort", "radius", "volume")}
    print(f"case fp32c append {n0} -> {n1} -> {full['n']}", flush=True)
    S = pa.Interaction(pa.Wake.from_dict(part), nu=full["nu"], uinf=full["uinf"],
                       precision=pa.FP32C)
    S.solve()
    assert S.variant()["block"] == 256
    for a, b in ((n0, n1), (n1, full["n"])):
        S.appendParticles(0, full["pos"][a:b], full["vort"][a:b], full["radius"][a:b],
                          full["volume"][a:b])
        v = S.variant()
        print(f"  after append to {b}: {v}", flush=True)
        assert v["block"] == (256 if b <= 524288 else 512)
    x = S.getStates()
    n = full["n"]
    assert np.array_equal(x[:3 * n], full["pos"].ravel())
    assert np.array_equal(x[3 * n:], full["vort"].ravel())
    S.solve()
    r = S.getRates()
    S.close()
    rf, vf = _fp32c_rates(full)
    assert vf["block"] == 512
    assert np.array_equal(r, rf)
