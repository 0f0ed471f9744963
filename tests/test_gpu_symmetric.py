"""The symmetric pair passes (each unordered pair once, both particles
updated, source-side sums in per-block slots), binary32 and binary64, against
the CPU oracle, on the shapes where their bookkeeping can go wrong: particle
counts around the 64-wide tiles and the target blocks, a block straddling a
chunk boundary, several wakes, non-uniform cores, the Gaussian kernel, a
freestream, far tiles (pairs wholly in binary32) and a set far from the
origin (where none may be far); and against the one-sided passes.  Tolerances: binary32 1e-6
relative RMS per rate field (velocity, vorticity rate); binary64 1e-10 per
particle, as for every result of that precision."""
import numpy as np
import pytest

import oracle as orc
import pawan_amd as pa
from pawan_amd import generators as gen
from helpers import per_particle_rel, rel_rms, sparse_lattice

pytestmark = pytest.mark.gpu

TOL32 = 1e-6


def _random_blob(n, seed, core=0.05, spread=0.5):
    u = gen.uniform(seed, 7 * n).reshape(n, 7)
    pos = spread * (u[:, :3] - 0.5)
    vort = 0.01 * (u[:, 3:6] - 0.5)
    return {"pos": np.ascontiguousarray(pos), "vort": np.ascontiguousarray(vort),
            "radius": np.full(n, core), "volume": np.full(n, 4.0 / 3.0 * np.pi * core ** 3),
            "n": n}


def _rates(w, nu, precision, scheme=pa.SCHEME_DEFAULT, wake_n=None, **kw):
    wake_n = wake_n or [w["pos"].shape[0]]
    offs = np.cumsum([0] + list(wake_n))
    wakes = [pa.Wake(w["pos"][a:b], w["vort"][a:b], w["radius"][a:b], w["volume"][a:b])
             for a, b in zip(offs[:-1], offs[1:])]
    S = pa.Interaction(*wakes, nu=nu, precision=precision, scheme=scheme, **kw)
    S.solve()
    r = S.getRates()
    S.close()
    return orc.unpack_state(r, wake_n)


@pytest.mark.parametrize("n", [2, 3, 63, 64, 65, 127, 200, 511, 513, 1000])
@pytest.mark.parametrize("block,tpt", [(64, 1), (128, 2), (256, 1), (256, 2), (512, 1)])
def test_symmetric_tile_edges(n, block, tpt):
    """Random blobs: every pair counted once.  Random strengths make the sums
    cancel far more than any vortex workload does, so the binary32 rounding
    shows at ~1e-6 here; the bound is 1e-5, far below what one missing or
    doubled pair would cost (~1/n of the RMS rate)."""
    w = _random_blob(n, seed=1000 + n)
    print(f"case symmetric n={n} block={block} tpt={tpt}", flush=True)
    nu = 1e-3
    vel, dv = _rates(w, nu, pa.FP32C, block=block, targets_per_thread=tpt)
    ob = min(block, 256)  # the one-sided pass has no 512-thread variant
    vo, do = _rates(w, nu, pa.FP32C, pa.SCHEME_ONESIDED, block=ob, targets_per_thread=tpt)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL)
    ev, ed = rel_rms(vel, rv), rel_rms(dv, rd)
    print(f"  symmetric vel {ev:.2e} dvort {ed:.2e}; one-sided vel {rel_rms(vo, rv):.2e} "
          f"dvort {rel_rms(do, rd):.2e}", flush=True)
    assert ev < 1e-5 and ed < 1e-5, (ev, ed)


def test_symmetric_single_particle():
    w = _random_blob(1, seed=7)
    vel, dv = _rates(w, 1e-3, pa.FP32C, uinf=(0.25, 0.0, -0.5))
    assert np.array_equal(vel, np.array([[0.25, 0.0, -0.5]])) and np.all(dv == 0.0)


@pytest.mark.parametrize("block,tpt,chunk", [(128, 2, 384), (64, 2, 192), (256, 2, 768)])
def test_symmetric_block_straddles_chunks(block, tpt, chunk):
    """Chunks that are not multiples of the target block: a block's own
    sources start in the middle of a chunk and run into the next one."""
    w = gen.vortex_ring(60, 30, R=1.0, core=0.1, seed=21)  # 1,800 particles
    print(f"case straddle block={block} tpt={tpt} chunk={chunk}", flush=True)
    nu = 1e-3
    vel, dv = _rates(w, nu, pa.FP32C, block=block, targets_per_thread=tpt, chunk=chunk)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.OMP_CLEAN)
    ev, ed = rel_rms(vel, rv), rel_rms(dv, rd)
    print(f"  vel {ev:.2e} dvort {ed:.2e}", flush=True)
    assert ev < TOL32 and ed < TOL32


def test_symmetric_wakes_nonuniform_gaussian_freestream():
    a = gen.vortex_ring(40, 16, R=1.0, core=0.1, seed=31)
    b = gen.vortex_ring(30, 16, R=0.7, core=0.08, seed=32, center=(0.1, 0, 0.3))
    w = {k: np.concatenate([a[k], b[k]]) for k in ("pos", "vort", "radius", "volume")}
    w["radius"] = w["radius"] * (1.0 + 0.3 * gen.uniform(33, w["pos"].shape[0]))
    wake_n = [a["n"], b["n"]]
    nu, uinf = 2e-3, (0.0, 0.05, -0.1)
    for kernel, kind in ((pa.KERNEL_HIGHORDER, orc.HIGHORDER), (pa.KERNEL_GAUSSIAN, orc.GAUSSIAN)):
        print(f"case wakes nonuniform kernel={kernel}", flush=True)
        vel, dv = _rates(w, nu, pa.FP32C, wake_n=wake_n, kernel=kernel, uinf=uinf)
        rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL,
                           kind=kind, wake_n=wake_n, uinf=uinf)
        ev, ed = rel_rms(vel, rv), rel_rms(dv, rd)
        print(f"  vel {ev:.2e} dvort {ed:.2e}", flush=True)
        assert ev < TOL32 and ed < TOL32


def test_symmetric_vs_onesided_rotor():
    w = gen.rotor_wake(32)  # 8,192 particles
    print(f"case rotor N={w['n']}", flush=True)
    nu = 1e-4
    v64, d64 = _rates(w, nu, pa.FP64)
    vs, ds = _rates(w, nu, pa.FP32C, pa.SCHEME_SYMMETRIC)
    vo, do = _rates(w, nu, pa.FP32C, pa.SCHEME_ONESIDED)
    for name, v, d in (("symmetric", vs, ds), ("one-sided", vo, do)):
        ev, ed = rel_rms(v, v64), rel_rms(d, d64)
        print(f"  {name}: vel {ev:.2e} dvort {ed:.2e}", flush=True)
        assert ev < TOL32 and ed < TOL32


@pytest.mark.parametrize("n", [2, 63, 64, 65, 200, 513, 1000])
@pytest.mark.parametrize("block,tpt", [(64, 1), (128, 2), (256, 1)])
def test_symmetric64_tile_edges(n, block, tpt):
    w = _random_blob(n, seed=2000 + n)
    print(f"case symmetric64 n={n} block={block} tpt={tpt}", flush=True)
    nu = 1e-3
    vel, dv = _rates(w, nu, pa.FP64, pa.SCHEME_SYMMETRIC, block=block, targets_per_thread=tpt)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL)
    ev, ed = per_particle_rel(vel, rv), per_particle_rel(dv, rd)
    assert ev < 1e-10 and ed < 1e-10, (ev, ed)


def test_symmetric64_wakes_nonuniform_gaussian_straddle():
    a = gen.vortex_ring(40, 16, R=1.0, core=0.1, seed=41)
    b = gen.vortex_ring(30, 16, R=0.7, core=0.08, seed=42, center=(0.1, 0, 0.3))
    w = {k: np.concatenate([a[k], b[k]]) for k in ("pos", "vort", "radius", "volume")}
    w["radius"] = w["radius"] * (1.0 + 0.3 * gen.uniform(43, w["pos"].shape[0]))
    wake_n = [a["n"], b["n"]]
    nu, uinf = 2e-3, (0.0, 0.05, -0.1)
    for kernel, kind in ((pa.KERNEL_HIGHORDER, orc.HIGHORDER), (pa.KERNEL_GAUSSIAN, orc.GAUSSIAN)):
        print(f"case symmetric64 wakes nonuniform kernel={kernel}", flush=True)
        vel, dv = _rates(w, nu, pa.FP64, pa.SCHEME_SYMMETRIC, wake_n=wake_n, kernel=kernel,
                         uinf=uinf, block=128, targets_per_thread=1, chunk=384)
        rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL,
                           kind=kind, wake_n=wake_n, uinf=uinf)
        ev, ed = per_particle_rel(vel, rv), per_particle_rel(dv, rd)
        print(f"  vel {ev:.2e} dvort {ed:.2e}", flush=True)
        assert ev < 1e-10 and ed < 1e-10


def test_block512_batch_alone_equals_mixed():
    """A batch built with 512 threads runs its binary32 jobs on the 512 x 1
    symmetric pass (half the source-side slot rows) and its binary64 jobs at
    256: every job bitwise the same alone and in the mix, the binary32 job
    within 1e-6 of the oracle after the steps."""
    jobs = gen.batch_jobs(3, seed=11, n_min=600, n_max=2500)
    precs = [pa.FP32C, pa.FP64, pa.FP32C]
    print("case batch block 512: n = " + ", ".join(str(j["n"]) for j in jobs), flush=True)
    res = []
    for members in ([0], [0, 1, 2]):
        B = pa.Batch(block=512, targets_per_thread=1)
        ids = [B.add(jobs[k], jobs[k]["nu"], jobs[k]["dt"], precs[k]) for k in members]
        B.step(2, graph=True)
        res.append(B.get(ids[0]))
        B.close()
    assert np.array_equal(res[0]["pos"], res[1]["pos"])
    assert np.array_equal(res[0]["vort"], res[1]["vort"])
    j = jobs[0]
    x = orc.rk4(orc.pack_state(j["pos"], j["vort"]), j["radius"], j["volume"], j["nu"], j["dt"], 2,
                mode=orc.OMP_CLEAN)
    p, v = orc.unpack_state(x)
    ep, ev = rel_rms(res[0]["pos"], p), rel_rms(res[0]["vort"], v)
    print(f"  binary32 512 x 1 after 2 steps vs oracle: pos {ep:.2e} vort {ev:.2e}", flush=True)
    assert ep < TOL32 and ev < TOL32


def test_symmetric64_rk4_batch_alone_equals_mixed():
    """RK4 through the symmetric binary64 pass: one job alone and in a batch
    with others gives bitwise the same states, and matches the oracle."""
    jobs = gen.batch_jobs(3, seed=7, n_min=500, n_max=3000)
    res = []
    for members in ([0], [0, 1, 2]):
        B = pa.Batch(block=128, targets_per_thread=2, scheme=pa.SCHEME_SYMMETRIC)
        ids = [B.add(jobs[k], jobs[k]["nu"], jobs[k]["dt"], pa.FP64) for k in members]
        B.step(3, graph=True)
        res.append(B.get(ids[0]))
        B.close()
    assert np.array_equal(res[0]["pos"], res[1]["pos"])
    assert np.array_equal(res[0]["vort"], res[1]["vort"])
    j = jobs[0]
    x = orc.rk4(orc.pack_state(j["pos"], j["vort"]), j["radius"], j["volume"], j["nu"], j["dt"], 3,
                mode=orc.OMP_CLEAN)
    p, v = orc.unpack_state(x)
    ep, ev = per_particle_rel(res[0]["pos"], p), per_particle_rel(res[0]["vort"], v)
    print(f"  symmetric64 rk4 vs oracle: pos {ep:.2e} vort {ev:.2e}", flush=True)
    assert ep < 1e-11 and ev < 1e-9


@pytest.mark.parametrize("block", [256, 512])
def test_far_tiles_sparse_lattice(block):
    """A sparse lattice ordered by 64-particle sub-cubes: most sub-cubes are
    farther from a target block than 1/8 of the coordinates, so their tiles
    run wholly in binary32, and they carry most of every velocity; the rates
    still meet the binary32 bar against the oracle."""
    w = sparse_lattice()
    nu = 1e-3
    print(f"case far tiles: sparse lattice N={w['n']} block={block}", flush=True)
    vel, dv = _rates(w, nu, pa.FP32C, block=block, targets_per_thread=1)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL)
    # the share of each velocity that comes from other sub-cubes
    own = np.zeros_like(rv)
    for c in range(w["n"] // 64):
        s = slice(64 * c, 64 * c + 64)
        own[s] = orc.rates(w["pos"][s], w["vort"][s], w["radius"][s], w["volume"][s], nu,
                           mode=orc.SERIAL)[0]
    share = np.linalg.norm(rv - own) / np.linalg.norm(rv)
    ev, ed = rel_rms(vel, rv), rel_rms(dv, rd)
    print(f"  velocity {ev:.2e} vorticity rate {ed:.2e}; other sub-cubes' share of the "
          f"velocity {share:.2f}", flush=True)
    assert share > 0.5
    assert ev < TOL32 and ed < TOL32, (ev, ed)


def test_far_test_scales_with_coordinates():
    """A ring 1,000 radii from the origin: binary32 positions there are ~6e-5
    apart, a thousandth of the particle spacing, so no tile of the ring may
    count as far (its size is far below 1/8 of its coordinates); the binary32
    rates match the oracle as they do at the origin."""
    w = gen.vortex_ring(128, 16, R=1.0, core=0.1, seed=3, center=(1000.0, 0.0, 0.0))
    nu = 1e-3
    print(f"case far test: ring at x = 1000, N={w['n']}", flush=True)
    vel, dv = _rates(w, nu, pa.FP32C)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL)
    ev, ed = rel_rms(vel, rv), rel_rms(dv, rd)
    print(f"  velocity {ev:.2e} vorticity rate {ed:.2e}", flush=True)
    assert ev < TOL32 and ed < TOL32, (ev, ed)
