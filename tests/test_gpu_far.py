"""The far-field form of the pair kernel against the CPU oracle.

Pairs with t = rho^2/sigma^2 >= 1024 (binary64: decided per pair of 64-
particle groups by their bounding boxes; in the symmetric option per target
block and 64-source tile) and far tiles with t >= 128 (the binary32 symmetric
pass) take the kernel values from one reciprocal square
root and a short polynomial, with Zs and the PSE terms dropped (below binary64
rounding there).  The oracle evaluates the reference's full kernel for every
pair (highorder.h:20-60, gaussian.h:20-58, interaction_utils.h:429-481).
Checked here: a particle set in which most of every velocity comes from far
pairs (both kernels, both precisions, both schemes of each), pairs of clusters whose boxes sit just
below, at and above the binary64 threshold, and that the far choice leaves the
one-sided variants bitwise equal.  Bars: binary64 1e-10 per particle (the
parity bar; the far form is measured at rounding level, printed); binary32
1e-6 relative RMS."""
import numpy as np
import pytest

import oracle as orc
import pawan_amd as pa
from pawan_amd import generators as gen
from helpers import per_particle_rel, rel_rms, sparse_lattice

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 1e-6
KIND = {pa.KERNEL_HIGHORDER: orc.HIGHORDER, pa.KERNEL_GAUSSIAN: orc.GAUSSIAN}


def _rates(w, nu, precision, kernel=pa.KERNEL_HIGHORDER, **kw):
    S = pa.Interaction(pa.Wake.from_dict(w), nu=nu, precision=precision, kernel=kernel, **kw)
    S.solve()
    r = S.getRates()
    S.close()
    n = w["n"]
    return r[:3 * n].reshape(n, 3), r[3 * n:].reshape(n, 3)


@pytest.mark.parametrize("kernel", [pa.KERNEL_HIGHORDER, pa.KERNEL_GAUSSIAN])
def test_far_sparse_lattice_both_precisions(kernel):
    """4,096 particles on a sparse lattice (sub-cubes of 64, 40 core radii
    between neighbouring sub-cubes' boxes): every pair between two sub-cubes
    is far, and those pairs carry most of every velocity."""
    w = sparse_lattice()
    nu = 1e-3
    print(f"case far lattice kernel={kernel} N={w['n']}", flush=True)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL,
                       kind=KIND[kernel])
    own = np.zeros_like(rv)
    for c in range(w["n"] // 64):
        s = slice(64 * c, 64 * c + 64)
        own[s] = orc.rates(w["pos"][s], w["vort"][s], w["radius"][s], w["volume"][s], nu,
                           mode=orc.SERIAL, kind=KIND[kernel])[0]
    share = np.linalg.norm(rv - own) / np.linalg.norm(rv)
    v64, d64 = _rates(w, nu, pa.FP64, kernel)
    e64v, e64d = per_particle_rel(v64, rv), per_particle_rel(d64, rd)
    vs, ds = _rates(w, nu, pa.FP64, kernel, scheme=pa.SCHEME_SYMMETRIC)
    esv, esd = per_particle_rel(vs, rv), per_particle_rel(ds, rd)
    v32, d32 = _rates(w, nu, pa.FP32C, kernel)
    e32v, e32d = rel_rms(v32, rv), rel_rms(d32, rd)
    vo, do = _rates(w, nu, pa.FP32C, kernel, scheme=pa.SCHEME_ONESIDED)
    eov, eod = rel_rms(vo, rv), rel_rms(do, rd)
    print(f"  far pairs' share of the velocity {share:.2f}; binary64 per particle: velocity "
          f"{e64v:.2e} vorticity rate {e64d:.2e} (symmetric {esv:.2e} {esd:.2e}); binary32 "
          f"rel RMS: {e32v:.2e} {e32d:.2e} (one-sided {eov:.2e} {eod:.2e})", flush=True)
    assert share > 0.5
    assert e64v < TOL64 and e64d < TOL64, (e64v, e64d)
    assert esv < TOL64 and esd < TOL64, (esv, esd)
    assert e32v < TOL32 and e32d < TOL32, (e32v, e32d)
    assert eov < TOL32 and eod < TOL32, (eov, eod)


def _two_clusters(gap, core=0.05, seed=11):
    """Two 64-particle clusters of side 0.5 sigma whose bounding boxes are
    `gap` core radii apart along x (sigma = core for a uniform core).  The
    second carries no vorticity: its velocities come from the first cluster
    alone, i.e. from the pairs between the two clusters."""
    u = gen.uniform(seed, 12 * 64).reshape(128, 6)
    side = 0.5 * core
    a = side * u[:64, :3]
    b = side * u[64:, :3]
    # b's box starts `gap` sigma past a's box along x
    b[:, 0] += a[:, 0].max() - b[:, 0].min() + gap * core
    pos = np.concatenate([a, b])
    vort = 0.01 * (u[:, 3:6] - 0.5)
    vort[64:] = 0.0
    n = 128
    return {"pos": np.ascontiguousarray(pos), "vort": np.ascontiguousarray(vort),
            "radius": np.full(n, core), "volume": np.full(n, 4.0 / 3.0 * np.pi * core ** 3),
            "n": n}


@pytest.mark.parametrize("gap", [20.0, 31.9, 32.0, 32.1, 45.0, 1000.0])
def test_far_threshold_two_clusters(gap):
    """Clusters around the binary64 threshold (box distance 32 sigma, t = 1024)
    and past the binary32 one (11.3 sigma): the near and far forms agree with
    the oracle on either side of it."""
    w = _two_clusters(gap)
    nu = 1e-3
    print(f"case far threshold gap={gap} sigma", flush=True)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL)
    v64, d64 = _rates(w, nu, pa.FP64)
    v32, d32 = _rates(w, nu, pa.FP32C)
    vs, ds = _rates(w, nu, pa.FP64, scheme=pa.SCHEME_SYMMETRIC)
    esv, esd = per_particle_rel(vs, rv), per_particle_rel(ds, rd)
    # the second cluster's velocities: all from pairs between the clusters
    eb = per_particle_rel(v64[64:], rv[64:])
    eb32 = rel_rms(v32[64:], rv[64:])
    e64v, e64d = per_particle_rel(v64, rv), per_particle_rel(d64, rd)
    e32v, e32d = rel_rms(v32, rv), rel_rms(d32, rd)
    print(f"  binary64 per particle: velocity {e64v:.2e} (second cluster {eb:.2e}) "
          f"vorticity rate {e64d:.2e}; binary32 rel RMS {e32v:.2e} (second cluster "
          f"{eb32:.2e}) {e32d:.2e}; binary64 symmetric {esv:.2e} {esd:.2e}", flush=True)
    assert e64v < TOL64 and e64d < TOL64 and eb < TOL64
    assert esv < TOL64 and esd < TOL64
    assert e32v < TOL32 and e32d < TOL32 and eb32 < TOL32


def test_far_choice_far_from_origin():
    """Two touching clusters (boxes 1 sigma apart, sigma = 1/16 so that the
    records' coordinates in units of sigma are exact) placed about 2^30 sigma
    from the origin, with the binary32 rounding boundary between the two
    boxes: rounded to nearest, the boxes would read 128 sigma apart.  Widened
    boxes keep these pairs in the full form, and every path matches the
    oracle as near the origin."""
    core = 1.0 / 16.0
    w = _two_clusters(1.0, core=core)
    w["pos"][:, 0] += (2.0 ** 30 + 62.75) * core
    nu = 1e-3
    lo_b = w["pos"][64:, 0].min() / core
    print(f"case far choice far from the origin: second box from {lo_b:.4f} sigma", flush=True)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL)
    out = []
    for prec, sch in ((pa.FP64, pa.SCHEME_ONESIDED), (pa.FP64, pa.SCHEME_SYMMETRIC),
                      (pa.FP32C, pa.SCHEME_ONESIDED), (pa.FP32C, pa.SCHEME_SYMMETRIC)):
        v, d = _rates(w, nu, prec, scheme=sch)
        if prec == pa.FP64:
            e = (per_particle_rel(v, rv), per_particle_rel(d, rd))
            assert e[0] < TOL64 and e[1] < TOL64, (prec, sch, e)
        else:
            e = (rel_rms(v, rv), rel_rms(d, rd))
            assert e[0] < TOL32 and e[1] < TOL32, (prec, sch, e)
        out.append(f"{prec}/{sch}: {e[0]:.2e} {e[1]:.2e}")
    print("  " + "; ".join(out), flush=True)


@pytest.mark.parametrize("cut", [27, 63])
def test_far_ragged_end(cut):
    """The sparse lattice with its last sub-cube cut short (N = 4,096 - cut):
    the job's last 64-particle group is partial (always the full form; its
    targets past N repeat particle N - 1 in the box), every other pair of
    sub-cubes is far.  Both precisions, both schemes, against the oracle, and
    the one-sided binary64 variants bitwise equal."""
    w = sparse_lattice()
    n = w["n"] - cut
    w = {k: (v[:n] if isinstance(v, np.ndarray) else v) for k, v in w.items()}
    w["n"] = n
    nu = 1e-3
    print(f"case far ragged end N={n}", flush=True)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL)
    out = []
    for prec, sch in ((pa.FP64, pa.SCHEME_ONESIDED), (pa.FP64, pa.SCHEME_SYMMETRIC),
                      (pa.FP32C, pa.SCHEME_ONESIDED), (pa.FP32C, pa.SCHEME_SYMMETRIC)):
        v, d = _rates(w, nu, prec, scheme=sch)
        if prec == pa.FP64:
            e = (per_particle_rel(v, rv), per_particle_rel(d, rd))
            assert e[0] < TOL64 and e[1] < TOL64, (prec, sch, e)
        else:
            e = (rel_rms(v, rv), rel_rms(d, rd))
            assert e[0] < TOL32 and e[1] < TOL32, (prec, sch, e)
        out.append(f"{prec}/{sch}: {e[0]:.2e} {e[1]:.2e}")
    print("  " + "; ".join(out), flush=True)
    ref = None
    for block, tpt in ((64, 1), (128, 2), (256, 1), (256, 2)):
        S = pa.Interaction(pa.Wake.from_dict(w), nu=nu, precision=pa.FP64, block=block,
                           chunk=1024, targets_per_thread=tpt, unroll=2,
                           scheme=pa.SCHEME_ONESIDED)
        S.solve()
        r = S.getRates()
        S.close()
        if ref is None:
            ref = r
        assert np.array_equal(r, ref), (block, tpt)


def test_far_onesided_variants_bitwise_identical():
    """The far choice is made per pair of 64-particle groups of the job, so
    every one-sided binary64 variant (block, targets per thread, unroll) gives
    the same bits on a set where most pairs are far."""
    w = sparse_lattice()
    print(f"case far variants N={w['n']}", flush=True)
    ref = None
    for block in (64, 128, 256):
        for tpt in (1, 2):
            for unroll in (1, 2, 4):
                S = pa.Interaction(pa.Wake.from_dict(w), nu=1e-3, precision=pa.FP64, block=block,
                                   chunk=1024, targets_per_thread=tpt, unroll=unroll,
                                   scheme=pa.SCHEME_ONESIDED)
                S.solve()
                r = S.getRates()
                S.close()
                if ref is None:
                    ref = r
                assert np.array_equal(r, ref), (block, tpt, unroll)


@pytest.mark.parametrize("kernel", [pa.KERNEL_HIGHORDER, pa.KERNEL_GAUSSIAN])
@pytest.mark.parametrize("offset", [0.0, 2.0 ** 20])
@pytest.mark.parametrize("length", [120.0, 136.0])
def test_far_wide_source_group(length, offset, kernel):
    """A compact 64-particle cluster and a 64-particle line `length` core
    radii long, 40 core radii beyond the cluster's box: the binary64 one-sided
    pass adds a far group by moments only when it is at most 128 core radii
    wide (the line at 120: moments, at 136: the full form; the cluster as a
    source group: moments either way), with every regrouped term at most
    1 + 2 x 128 / 32 times the per-pair term it stands for.  The whole set is
    also placed 2^20 core radii from the origin, where the moments about each
    group's first particle keep the same bound (sigma = 1/16, so that the
    records' coordinates in units of sigma are exact there, as the oracle's
    are).  Both kernels, against the oracle."""
    core = 1.0 / 16.0
    u = gen.uniform(23, 12 * 64).reshape(128, 6)
    a = 0.5 * core * u[:64, :3]
    b = np.zeros((64, 3))
    b[:, 1] = np.linspace(0.0, length * core, 64)
    b[:, 2] = 0.25 * core * u[64:, 2]
    b[:, 0] = a[:, 0].max() + 40.0 * core + 0.25 * core * u[64:, 0]
    pos = np.concatenate([a, b]) + offset * core
    vort = 0.01 * (u[:, 3:6] - 0.5)
    n = 128
    w = {"pos": np.ascontiguousarray(pos), "vort": np.ascontiguousarray(vort),
         "radius": np.full(n, core), "volume": np.full(n, 4.0 / 3.0 * np.pi * core ** 3), "n": n}
    nu = 1e-3
    print(f"case far wide source group length={length} offset={offset} kernel={kernel}",
          flush=True)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL,
                       kind=KIND[kernel])
    v64, d64 = _rates(w, nu, pa.FP64, kernel)
    e64v, e64d = per_particle_rel(v64, rv), per_particle_rel(d64, rd)
    # the line's velocities come from its own pairs and from the cluster
    # (a far source group); the cluster's from itself and from the line
    print(f"  binary64 per particle: velocity {e64v:.2e} vorticity rate {e64d:.2e}", flush=True)
    assert e64v < TOL64 and e64d < TOL64, (e64v, e64d)


def test_far_lattice_away_from_origin():
    """The sparse lattice moved 2^20 core radii from the origin (sigma = 1/4,
    so that the records' coordinates in units of sigma are exact).  The
    binary32 pass measures a far-field tile's coordinates from the centre of
    the target block's box, so these tiles still run wholly in binary32; both
    binary32 schemes and the binary64 pass against the oracle, as at the
    origin."""
    core = 0.25
    w = sparse_lattice(core=core)
    w["pos"] = w["pos"] + 2.0 ** 20 * core
    nu = 1e-3
    print(f"case far lattice at 2^20 sigma N={w['n']}", flush=True)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL)
    v64, d64 = _rates(w, nu, pa.FP64)
    e64v, e64d = per_particle_rel(v64, rv), per_particle_rel(d64, rd)
    out = [f"binary64 {e64v:.2e} {e64d:.2e}"]
    assert e64v < TOL64 and e64d < TOL64, (e64v, e64d)
    for sch in (pa.SCHEME_SYMMETRIC, pa.SCHEME_ONESIDED):
        v, d = _rates(w, nu, pa.FP32C, scheme=sch)
        e = (rel_rms(v, rv), rel_rms(d, rd))
        out.append(f"binary32 scheme {sch}: {e[0]:.2e} {e[1]:.2e}")
        assert e[0] < TOL32 and e[1] < TOL32, (sch, e)
    print("  " + "; ".join(out), flush=True)


@pytest.mark.parametrize("kernel", [pa.KERNEL_HIGHORDER, pa.KERNEL_GAUSSIAN])
@pytest.mark.parametrize("spacing", [2.25, 3.0])
def test_far_lattice_close_tiles(spacing, kernel):
    """The sparse lattice with neighbouring sub-cubes' boxes 9.1 and 12.2 core
    radii apart: far binary32 tiles whose pairs all have 64 <= t < 128 (the
    far-field form with the u^4 term) or t >= 128, beside near ones, against
    the oracle in both binary32 schemes and binary64, for both kernels."""
    w = sparse_lattice(spacing=spacing)
    nu = 1e-3
    print(f"case far lattice spacing {spacing} kernel={kernel} N={w['n']}", flush=True)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], nu, mode=orc.SERIAL,
                       kind=KIND[kernel])
    v64, d64 = _rates(w, nu, pa.FP64, kernel)
    e64v, e64d = per_particle_rel(v64, rv), per_particle_rel(d64, rd)
    out = [f"binary64 {e64v:.2e} {e64d:.2e}"]
    assert e64v < TOL64 and e64d < TOL64, (e64v, e64d)
    for sch in (pa.SCHEME_SYMMETRIC, pa.SCHEME_ONESIDED):
        v, d = _rates(w, nu, pa.FP32C, kernel, scheme=sch)
        e = (rel_rms(v, rv), rel_rms(d, rd))
        out.append(f"binary32 scheme {sch}: {e[0]:.2e} {e[1]:.2e}")
        assert e[0] < TOL32 and e[1] < TOL32, (sch, e)
    print("  " + "; ".join(out), flush=True)
