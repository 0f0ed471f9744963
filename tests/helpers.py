"""Shared test helpers: error metrics and test particle sets used by the
parity tests."""
import numpy as np

from pawan_amd import generators as gen


def per_particle_rel(got, ref):
    """max_i |got_i - ref_i| / max(|ref_i|, 1e-3 * rms_k |ref_k|) over particles
    (rows of N x 3 arrays)."""
    got = np.asarray(got).reshape(-1, 3)
    ref = np.asarray(ref).reshape(-1, 3)
    nr = np.linalg.norm(ref, axis=1)
    floor = 1e-3 * np.sqrt(np.mean(nr * nr))
    return float(np.max(np.linalg.norm(got - ref, axis=1) / np.maximum(nr, floor)))


def rel_rms(got, ref):
    """sqrt(sum |got - ref|^2 / sum |ref|^2)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.sqrt(np.sum((got - ref) ** 2) / np.sum(ref ** 2)))


def sample_targets(n, k, seed=0, edges=(64, 256, 512, 16384)):
    """A stratified sample of about k target indices of an n-particle system:
    one index drawn from each of k equal strata (seeded), plus the ends and the
    indices on either side of every multiple of the tile, block and chunk
    lengths in `edges` (where a pass's partial sums are split).  Sorted, unique."""
    rng = np.random.default_rng(seed)
    k = max(1, min(k, n))
    bounds = np.linspace(0, n, k + 1).astype(np.int64)
    pick = [int(rng.integers(a, b)) for a, b in zip(bounds[:-1], bounds[1:]) if b > a]
    extra = [0, 1, n - 2, n - 1]
    for e in edges:
        for m in range(e, n, e * max(1, n // (e * 8))):
            extra += [m - 1, m]
    idx = np.unique(np.clip(np.array(pick + extra, dtype=np.int64), 0, n - 1))
    return idx


def sparse_lattice(sub=4, cubes=4, spacing=10.0, core=0.2, seed=7):
    """Particles `spacing` apart on a (sub x cubes)^3 lattice, ordered sub-cube
    by sub-cube (sub^3 = 64 particles: one tile each), jittered by 10%;
    vorticity mostly along z.  Every distance appears as often as its
    influence decays, so the pairs between different sub-cubes carry most of
    every velocity.  With the defaults the boxes of neighbouring sub-cubes
    are at least 0.8 spacings = 40 core radii apart."""
    g = np.stack(np.meshgrid(*[np.arange(sub)] * 3, indexing="ij"), -1).reshape(-1, 3)
    c = np.stack(np.meshgrid(*[np.arange(cubes)] * 3, indexing="ij"), -1).reshape(-1, 3)
    lat = (c[:, None, :] * sub + g[None, :, :]).reshape(-1, 3).astype(float)
    n = lat.shape[0]
    u = gen.uniform(seed, 6 * n).reshape(n, 6)
    pos = (lat - (sub * cubes - 1) / 2.0 + 0.1 * (2.0 * u[:, :3] - 1.0)) * spacing
    vort = 0.01 * np.concatenate([0.2 * (u[:, 3:5] - 0.5), 0.5 + u[:, 5:6]], -1)
    return {"pos": np.ascontiguousarray(pos), "vort": np.ascontiguousarray(vort),
            "radius": np.full(n, core), "volume": np.full(n, 4.0 / 3.0 * np.pi * core ** 3),
            "n": n}
