"""Shared test helpers: error metrics used by the parity tests."""
import numpy as np


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
