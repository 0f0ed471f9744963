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
