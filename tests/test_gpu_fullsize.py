"""Size-independent properties at the full metric size (N = 262,144), where
the CPU oracle is too slow to run in a test: the oracle-checked binary64 path
must conserve total vorticity rate (pair terms are odd under exchange), be
translation invariant, agree with itself across kernel variants and chunk
lengths, and the compensated binary32 path must stay within 1e-6 of it."""
import numpy as np
import pytest

import pawan_amd as pa
from pawan_amd import generators as gen
from helpers import per_particle_rel, rel_rms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wake():
    return gen.config_rotor_wake(1024)


def _rates(w, **kw):
    S = pa.Interaction(pa.Wake.from_dict(w), nu=w["nu"], uinf=w["uinf"], **kw)
    S.solve()
    r = S.getRates()
    S.close()
    n = w["n"]
    return r[:3 * n].reshape(n, 3), r[3 * n:].reshape(n, 3)


def test_fullsize_conservation_translation_variants(wake):
    w = wake
    print(f"case full size N={w['n']}", flush=True)
    v, d = _rates(w)
    # sum of dalpha/dt vanishes (each pair adds +da to one particle, -da to the other)
    assert np.abs(d.sum(axis=0)).max() <= 1e-12 * np.abs(d).sum()
    # translation invariance (shifted positions, same rates)
    w2 = dict(w)
    w2["pos"] = w["pos"] + np.array([0.75, -0.5, 1.25])
    v2, d2 = _rates(w2)
    ev, ed = per_particle_rel(v2, v), per_particle_rel(d2, d)
    print(f"  translation: vel {ev:.3e} dvort {ed:.3e}", flush=True)
    assert ev < 1e-10 and ed < 1e-9
    # another variant and chunk length: same sums up to reassociation
    v3, d3 = _rates(w, block=128, targets_per_thread=2, unroll=4, chunk=4096)
    ev, ed = per_particle_rel(v3, v), per_particle_rel(d3, d)
    print(f"  variant/chunk: vel {ev:.3e} dvort {ed:.3e}", flush=True)
    assert ev < 1e-12 and ed < 1e-10


def test_fullsize_fp32c_within_1e6(wake):
    w = wake
    v, d = _rates(w)
    v32, d32 = _rates(w, precision=pa.FP32C)
    ev, ed = rel_rms(v32, v), rel_rms(d32, d)
    print(f"  fp32c vs fp64 at N={w['n']}: velocity {ev:.3e} vorticity rate {ed:.3e}", flush=True)
    assert ev < 1e-6 and ed < 1e-6


def test_twice_fullsize_fp32c_default_within_1e6():
    """N = 524,288, the largest size whose binary32 default keeps 256-target
    blocks (43 GB of source-side slots): the symmetric pass stays within
    1e-6 of the binary64 path and conserves the total vorticity rate to
    binary32 rounding (a pair's +s and -s are summed in different binary32
    runs: measured 6e-8 of the summed magnitudes)."""
    w = gen.config_rotor_wake(2048)
    print(f"case N={w['n']} binary32 default variant", flush=True)
    v, d = _rates(w)
    v32, d32 = _rates(w, precision=pa.FP32C)
    ev, ed = rel_rms(v32, v), rel_rms(d32, d)
    print(f"  fp32c vs fp64: velocity {ev:.3e} vorticity rate {ed:.3e}", flush=True)
    assert ev < 1e-6 and ed < 1e-6
    assert np.abs(d32.sum(axis=0)).max() <= 1e-6 * np.abs(d32).sum()


def test_fullsize_fp32c_default_beyond_524288():
    """N = 786,432: past 524,288 particles the binary32 default is the 512 x 1
    symmetric pass (one slot row per 512 targets, 48 GB here): within 1e-6 of
    the binary64 path, total vorticity rate conserved to binary32 rounding."""
    w = gen.config_rotor_wake(3072)
    print(f"case N={w['n']} binary32 default variant (512 x 1)", flush=True)
    v, d = _rates(w)
    v32, d32 = _rates(w, precision=pa.FP32C)
    ev, ed = rel_rms(v32, v), rel_rms(d32, d)
    print(f"  fp32c vs fp64: velocity {ev:.3e} vorticity rate {ed:.3e}", flush=True)
    assert ev < 1e-6 and ed < 1e-6
    assert np.abs(d32.sum(axis=0)).max() <= 1e-6 * np.abs(d32).sum()
