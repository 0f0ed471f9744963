"""Kernel variants (workgroup size, targets per thread, unroll) and the
autotuner: every variant gives bitwise the same rates; the tuner times them."""
import numpy as np
import pytest

import pawan_amd as pa
from pawan_amd import generators as gen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [pa.FP64, pa.FP32C])
def test_variants_bitwise_identical(precision):
    w = gen.vortex_ring(100, 30, R=1.0, core=0.1, seed=41)  # 3,000 particles
    print(f"case variants precision={precision} N={w['n']}", flush=True)
    ref = None
    for block in (64, 128, 256):
        for tpt in (1, 2):
            for unroll in (1, 2, 4):
                S = pa.Interaction(pa.Wake.from_dict(w), nu=1e-3, precision=precision, block=block,
                                   chunk=768, targets_per_thread=tpt, unroll=unroll)
                S.solve()
                r = S.getRates()
                S.close()
                if ref is None:
                    ref = r
                assert np.array_equal(r, ref), (block, tpt, unroll)


def test_autotune_small(tmp_path):
    w = gen.vortex_ring(64, 16, seed=42)
    best, allv = pa.autotune(w, pa.FP64, reps=1, cache=str(tmp_path / "tune.txt"))
    print(f"  best {best}", flush=True)
    assert len(allv) == 18 and best["pair_ms"] == min(e["pair_ms"] for e in allv)
    assert (tmp_path / "tune.txt").read_text().split()[3:6] == [
        str(best["block"]), str(best["targets_per_thread"]), str(best["unroll"])]
