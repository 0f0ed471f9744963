"""Kernel variants (workgroup size, targets per thread, unroll, pair scheme)
and the autotuner.  One-sided variants give bitwise the same rates; the
symmetric binary32 pass groups its sums by block, so its variants agree to
binary32 rounding; the tuner times them all."""
import numpy as np
import pytest

import pawan_amd as pa
from pawan_amd import generators as gen
from helpers import rel_rms

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [pa.FP64, pa.FP32C])
def test_onesided_variants_bitwise_identical(precision):
    w = gen.vortex_ring(100, 30, R=1.0, core=0.1, seed=41)  # 3,000 particles
    print(f"case variants precision={precision} N={w['n']}", flush=True)
    ref = None
    for block in (64, 128, 256):
        for tpt in (1, 2):
            for unroll in (1, 2, 4):
                S = pa.Interaction(pa.Wake.from_dict(w), nu=1e-3, precision=precision, block=block,
                                   chunk=768, targets_per_thread=tpt, unroll=unroll,
                                   scheme=pa.SCHEME_ONESIDED)
                S.solve()
                r = S.getRates()
                S.close()
                if ref is None:
                    ref = r
                assert np.array_equal(r, ref), (block, tpt, unroll)


def test_symmetric_variants_agree():
    w = gen.vortex_ring(100, 30, R=1.0, core=0.1, seed=41)
    n = w["n"]
    S = pa.Interaction(pa.Wake.from_dict(w), nu=1e-3)
    S.solve()
    r64 = S.getRates()
    S.close()
    for block in (64, 128, 256):
        for tpt in (1, 2):
            for chunk in (0, 768):
                S = pa.Interaction(pa.Wake.from_dict(w), nu=1e-3, precision=pa.FP32C, block=block,
                                   chunk=chunk, targets_per_thread=tpt,
                                   scheme=pa.SCHEME_SYMMETRIC)
                S.solve()
                r = S.getRates()
                S.close()
                ev, ed = rel_rms(r[:3 * n], r64[:3 * n]), rel_rms(r[3 * n:], r64[3 * n:])
                print(f"  block {block} tpt {tpt} chunk {chunk}: vel {ev:.2e} dvort {ed:.2e}",
                      flush=True)
                assert ev < 1e-6 and ed < 1e-6, (block, tpt, chunk)


def test_autotune_small(tmp_path):
    w = gen.vortex_ring(64, 16, seed=42)
    best, allv = pa.autotune(w, pa.FP64, reps=1, cache=str(tmp_path / "tune.txt"))
    print(f"  best {best}", flush=True)
    assert {e["scheme"] for e in allv} == {pa.SCHEME_SYMMETRIC, pa.SCHEME_ONESIDED}
    assert len(allv) == 6 + 18 and best["pair_ms"] == min(e["pair_ms"] for e in allv)
    assert (tmp_path / "tune.txt").read_text().split()[3:7] == [
        str(best["block"]), str(best["targets_per_thread"]), str(best["unroll"]),
        str(best["scheme"])]


def test_autotune_fp32c_both_schemes(tmp_path):
    w = gen.vortex_ring(64, 16, seed=43)
    best, allv = pa.autotune(w, pa.FP32C, reps=1)
    print(f"  best {best}", flush=True)
    schemes = {e["scheme"] for e in allv}
    assert schemes == {pa.SCHEME_SYMMETRIC, pa.SCHEME_ONESIDED}
    # symmetric: 3 blocks x 2 targets per thread, plus 512 x 1; one-sided: 3 x 2 x 3 unrolls
    assert len(allv) == 7 + 18 and best["pair_ms"] == min(e["pair_ms"] for e in allv)
    big = [(e["targets_per_thread"], e["scheme"]) for e in allv if e["block"] == 512]
    assert big == [(1, pa.SCHEME_SYMMETRIC)]
