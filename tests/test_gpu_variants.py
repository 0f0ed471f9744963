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


def _check_chunk_entries(allv, nbase):
    """The tuner's second sweep: the fastest variant of each scheme at other
    chunk counts (chunks > 0), after the nbase policy entries (chunks 0)."""
    base = [e for e in allv if e["chunks"] == 0]
    assert len(base) == nbase and allv[:nbase] == base
    key = ("block", "targets_per_thread", "unroll", "scheme")
    for e in allv[nbase:]:
        fastest = min((b for b in base if b["scheme"] == e["scheme"]), key=lambda b: b["pair_ms"])
        assert e["chunks"] > 0 and all(e[k] == fastest[k] for k in key), e
    return allv[nbase:]


def test_autotune_small(tmp_path):
    w = gen.vortex_ring(64, 16, seed=42)
    best, allv = pa.autotune(w, pa.FP64, reps=1, cache=str(tmp_path / "tune.txt"))
    print(f"  best {best}", flush=True)
    assert {e["scheme"] for e in allv} == {pa.SCHEME_SYMMETRIC, pa.SCHEME_ONESIDED}
    extra = _check_chunk_entries(allv, 6 + 18)
    print(f"  chunk counts tried: {[(e['scheme'], e['chunks']) for e in extra]}", flush=True)
    assert best["pair_ms"] == min(e["pair_ms"] for e in allv)
    fields = (tmp_path / "tune.txt").read_text().split()
    assert fields[3:7] == [str(best["block"]), str(best["targets_per_thread"]),
                           str(best["unroll"]), str(best["scheme"])]
    assert fields[8] == str(best["chunks"])


def test_autotune_fp32c_both_schemes(tmp_path):
    w = gen.vortex_ring(64, 16, seed=43)
    best, allv = pa.autotune(w, pa.FP32C, reps=1)
    print(f"  best {best}", flush=True)
    schemes = {e["scheme"] for e in allv}
    assert schemes == {pa.SCHEME_SYMMETRIC, pa.SCHEME_ONESIDED}
    # symmetric: 3 blocks x 2 targets per thread, plus 512 x 1; one-sided: 3 x 2 x 3 unrolls
    _check_chunk_entries(allv, 7 + 18)
    assert best["pair_ms"] == min(e["pair_ms"] for e in allv)
    big = {(e["targets_per_thread"], e["scheme"]) for e in allv if e["block"] == 512}
    assert big == {(1, pa.SCHEME_SYMMETRIC)}


def _chunk_count(n, block, tpt, sym, chunks):
    """Source chunks of a job asked for `chunks` (the library's rounding:
    whole blocks of block x tpt targets on the symmetric pass, whole tiles of
    block one-sided)."""
    unit = block * tpt if sym else block
    nunit = -(-n // unit)
    L = unit * -(-nunit // chunks)
    return -(-n // L)


@pytest.mark.parametrize("precision,scheme,tpt", [(pa.FP64, pa.SCHEME_ONESIDED, 2),
                                                  (pa.FP64, pa.SCHEME_SYMMETRIC, 1),
                                                  (pa.FP32C, pa.SCHEME_SYMMETRIC, 1)])
def test_chunk_counts_vs_oracle_batch_and_cache(precision, scheme, tpt, tmp_path, monkeypatch):
    """A job asked for a source chunk count (pw_config.chunks, a batch's
    pw_variant.chunks, the tuning cache's last field) runs with that many
    chunks (pw_get_variant), matches the oracle on sampled targets (binary64
    1e-10 per particle, binary32 1e-6 relative RMS), and gives bitwise the
    same rates in a batch as in a context of the same block, targets per
    thread and count."""
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
    import oracle as orc
    from helpers import per_particle_rel, sample_targets
    w = gen.rotor_wake(64)  # 16,384 particles
    n, nu = w["n"], 1e-4
    sym = scheme == pa.SCHEME_SYMMETRIC
    tg = sample_targets(n, 256, seed=5)
    ov, ost, odf = orc.rates_sample(w["pos"], w["vort"], w["radius"], w["volume"], nu, tg)
    od = ost + odf
    for chunks in (0, 8, 40):
        print(f"case chunks precision={precision} scheme={scheme} N={n} chunks={chunks}",
              flush=True)
        S = pa.Interaction(pa.Wake.from_dict(w), nu=nu, precision=precision, scheme=scheme,
                           block=256, targets_per_thread=tpt, chunks=chunks)
        S.solve()
        r = S.getRates()
        got = S.variant()["chunks"]
        S.close()
        if chunks:
            assert got == _chunk_count(n, 256, tpt, sym, chunks)
        vel, dv = r[:3 * n].reshape(n, 3)[tg], r[3 * n:].reshape(n, 3)[tg]
        if precision == pa.FP64:
            ev, ed = per_particle_rel(vel, ov), per_particle_rel(dv, od)
            assert ev < 1e-10 and ed < 1e-10, (chunks, ev, ed)
        else:
            ev, ed = rel_rms(vel, ov), rel_rms(dv, od)
            assert ev < 1e-6 and ed < 1e-6, (chunks, ev, ed)
        print(f"  {got} chunks: velocity {ev:.2e} vorticity rate {ed:.2e}", flush=True)
        if chunks:
            B = pa.Batch(block=256, targets_per_thread=tpt, scheme=scheme, chunks=chunks)
            jid = B.add(w, nu, 0.0, precision)
            B.solve()
            g = B.get(jid)
            B.close()
            assert np.array_equal(g["vel"].ravel(), r[:3 * n])
            assert np.array_equal(g["dvort"].ravel(), r[3 * n:])
    # the tuning cache's chunk count reaches a context that leaves the variant at 0
    cache = tmp_path / "tune.txt"
    bucket = n.bit_length() - 1
    cache.write_text(f"0 {precision} {bucket} 256 {tpt} 0 {scheme} 1.0 8\n")
    monkeypatch.setenv("PW_TUNE_CACHE", str(cache))
    S = pa.Interaction(pa.Wake.from_dict(w), nu=nu, precision=precision)
    S.solve()
    v = S.variant()
    S.close()
    assert (v["block"], v["targets_per_thread"], v["scheme"]) == (256, tpt, scheme)
    assert v["chunks"] == _chunk_count(n, 256, tpt, sym, 8)
