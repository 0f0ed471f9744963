"""The C ABI library: loads, exports every function include/pawan_hip.h
declares, and its host-only entry points behave (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

import pawan_amd
from pawan_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pawan_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pw_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared_functions()
    assert len(names) >= 35
    for must in ("pw_create", "pw_solve", "pw_rk4_steps", "pw_batch_step", "pw_server_step"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_host_only_entry_points():
    lib = _lib.load()
    assert lib.pw_version().decode().startswith("pawan_hip")
    cfg = _lib.PwConfig()
    assert lib.pw_config_default(C.byref(cfg)) == 0
    assert cfg.nu == 2.5e-3 and cfg.kernel == 0 and cfg.precision == 0
    # chunk policy: multiple of the block; 16 chunks, or enough for ~4,096
    # tile x chunk items when the tiles are few; a function of (n, block) only
    for n in (1, 100, 1024, 2048, 65536, 262144, 1000003):
        for b in (64, 128, 256):
            L = lib.pw_chunk_for(n, b)
            ntile, nchunk = -(-n // b), -(-n // L)
            assert L % b == 0 and nchunk <= max(16, -(-4096 // ntile)), (n, b, nchunk)
            assert nchunk >= min(16, ntile) or L == b, (n, b, nchunk)
    assert -(-2048 // lib.pw_chunk_for(2048, 64)) == 32     # one 64-source tile per chunk
    assert -(-262144 // lib.pw_chunk_for(262144, 256)) == 16
    assert lib.pw_job_footprint(262144, 0) > 262144 * 64
    # binary32 records carry binary64 strengths too: 80 B instead of 64 B.
    # (Contexts under 8,192 particles take 64-thread blocks.)
    # The symmetric pass keeps NCOMP_S = 10 binary64 sums per particle and
    # partial row: blocks of tw = 256 targets (256 x 1) up to 524,288
    # particles, 512 (512 x 1) beyond; source chunks and strips of L =
    # tw x ceil(nblk / nstrip) particles, nstrip = min(nblk, max(128, nblk / 8));
    # target-side row c holds the targets below (c + 1) L, source-side row s
    # the particles from s L on.  The one-sided binary64 pass keeps 10 sums
    # per particle and chunk, 16 chunks.
    for n, tw in ((262144, 256), (524288, 256), (786432, 512), (65536, 256), (131072, 256),
                  (4096, 64)):
        npad = n
        nblk = npad // tw
        nstrip = min(nblk, max(128, nblk // 8))
        L = tw * -(-nblk // nstrip)
        tri = sum(min((c + 1) * L, npad) for c in range(-(-n // L)))
        rows = sum(npad - s * L for s in range(-(-npad // L)))
        nch64 = -(-n // lib.pw_chunk_for(n, tw if tw < 512 else 256))
        want = 16 * n + 8 * 10 * (tri + rows) - 8 * 10 * nch64 * npad
        assert lib.pw_job_footprint(n, 1) - lib.pw_job_footprint(n, 0) == want, n
    # the slots of a 262,144-particle binary32 job: under 3 GB (1.35 + 1.35)
    assert lib.pw_job_footprint(262144, 1) < 3.0e9


def test_null_arguments_rejected():
    lib = _lib.load()
    assert lib.pw_create(None, None) == _lib.PW_EINVAL
    assert lib.pw_destroy(None) == _lib.PW_EINVAL
    assert lib.pw_solve(None) == _lib.PW_EINVAL
    assert lib.pw_rk4_steps(None, 0.01, 1) == _lib.PW_EINVAL
    assert lib.pw_batch_step(None, 1, 1) == _lib.PW_EINVAL
    assert lib.pw_server_step(None, None) == _lib.PW_EINVAL
    assert lib.pw_config_default(None) == _lib.PW_EINVAL


def test_invalid_config_rejected_before_device_use():
    lib = _lib.load()
    cfg = _lib.PwConfig()
    lib.pw_config_default(C.byref(cfg))
    h = C.c_void_p()
    for field, bad in (("kernel", 7), ("precision", 5), ("block", 96), ("chunks", -1)):
        c2 = _lib.PwConfig.from_buffer_copy(cfg)
        setattr(c2, field, bad)
        assert lib.pw_create(C.byref(h), C.byref(c2)) == _lib.PW_EINVAL
        assert h.value is None
    # block 512 exists only for the binary32 symmetric HIGHORDER pass at one
    # target per thread
    for changes in ({"precision": _lib.FP64}, {"scheme": _lib.SCHEME_ONESIDED},
                    {"kernel": _lib.KERNEL_GAUSSIAN}, {"targets_per_thread": 2}):
        c2 = _lib.PwConfig.from_buffer_copy(cfg)
        c2.block, c2.precision = 512, _lib.FP32C
        for field, v in changes.items():
            setattr(c2, field, v)
        assert lib.pw_create(C.byref(h), C.byref(c2)) == _lib.PW_EINVAL, changes
        assert h.value is None


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_device_is_an_error_not_a_crash():
    lib = _lib.load()
    h = C.c_void_p()
    rc = lib.pw_create(C.byref(h), None)
    assert rc in (_lib.PW_EHIP, _lib.PW_EINVAL) and h.value is None
    assert lib.pw_last_error(None)


def test_tune_cache_roundtrip(tmp_path):
    lib = _lib.load()
    path = str(tmp_path / "tune.txt").encode()
    v = _lib.PwVariant()
    assert lib.pw_tune_lookup(path, 0, 0, 262144, C.byref(v)) == _lib.PW_ENOTFOUND
    e = _lib.PwTuneEntry(_lib.PwVariant(256, 2, 4), 123.0, 5e11)
    assert lib.pw_tune_save(path, 0, 0, 262144, C.byref(e)) == 0
    e2 = _lib.PwTuneEntry(_lib.PwVariant(64, 2, 0, _lib.SCHEME_SYMMETRIC), 1.0, 1e9)
    assert lib.pw_tune_save(path, 0, 1, 2048, C.byref(e2)) == 0
    e3 = _lib.PwTuneEntry(_lib.PwVariant(128, 1, 2), 120.0, 5e11)
    assert lib.pw_tune_save(path, 0, 0, 300000, C.byref(e3)) == 0  # same log2 bucket: replaces
    assert lib.pw_tune_lookup(path, 0, 0, 262144, C.byref(v)) == 0
    assert (v.block, v.targets_per_thread, v.unroll) == (128, 1, 2)
    assert lib.pw_tune_lookup(path, 0, 1, 3000, C.byref(v)) == 0
    assert (v.block, v.targets_per_thread, v.unroll, v.scheme) == (64, 2, 0, _lib.SCHEME_SYMMETRIC)
    assert lib.pw_tune_lookup(path, 1, 0, 262144, C.byref(v)) == _lib.PW_ENOTFOUND
    assert len(open(path.decode()).read().splitlines()) == 2
    # the 512-thread binary32 symmetric variant the tuner picks at N = 262,144
    e4 = _lib.PwTuneEntry(_lib.PwVariant(512, 1, 0, _lib.SCHEME_SYMMETRIC), 60.6, 1.13e12)
    assert lib.pw_tune_save(path, 0, 1, 262144, C.byref(e4)) == 0
    assert lib.pw_tune_lookup(path, 0, 1, 262144, C.byref(v)) == 0
    assert (v.block, v.targets_per_thread, v.unroll, v.scheme) == (512, 1, 0, _lib.SCHEME_SYMMETRIC)
    assert v.chunks == 0
    assert len(open(path.decode()).read().splitlines()) == 3
    # a tuned chunk count is the last field of a line
    e5 = _lib.PwTuneEntry(_lib.PwVariant(512, 1, 0, _lib.SCHEME_SYMMETRIC, 96), 60.1, 1.14e12)
    assert lib.pw_tune_save(path, 0, 1, 262144, C.byref(e5)) == 0
    assert lib.pw_tune_lookup(path, 0, 1, 262144, C.byref(v)) == 0
    assert (v.block, v.scheme, v.chunks) == (512, _lib.SCHEME_SYMMETRIC, 96)
    line = [l for l in open(path.decode()).read().splitlines() if l.startswith("0 1 18 ")]
    assert line and line[0].split()[8] == "96"


def test_tune_cache_line_without_chunks(tmp_path):
    """Cache lines written before chunk counts were tuned end at ms: they
    read as the chunk policy (chunks 0)."""
    lib = _lib.load()
    p = tmp_path / "old.txt"
    p.write_text("0 0 18 256 2 2 1 117.4\n")
    v = _lib.PwVariant(0, 0, 0, 0, 77)
    assert lib.pw_tune_lookup(str(p).encode(), 0, 0, 262144, C.byref(v)) == 0
    assert (v.block, v.targets_per_thread, v.unroll, v.scheme, v.chunks) == (256, 2, 2, 1, 0)
    p.write_text("0 0 18 256 2 2 1 117.4 -3\n")
    assert lib.pw_tune_lookup(str(p).encode(), 0, 0, 262144, C.byref(v)) == _lib.PW_EINVAL
