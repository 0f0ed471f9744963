"""pw_rates_noalloc: the rates of one wake with no device allocation (every
intermediate in the library's fixed per-device scratch array), called through
the C ABI on device arrays, against the context entry point it is a variant
of (pw_solve + pw_get_rates): binary64 rates to 1e-10 per particle, bit for
bit where both run the same pass (256 threads x 1 target, non-uniform cores).
Two calls in a row, the second smaller; the scratch keeps the first call's
contents, which the second must not see."""
import ctypes as C

import numpy as np
import pytest

import pawan_amd as pa
from pawan_amd import _lib
from pawan_amd import generators as gen
from helpers import per_particle_rel

pytestmark = pytest.mark.gpu


def _context_rates(w, nu, uinf, kernel):
    S = pa.Interaction(pa.Wake.from_dict(w), nu=nu, uinf=uinf, kernel=kernel, block=256,
                       targets_per_thread=1, unroll=2)
    S.solve()
    r = S.getRates()
    S.close()
    n = w["pos"].shape[0]
    return r[:3 * n].reshape(n, 3), r[3 * n:].reshape(n, 3)


class _Hip:
    """The caller's side: device arrays from the HIP runtime the library
    itself links (its already-loaded libamdhip64), through ctypes."""

    def __init__(self):
        _lib.load()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        self.live = []

    def ck(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.ck(self.h.hipMalloc(C.byref(p), nbytes))
        self.live.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.alloc(a.nbytes)
        self.ck(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1))
        return p

    def download(self, p, shape):
        out = np.empty(shape, dtype=np.float64)
        self.ck(self.h.hipDeviceSynchronize())
        self.ck(self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2))
        return out

    def free_bytes(self):
        f, t = C.c_size_t(), C.c_size_t()
        self.ck(self.h.hipDeviceSynchronize())
        self.ck(self.h.hipMemGetInfo(C.byref(f), C.byref(t)))
        return f.value

    def release(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _noalloc(hip, w, nu, uinf, kernel):
    L = _lib.load()
    n = w["pos"].shape[0]
    d = {k: hip.upload(w[k]) for k in ("pos", "vort", "radius", "volume")}
    nan = np.full((n, 3), np.nan)
    vel, dvort = hip.upload(nan), hip.upload(nan)
    u = np.asarray(uinf, dtype=np.float64)
    rc = L.pw_rates_noalloc(kernel, nu, u.ctypes.data_as(C.POINTER(C.c_double)), n, d["pos"],
                            d["vort"], d["radius"], d["volume"], vel, dvort, None)
    assert rc == 0, L.pw_last_error(None)
    out = hip.download(vel, (n, 3)), hip.download(dvort, (n, 3))
    hip.release()
    return out


def test_rates_noalloc_matches_context_solve():
    hip = _Hip()
    big = gen.rotor_wake(40)  # 10,240 particles, one core size
    small = gen.vortex_ring(50, 20, R=1.0, core=0.1, seed=8)  # 1,000 particles
    small["radius"] = small["radius"] * (1.0 + 0.3 * gen.uniform(17, small["n"]))
    cases = [("rotor wake", big, 1e-4, (0.0, 0.0, -0.05), pa.KERNEL_HIGHORDER),
             ("ring, non-uniform cores", small, 1e-3, (0.0, 0.0, 0.0), pa.KERNEL_HIGHORDER),
             ("ring, non-uniform cores, GAUSSIAN", small, 1e-3, (0.01, 0.0, 0.0),
              pa.KERNEL_GAUSSIAN)]
    for name, w, nu, uinf, kernel in cases:
        n = w["pos"].shape[0]
        print(f"case no-alloc {name} N={n} kernel={kernel}", flush=True)
        v, d = _noalloc(hip, w, nu, uinf, kernel)
        rv, rd = _context_rates(w, nu, uinf, kernel)
        ev, ed = per_particle_rel(v, rv), per_particle_rel(d, rd)
        print(f"  vs pw_solve: velocity {ev:.3e} vorticity rate {ed:.3e}", flush=True)
        assert np.all(np.isfinite(v)) and np.all(np.isfinite(d))
        assert ev < 1e-10 and ed < 1e-10
        if name.startswith("ring"):  # same pass, same cores: the same sums
            assert np.array_equal(v, rv) and np.array_equal(d, rd)
    # a call allocates nothing: with the caller's arrays in place, the
    # device's free memory is the same after it
    L = _lib.load()
    w = small
    n = w["pos"].shape[0]
    d = {k: hip.upload(w[k]) for k in ("pos", "vort", "radius", "volume")}
    out = hip.alloc(2 * 3 * n * 8)
    free0 = hip.free_bytes()
    rc = L.pw_rates_noalloc(pa.KERNEL_HIGHORDER, 1e-3, None, n, d["pos"], d["vort"], d["radius"],
                            d["volume"], out, C.c_void_p(out.value + 3 * n * 8), None)
    assert rc == 0
    free1 = hip.free_bytes()
    hip.release()
    print(f"  free device memory before/after a call: {free0} / {free1}", flush=True)
    assert free1 == free0


def test_rates_noalloc_rejects_beyond_limit():
    hip = _Hip()
    L = _lib.load()
    n = 16385
    buf = hip.alloc(3 * n * 8)
    one = hip.alloc(n * 8)
    rc = L.pw_rates_noalloc(pa.KERNEL_HIGHORDER, 1e-3, None, n, buf, buf, one, one, buf, buf,
                            None)
    hip.release()
    msg = L.pw_last_error(None).decode()
    print(f"  n = {n}: status {rc}, '{msg}'", flush=True)
    assert rc == 1 and "16384" in msg
    assert L.pw_noalloc_scratch_bytes() < 32 << 20
