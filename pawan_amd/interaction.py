"""Python mirror of pawan's interaction / integration API over the HIP library.

The classes follow the reference's C++ classes method for method (names,
argument meaning, state layout) so that code written against pawan reads
the same:

    pawan                                    pawan_amd
    __interaction(W) / (W1, W2)   -> Interaction(W) / Interaction(W1, W2)
      addWake, solve, setStates, getStates, getRates, diagnose
    __integration(t, n).integrate -> Integration(t, n).integrate
    __rk4(t, n)                   -> RK4(t, n)

A wake is any object with the reference's public particle members
(``_position`` N x 3, ``_vorticity`` N x 3, ``_radius`` N, ``_volume`` N,
src/wake/wake.h:33-42); ``Wake`` below is a minimal one.  Velocities and
vorticity rates are written back to ``_velocity`` / ``_retvorcity`` on
``getRates``/``sync_wakes``, as the reference's matrices are.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

_lib_h = None


def lib():
    global _lib_h
    if _lib_h is None:
        _lib_h = _lib.load()
    return _lib_h


def _check(rc: int, handle, errfn: str):
    if rc != _lib.PW_OK:
        msg = getattr(lib(), errfn)(handle)
        raise _lib.PawanError(rc, msg.decode() if msg else "")


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None and a.shape != shape:
        a = a.reshape(shape)
    return a


class Wake:
    """Particle storage with the reference's public members (wake.h:33-42)."""

    def __init__(self, pos, vort, radius, volume):
        self._position = _f64(pos).reshape(-1, 3).copy()
        self._numParticles = self._position.shape[0]
        self._vorticity = _f64(vort).reshape(-1, 3).copy()
        self._radius = _f64(radius).reshape(-1).copy()
        self._volume = _f64(volume).reshape(-1).copy()
        self._velocity = np.zeros_like(self._position)
        self._retvorcity = np.zeros_like(self._position)
        self._size = 6 * self._numParticles  # wake.cpp:12

    @classmethod
    def from_dict(cls, w):
        return cls(w["pos"], w["vort"], w["radius"], w["volume"])


class Interaction:
    """GPU counterpart of pawan::__interaction (interaction.h:21-235)."""

    def __init__(self, *wakes, nu: float = 2.5e-3, kernel: int = _lib.KERNEL_HIGHORDER,
                 precision: int = _lib.FP64, uinf=(0.0, 0.0, 0.0), device: int = 0,
                 block: int = 0, chunk: int = 0, targets_per_thread: int = 0, unroll: int = 0,
                 scheme: int = _lib.SCHEME_DEFAULT, chunks: int = 0):
        L = lib()
        cfg = _lib.PwConfig()
        L.pw_config_default(C.byref(cfg))
        cfg.device, cfg.kernel, cfg.precision, cfg.nu = device, kernel, precision, nu
        cfg.uinf[:] = list(uinf)
        cfg.block, cfg.chunk = block, chunk
        cfg.targets_per_thread, cfg.unroll = targets_per_thread, unroll
        cfg.scheme, cfg.chunks = scheme, chunks
        h = C.c_void_p()
        rc = L.pw_create(C.byref(h), C.byref(cfg))
        if rc != _lib.PW_OK:
            raise _lib.PawanError(rc, L.pw_last_error(None).decode())
        self._h = h
        self._W = []
        self._nu = nu
        self._size = 0
        for w in wakes:
            self.addWake(w)

    def close(self):
        if getattr(self, "_h", None):
            lib().pw_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        _check(rc, self._h, "pw_last_error")

    # interaction.cpp:35-39
    def addWake(self, W):
        self._ck(lib().pw_add_wake(self._h, W._numParticles, _lib.ptr(W._position),
                                   _lib.ptr(W._vorticity), _lib.ptr(W._radius),
                                   _lib.ptr(W._volume)))
        self._W.append(W)
        self._size = lib().pw_size(self._h)

    def setNu(self, nu: float):
        self._nu = nu
        self._ck(lib().pw_set_nu(self._h, nu))

    # interaction.cpp:41-43
    def solve(self):
        self._ck(lib().pw_solve(self._h))

    # interaction.cpp:204-229
    def setStates(self, state):
        s = _f64(state)
        assert s.size == self._size
        self._ck(lib().pw_set_state(self._h, _lib.ptr(s)))

    def getStates(self, state=None):
        out = np.empty(self._size) if state is None else state
        self._ck(lib().pw_get_state(self._h, _lib.ptr(out)))
        return out

    def getRates(self, rate=None):
        out = np.empty(self._size) if rate is None else rate
        self._ck(lib().pw_get_rates(self._h, _lib.ptr(out)))
        return out

    def sync_wakes(self):
        """Copy device positions, vorticities and last rates back into the
        wakes' matrices."""
        for k, W in enumerate(self._W):
            self._ck(lib().pw_get_wake(self._h, k, _lib.ptr(W._position), _lib.ptr(W._vorticity),
                                       _lib.ptr(W._velocity), _lib.ptr(W._retvorcity)))

    def rk4(self, dt: float, nsteps: int = 1):
        self._ck(lib().pw_rk4_steps(self._h, dt, nsteps))

    def euler(self, dt: float, nsteps: int = 1):
        self._ck(lib().pw_euler_steps(self._h, dt, nsteps))

    def diagnose(self, ke: bool = True):
        """Total vorticity, linear impulse (interaction.cpp:240-259) and
        kinetic energy (interaction.cpp:351-383)."""
        om = np.zeros(3)
        im = np.zeros(3)
        k = C.c_double(0.0)
        self._ck(lib().pw_diagnostics(self._h, _lib.ptr(om), _lib.ptr(im),
                                      C.cast(C.byref(k), C.POINTER(C.c_double)) if ke else None))
        return {"total_vorticity": om, "linear_impulse": im,
                "kinetic_energy": k.value if ke else None}

    def diagnose_all(self):
        """Everything diagnose() reports (interaction.cpp:49-93) plus kinetic
        energy, computed on the device."""
        d = _lib.PwDiag()
        self._ck(lib().pw_diagnose(self._h, C.byref(d)))
        return {"total_vorticity": np.array(d.total_vorticity[:]),
                "linear_impulse": np.array(d.linear_impulse[:]),
                "angular_impulse": np.array(d.angular_impulse[:]),
                "helicity": d.helicity, "enstrophy": d.enstrophy,
                "kinetic_energy": d.kinetic_energy}

    # interaction.cpp:45-47 -> influence()
    def resolve(self):
        """Vorticity field at every particle (N x 3, wakes in order)."""
        out = np.empty((self._size // 6, 3))
        self._ck(lib().pw_resolve(self._h, _lib.ptr(out)))
        return out

    # interaction.cpp:197-202
    def write(self, path, t=None, append=True):
        """Append a frame to a .wake file (a path, or an open WakeFile): the
        time (when given, as integrate() writes it) then nWake and each
        wake's particles.  Synchronous."""
        if isinstance(path, WakeFile):
            if t is not None:
                path.write_time(t)
            self._ck(lib().pw_write(self._h, path.handle))
            return
        with WakeFile(path, append=append) as f:
            if t is not None:
                f.write_time(t)
            self._ck(lib().pw_write(self._h, f.handle))

    def write_async(self, f: "WakeFile", t=None):
        """Queue a frame of the current state into the open ``f`` (time
        first when given) and return at once; ``write_wait`` or closing ``f``
        waits for it.  The bytes are those of ``write``."""
        f.users.add(self)
        self._ck(lib().pw_write_async(self._h, f.handle, 0.0 if t is None else float(t),
                                      0 if t is None else 1))

    def write_wait(self):
        """Wait until every queued frame is in its file."""
        self._ck(lib().pw_write_wait(self._h))

    def appendParticles(self, wake: int, pos, vort, radius, volume):
        """Emit particles into wake ``wake`` (the state grows by 6 per particle)."""
        pos, vort = _f64(pos).reshape(-1, 3), _f64(vort).reshape(-1, 3)
        radius, volume = _f64(radius).reshape(-1), _f64(volume).reshape(-1)
        self._ck(lib().pw_append_particles(self._h, wake, pos.shape[0], _lib.ptr(pos),
                                           _lib.ptr(vort), _lib.ptr(radius), _lib.ptr(volume)))
        W = self._W[wake]
        W._position = np.concatenate([W._position, pos])
        W._vorticity = np.concatenate([W._vorticity, vort])
        W._radius = np.concatenate([W._radius, radius])
        W._volume = np.concatenate([W._volume, volume])
        W._numParticles = W._position.shape[0]
        W._velocity = np.zeros_like(W._position)
        W._retvorcity = np.zeros_like(W._position)
        W._size = 6 * W._numParticles
        self._size = lib().pw_size(self._h)

    def variant(self):
        """The pair-pass variant this context runs (pw_get_variant); chunks is
        the source chunk count of its job."""
        v = _lib.PwVariant()
        self._ck(lib().pw_get_variant(self._h, C.byref(v)))
        return {"block": v.block, "targets_per_thread": v.targets_per_thread,
                "unroll": v.unroll, "scheme": v.scheme, "chunks": v.chunks}

    def last_timing(self):
        t, p, n = C.c_double(), C.c_double(), C.c_double()
        dp = C.POINTER(C.c_double)
        lib().pw_last_timing(self._h, C.cast(C.byref(t), dp), C.cast(C.byref(p), dp),
                             C.cast(C.byref(n), dp))
        return {"total_ms": t.value, "pair_ms": p.value, "pairs": n.value}


class WakeFile:
    """An open .wake file (C stdio), for ``Interaction.write_async``.
    Closing it first waits for the frames queued into it."""

    def __init__(self, path, append: bool = False):
        libc = C.CDLL(None)
        libc.fopen.restype = C.c_void_p
        libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
        libc.fclose.argtypes = [C.c_void_p]
        libc.fwrite.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p]
        libc.fwrite.restype = C.c_size_t
        self._libc = libc
        self.handle = libc.fopen(str(path).encode(), b"ab" if append else b"wb")
        if not self.handle:
            raise OSError(f"cannot open {path}")
        self.users = set()

    def write_time(self, t: float):
        tt = C.c_double(t)
        if self._libc.fwrite(C.byref(tt), 8, 1, self.handle) != 1:
            raise OSError("write to the file failed")

    def close(self):
        if self.handle:
            # every writer of this file drains before the FILE is closed; the
            # first failure is raised afterwards
            err = None
            for S in self.users:
                if getattr(S, "_h", None):
                    try:
                        S.write_wait()
                    except Exception as e:  # noqa: BLE001
                        err = err or e
            self._libc.fclose(self.handle)
            self.handle = None
            if err is not None:
                raise err

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Integration:
    """pawan::__integration (integration.h:20-58): forward Euler stepping."""

    def __init__(self, t: float, n: int):
        self._t, self._n = t, n
        self._dt = t / n  # integration.cpp:10-14

    def step(self, dt: float, S: Interaction):
        S.euler(dt, 1)

    def integrate(self, S: Interaction, writer=None, path=None):
        """Advance S by n steps; ``writer(t, S)`` is called at t = 0 and after
        every step, as integrate() writes a frame per step
        (integration.cpp:16-35).  With ``path``, the frames (time, then the
        state) go to that .wake file asynchronously: each step is queued
        behind the previous frame's snapshot without waiting for its copy or
        its file write."""
        if path is not None:
            with WakeFile(path) as f:
                S.write_async(f, 0.0)
                for i in range(1, self._n + 1):
                    self.step(self._dt, S)
                    S.write_async(f, i * self._dt)
            return
        t = 0.0
        if writer:
            writer(t, S)
        for i in range(1, self._n + 1):
            t = i * self._dt
            self.step(self._dt, S)
            if writer:
                writer(t, S)


class RK4(Integration):
    """pawan::__rk4 (rk4.h:14-41): classic RK4, stages fused on device."""

    def step(self, dt: float, S: Interaction):
        S.rk4(dt, 1)

    def integrate(self, S: Interaction, writer=None, path=None):
        if writer is None and path is None:
            S.rk4(self._dt, self._n)
            return
        super().integrate(S, writer, path)


class Batch:
    """Many independent simulations stepped in one persistent launch."""

    def __init__(self, device: int = 0, kernel: int = _lib.KERNEL_HIGHORDER, block: int = 0,
                 targets_per_thread: int = 0, unroll: int = 0, scheme: int = _lib.SCHEME_DEFAULT,
                 chunks: int = 0):
        h = C.c_void_p()
        v = _lib.PwVariant(block, targets_per_thread, unroll, scheme, chunks)
        rc = lib().pw_batch_create(C.byref(h), device, kernel, C.byref(v))
        if rc != _lib.PW_OK:
            raise _lib.PawanError(rc, lib().pw_batch_last_error(None).decode())
        self._h = h
        self.n = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().pw_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        _check(rc, self._h, "pw_batch_last_error")

    def add(self, w: dict, nu: float, dt: float, precision: int = _lib.FP64, uinf=(0, 0, 0)):
        spec, keep = job_spec(w, nu, dt, precision, uinf)
        jid = C.c_int64()
        self._ck(lib().pw_batch_add(self._h, C.byref(spec), C.byref(jid)))
        self.n[jid.value] = int(w["pos"].shape[0])
        return jid.value

    def remove(self, jid):
        self._ck(lib().pw_batch_remove(self._h, jid))
        self.n.pop(jid, None)

    def step(self, nsteps: int = 1, graph: bool = True):
        self._ck(lib().pw_batch_step(self._h, nsteps, 1 if graph else 0))

    def solve(self):
        self._ck(lib().pw_batch_solve(self._h))

    def get(self, jid):
        n = self.n[jid]
        out = [np.empty((n, 3)) for _ in range(4)]
        self._ck(lib().pw_batch_get(self._h, jid, *[_lib.ptr(a) for a in out]))
        return {"pos": out[0], "vort": out[1], "vel": out[2], "dvort": out[3]}

    def status(self, jid):
        return lib().pw_batch_status(self._h, jid)

    def last_timing(self):
        t, p, n = C.c_double(), C.c_double(), C.c_double()
        dp = C.POINTER(C.c_double)
        lib().pw_batch_last_timing(self._h, C.cast(C.byref(t), dp), C.cast(C.byref(p), dp),
                                   C.cast(C.byref(n), dp))
        return {"total_ms": t.value, "pair_ms": p.value, "pairs": n.value}


def job_spec(w: dict, nu: float, dt: float, precision: int, uinf=(0, 0, 0)):
    """pw_job_spec for a generated particle set; returns (spec, arrays to keep alive)."""
    keep = [_f64(w[k]) for k in ("pos", "vort", "radius", "volume")]
    spec = _lib.PwJobSpec()
    spec.n = keep[0].shape[0]
    spec.pos, spec.vort, spec.radius, spec.volume = [_lib.ptr(a) for a in keep]
    spec.nu, spec.dt, spec.precision = nu, dt, precision
    spec.uinf[:] = list(uinf)
    return spec, keep


class Server:
    """Continuous-batching request loop (pw_server_*)."""

    QUEUED, RUNNING, DONE, FAILED, CANCELLED = 0, 1, 2, 3, 4

    def __init__(self, hbm_budget_bytes: int, device: int = 0,
                 kernel: int = _lib.KERNEL_HIGHORDER, block: int = 0, targets_per_thread: int = 0,
                 unroll: int = 0, scheme: int = _lib.SCHEME_DEFAULT, chunks: int = 0):
        h = C.c_void_p()
        v = _lib.PwVariant(block, targets_per_thread, unroll, scheme, chunks)
        rc = lib().pw_server_create(C.byref(h), device, kernel, hbm_budget_bytes, C.byref(v))
        if rc != _lib.PW_OK:
            raise _lib.PawanError(rc, lib().pw_server_last_error(None).decode())
        self._h = h
        self.n = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().pw_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        _check(rc, self._h, "pw_server_last_error")

    def now_ms(self):
        return lib().pw_server_now_ms(self._h)

    def submit(self, w, nu, dt, nsteps, precision=_lib.FP64, uinf=(0, 0, 0), arrival_ms=-1.0):
        spec, keep = job_spec(w, nu, dt, precision, uinf)
        jid = C.c_int64()
        self._ck(lib().pw_server_submit(self._h, C.byref(spec), nsteps, arrival_ms, C.byref(jid)))
        self.n[jid.value] = int(w["pos"].shape[0])
        return jid.value

    def cancel(self, jid):
        self._ck(lib().pw_server_cancel(self._h, jid))

    def step(self):
        r = C.c_int64()
        self._ck(lib().pw_server_step(self._h, C.byref(r)))
        return r.value

    def state(self, jid):
        d = C.c_int64()
        s = lib().pw_server_job_state(self._h, jid, C.byref(d))
        return s, d.value

    def fetch(self, jid):
        n = self.n[jid]
        pos, vort = np.empty((n, 3)), np.empty((n, 3))
        self._ck(lib().pw_server_fetch(self._h, jid, _lib.ptr(pos), _lib.ptr(vort)))
        self.n.pop(jid)
        return pos, vort

    def stats(self):
        st = _lib.PwServerStats()
        self._ck(lib().pw_server_get_stats(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}


def autotune(w: dict, precision: int = _lib.FP64, kernel: int = _lib.KERNEL_HIGHORDER,
             reps: int = 2, device: int = 0, cache: str = None):
    """Time every compiled pair-kernel variant on the particles of ``w`` and
    return (best, all) as dicts; optionally record the best in a tuning cache
    file (read by Interaction through the PW_TUNE_CACHE variable)."""
    keep = [_f64(w[k]) for k in ("pos", "vort", "radius", "volume")]
    n = keep[0].reshape(-1, 3).shape[0]
    allv = (_lib.PwTuneEntry * 64)()
    cnt = C.c_size_t()
    best = _lib.PwTuneEntry()
    rc = lib().pw_autotune(device, kernel, precision, n, *[_lib.ptr(a) for a in keep], reps, allv,
                           64, C.byref(cnt), C.byref(best))
    if rc != _lib.PW_OK:
        raise _lib.PawanError(rc, lib().pw_last_error(None).decode())

    def d(e):
        return {"block": e.variant.block, "targets_per_thread": e.variant.targets_per_thread,
                "unroll": e.variant.unroll, "scheme": e.variant.scheme,
                "chunks": e.variant.chunks, "pair_ms": e.pair_ms, "pairs_per_s": e.pairs_per_s}

    if cache:
        lib().pw_tune_save(str(cache).encode(), kernel, precision, n, C.byref(best))
    return d(best), [d(allv[i]) for i in range(cnt.value)]
