"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, never by the pawan_amd package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

HIGHORDER, GAUSSIAN = 0, 1
SERIAL, OMP_FAITHFUL, OMP_CLEAN, ONESIDED_LD = 0, 1, 2, 3

_dp = C.POINTER(C.c_double)
_sp = C.POINTER(C.c_size_t)
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run make -C oracle")
    L = C.CDLL(LIB_PATH)
    sig = {
        "orc_kernel": (None, [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]),
        "orc_dnrm2": (C.c_double, [C.c_size_t, _dp]),
        "orc_pair": (None, [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                            C.c_double, C.c_double, _dp, _dp, _dp, _dp]),
        "orc_rates": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_size_t, _sp, _dp, _dp, _dp, _dp,
                                _dp, _dp, _dp]),
        "orc_rates_split_ld": (C.c_int, [C.c_int, C.c_double, C.c_size_t, _dp, _dp, _dp, _dp,
                                         _dp, _dp, _dp]),
        "orc_rates_sample": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_size_t, _dp, _dp, _dp,
                                       _dp, _dp, C.c_size_t, _sp, _dp, _dp, _dp]),
        "orc_pack_state": (None, [C.c_size_t, _sp, _dp, _dp, _dp]),
        "orc_unpack_state": (None, [C.c_size_t, _sp, _dp, _dp, _dp]),
        "orc_rk4": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_size_t, _sp, _dp, _dp, _dp, _dp,
                              C.c_double, C.c_size_t]),
        "orc_euler": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_size_t, _sp, _dp, _dp, _dp,
                                _dp, C.c_double, C.c_size_t]),
        "orc_total_vorticity": (None, [C.c_size_t, _sp, _dp, _dp]),
        "orc_linear_impulse": (None, [C.c_size_t, _sp, _dp, _dp, _dp]),
        "orc_angular_impulse": (None, [C.c_size_t, _sp, _dp, _dp, _dp]),
        "orc_kinetic_energy": (C.c_double, [C.c_size_t, _sp, _dp, _dp, _dp]),
        "orc_kinetic_energy_cross": (C.c_double, [C.c_size_t, C.c_size_t, _dp, _dp, _dp]),
        "orc_enstrophy": (C.c_double, [C.c_size_t, _sp, _dp, _dp, _dp]),
        "orc_helicity": (C.c_double, [C.c_size_t, _sp, _dp, _dp, _dp]),
        "orc_influence": (C.c_int, [C.c_int, C.c_size_t, _sp, _dp, _dp, _dp, _dp]),
        "orc_omp_rows": (C.c_double, [C.c_int, C.c_int, C.c_double, C.c_size_t, _dp, _dp, _dp,
                                      _dp, C.c_size_t, C.c_size_t, C.c_size_t, _dp, _dp]),
        "orc_max_threads": (C.c_int, []),
        "orc_wtime": (C.c_double, []),
    }
    for k, (r, a) in sig.items():
        f = getattr(L, k)
        f.restype = r
        f.argtypes = a
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _f(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _wakes(wake_n, n):
    w = np.ascontiguousarray(np.asarray(wake_n if wake_n is not None else [n], dtype=np.uintp))
    return w, len(w), w.ctypes.data_as(_sp)


def kernel(rho, sigma, kind=HIGHORDER):
    q, F, Z = C.c_double(), C.c_double(), C.c_double()
    load().orc_kernel(kind, rho, sigma, C.byref(q), C.byref(F), C.byref(Z))
    return q.value, F.value, Z.value


def dnrm2(x):
    x = _f(x)
    return load().orc_dnrm2(x.size, _p(x))


def rates(pos, vort, radius, volume, nu, mode=SERIAL, kind=HIGHORDER, wake_n=None, uinf=None):
    pos, vort, radius, volume = _f(pos), _f(vort), _f(radius), _f(volume)
    n = radius.size
    w, nw, wp = _wakes(wake_n, n)
    vel, dv = np.zeros((n, 3)), np.zeros((n, 3))
    u = None if uinf is None else _f(uinf)
    rc = load().orc_rates(mode, kind, nu, nw, wp, _p(pos), _p(vort), _p(radius), _p(volume),
                          _p(u), _p(vel), _p(dv))
    if rc:
        raise ValueError("orc_rates failed")
    return vel, dv


def rates_split(pos, vort, radius, volume, nu, kind=HIGHORDER):
    pos, vort, radius, volume = _f(pos), _f(vort), _f(radius), _f(volume)
    n = radius.size
    vel, st, df = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    load().orc_rates_split_ld(kind, nu, n, _p(pos), _p(vort), _p(radius), _p(volume), _p(vel),
                              _p(st), _p(df))
    return vel, st, df


def rates_sample(pos, vort, radius, volume, nu, targets, kind=HIGHORDER, ld=False, uinf=None,
                 split=True):
    """Rates of the particles `targets` against all particles (wakes
    concatenated): (vel, stretching, diffusion), or (vel, dvort) with
    split=False, each len(targets) x 3.  ld=False: the reference's binary64
    pair arithmetic summed in source order; ld=True: long double."""
    pos, vort, radius, volume = _f(pos), _f(vort), _f(radius), _f(volume)
    n = radius.size
    t = np.ascontiguousarray(np.asarray(targets, dtype=np.uintp))
    ns = t.size
    vel, st = np.zeros((ns, 3)), np.zeros((ns, 3))
    df = np.zeros((ns, 3)) if split else None
    u = None if uinf is None else _f(uinf)
    rc = load().orc_rates_sample(1 if ld else 0, kind, nu, n, _p(pos), _p(vort), _p(radius),
                                 _p(volume), _p(u), ns, t.ctypes.data_as(_sp), _p(vel), _p(st),
                                 _p(df))
    if rc:
        raise ValueError("orc_rates_sample failed")
    return (vel, st, df) if split else (vel, st)


def pack_state(pos, vort, wake_n=None):
    pos, vort = _f(pos), _f(vort)
    n = pos.shape[0]
    w, nw, wp = _wakes(wake_n, n)
    s = np.zeros(6 * n)
    load().orc_pack_state(nw, wp, _p(pos), _p(vort), _p(s))
    return s


def unpack_state(state, wake_n=None):
    state = _f(state)
    n = state.size // 6
    w, nw, wp = _wakes(wake_n, n)
    pos, vort = np.zeros((n, 3)), np.zeros((n, 3))
    load().orc_unpack_state(nw, wp, _p(state), _p(pos), _p(vort))
    return pos, vort


def rk4(state, radius, volume, nu, dt, nsteps, mode=OMP_CLEAN, kind=HIGHORDER, wake_n=None,
        uinf=None, euler=False):
    state = _f(state).copy()
    radius, volume = _f(radius), _f(volume)
    w, nw, wp = _wakes(wake_n, radius.size)
    u = None if uinf is None else _f(uinf)
    fn = load().orc_euler if euler else load().orc_rk4
    rc = fn(mode, kind, nu, nw, wp, _p(radius), _p(volume), _p(u), _p(state), dt, nsteps)
    if rc:
        raise ValueError("orc_rk4 failed")
    return state


def diagnostics(pos, vort, radius, wake_n=None, ke=True):
    pos, vort, radius = _f(pos), _f(vort), _f(radius)
    w, nw, wp = _wakes(wake_n, radius.size)
    L = load()
    om, li, ai = np.zeros(3), np.zeros(3), np.zeros(3)
    L.orc_total_vorticity(nw, wp, _p(vort), _p(om))
    L.orc_linear_impulse(nw, wp, _p(pos), _p(vort), _p(li))
    L.orc_angular_impulse(nw, wp, _p(pos), _p(vort), _p(ai))
    out = {"total_vorticity": om, "linear_impulse": li, "angular_impulse": ai}
    if ke:
        out["kinetic_energy"] = L.orc_kinetic_energy(nw, wp, _p(pos), _p(vort), _p(radius))
    return out


def kinetic_energy_cross(pos, vort, radius, n1):
    """The reference's calculateKineticEnergy(W1, W2) with W1 = first n1 particles."""
    pos, vort, radius = _f(pos), _f(vort), _f(radius)
    return load().orc_kinetic_energy_cross(n1, radius.size - n1, _p(pos), _p(vort), _p(radius))


def enstrophy(pos, vort, radius, wake_n=None):
    pos, vort, radius = _f(pos), _f(vort), _f(radius)
    w, nw, wp = _wakes(wake_n, radius.size)
    return load().orc_enstrophy(nw, wp, _p(pos), _p(vort), _p(radius))


def helicity(pos, vort, radius, wake_n=None):
    pos, vort, radius = _f(pos), _f(vort), _f(radius)
    w, nw, wp = _wakes(wake_n, radius.size)
    return load().orc_helicity(nw, wp, _p(pos), _p(vort), _p(radius))


def influence(pos, vort, radius, kind=HIGHORDER, wake_n=None):
    pos, vort, radius = _f(pos), _f(vort), _f(radius)
    n = radius.size
    w, nw, wp = _wakes(wake_n, n)
    out = np.zeros((n, 3))
    load().orc_influence(kind, nw, wp, _p(pos), _p(vort), _p(radius), _p(out))
    return out


def omp_rows(pos, vort, radius, volume, nu, row_begin, stride, nrows, faithful=True,
             kind=HIGHORDER):
    """Time the OpenMP all-pairs loop over a sample of source rows; returns
    (unordered pairs evaluated, seconds)."""
    pos, vort, radius, volume = _f(pos), _f(vort), _f(radius), _f(volume)
    n = radius.size
    vel, dv = np.zeros((n, 3)), np.zeros((n, 3))
    L = load()
    t0 = L.orc_wtime()
    pairs = L.orc_omp_rows(1 if faithful else 0, kind, nu, n, _p(pos), _p(vort), _p(radius),
                           _p(volume), row_begin, stride, nrows, _p(vel), _p(dv))
    return pairs, L.orc_wtime() - t0


def max_threads():
    return load().orc_max_threads()
