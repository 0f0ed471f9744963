"""Seeded particle generators for the vortex-ring and rotor-wake workloads.

pawan's own generators cannot produce these layouts (``__vring`` makes
(2n+1)^2 particles per cross-section, src/wake/vring.cpp:55; ``__ring`` makes
(2n-1)^2, src/wake/ring.cpp:90), so these are this package's own.  They keep
the reference's conventions:

* azimuthal unit vector e_psi = (-sin psi, cos psi, 0)  (ring.cpp:32-33)
* strength = circulation x arc length along e_psi        (ring.cpp:25)
* isolated-particle volume 4/3 pi s^3                    (ring.cpp:26)

Randomness comes only from SplitMix64 (Steele, Lea & Flood 2014) in counter
form, so the same seed gives the same particles on any platform:
``x_k = mix(seed + (k+1) * 0x9E3779B97F4A7C15)``, ``u_k = (x_k >> 11) * 2^-53``.

Every generator returns a dict with float64 arrays ``pos`` (N,3), ``vort``
(N,3), ``radius`` (N,), ``volume`` (N,) in the reference's storage layout
(gsl_matrix N x 3 row-major), plus the parameters that define the workload.
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_MASK = (1 << 64) - 1


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """``count`` outputs of the SplitMix64 stream started at ``seed``,
    skipping the first ``offset`` outputs."""
    k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _MASK) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def cross_section(n_cs: int, core: float):
    """Particle offsets (xi radial, eta axial), spacing h and relative
    strengths over a core of radius ``core``.

    A perfect square n_cs = m^2 is laid out as the cell centres of an m x m
    grid over [-core, core]^2 (h = 2 core/m); any other count as a Vogel
    sunflower disk r_k = core sqrt((k+1/2)/n_cs), theta_k = k * golden angle
    (h = core sqrt(pi/n_cs), the side of an equal-area cell).  Strengths
    follow exp(-r^2/core^2), normalised to sum to 1."""
    m = int(round(math.sqrt(n_cs)))
    if m * m == n_cs:
        h = 2.0 * core / m
        c = -core + h * (np.arange(m) + 0.5)
        xi, eta = np.meshgrid(c, c, indexing="ij")
        xi, eta = xi.ravel(), eta.ravel()
    else:
        h = core * math.sqrt(math.pi / n_cs)
        k = np.arange(n_cs, dtype=np.float64)
        r = core * np.sqrt((k + 0.5) / n_cs)
        th = k * (math.pi * (3.0 - math.sqrt(5.0)))
        xi, eta = r * np.cos(th), r * np.sin(th)
    w = np.exp(-(xi * xi + eta * eta) / (core * core))
    return xi, eta, h, w / w.sum()


def rotation_from_uniforms(u1: float, u2: float, u3: float) -> np.ndarray:
    """Uniformly distributed rotation (Shoemake's unit quaternion method)."""
    a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
    qx, qy = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2)
    qz, qw = b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
    return np.array([
        [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
        [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
        [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)],
    ])


def vortex_ring(n_azi: int, n_cs: int, R: float = 1.0, core: float = 0.1, gamma: float = 1.0,
                seed: int = 0, center=(0.0, 0.0, 0.0), rotation=None, jitter: float = 0.01,
                s_factor: float = 1.3) -> dict:
    """A thick vortex ring about the z axis (then rotated and translated).

    n_azi azimuthal stations x n_cs cross-section particles, ordered station
    by station.  Particle p of station a sits at ((R+xi) cos psi, (R+xi) sin psi,
    eta) with psi = 2 pi a/n_azi, carries alpha = Gamma_p (R+xi) dpsi e_psi,
    core radius s = s_factor h and volume h^2 (R+xi) dpsi.  Each coordinate is
    jittered uniformly by +-jitter*h from the SplitMix64 stream of ``seed``."""
    xi, eta, h, w = cross_section(n_cs, core)
    dpsi = 2.0 * math.pi / n_azi
    psi = dpsi * np.arange(n_azi)
    rho = np.broadcast_to(R + xi[None, :], (n_azi, n_cs))
    c, s = np.cos(psi)[:, None], np.sin(psi)[:, None]
    pos = np.stack([rho * c, rho * s, np.broadcast_to(eta[None, :], rho.shape)], axis=-1)
    mag = gamma * w[None, :] * rho * dpsi
    vort = np.stack([-mag * s, mag * c, np.zeros_like(mag)], axis=-1)
    n = n_azi * n_cs
    pos = pos.reshape(n, 3)
    vort = vort.reshape(n, 3)
    pos = pos + jitter * h * (2.0 * uniform(seed, 3 * n).reshape(n, 3) - 1.0)
    if rotation is not None:
        rot = np.asarray(rotation, dtype=np.float64)
        pos = pos @ rot.T
        vort = vort @ rot.T
    pos = pos + np.asarray(center, dtype=np.float64)
    return {
        "pos": np.ascontiguousarray(pos), "vort": np.ascontiguousarray(vort),
        "radius": np.full(n, s_factor * h), "volume": (h * h * rho * dpsi).reshape(n).copy(),
        "n": n, "h": h, "R": R, "core": core, "gamma": gamma, "seed": seed,
    }


def rotor_wake(n_shed: int, blades: int = 4, stations: int = 64, R: float = 1.0,
               gamma0: float = 1.0, dpsi: float = 0.1, dt: float = 0.1, inflow: float = 0.05,
               core: float = 0.02) -> dict:
    """Frozen synthetic rotor wake after ``n_shed`` shedding steps.

    Blade b is at azimuth 2 pi b/blades + dpsi*m at step m; every step each
    blade sheds one particle per station r_k = (k + 1/2) R/stations (never
    r = 0, where the blades meet).  A particle shed at step m has been carried
    down the axis by the inflow for (n_shed - 1 - m) steps: z = -inflow dt
    (n_shed - 1 - m).  Strength along e_psi of the shedding azimuth over the
    arc r_k dpsi swept in one step: stations k < stations-1 trail the drop of
    the elliptic bound circulation Gamma(r) = gamma0 sqrt(1 - (r/R)^2) across
    their span cell; the tip station carries a concentrated tip vortex of
    strength gamma0.  Core radius ``core``, volume 4/3 pi core^3.
    Ordering: step-major, then blade, then station (N = n_shed*blades*stations)."""
    k = np.arange(stations, dtype=np.float64)
    dr = R / stations
    r = (k + 0.5) * dr

    def circ(x):
        x = np.clip(x / R, 0.0, 1.0)
        return gamma0 * np.sqrt(1.0 - x * x)

    g = circ(r - 0.5 * dr) - circ(r + 0.5 * dr)
    g[-1] = gamma0
    m = np.arange(n_shed, dtype=np.float64)
    b = np.arange(blades, dtype=np.float64)
    psi = 2.0 * math.pi * b[None, :] / blades + dpsi * m[:, None]        # (n_shed, blades)
    z = -inflow * dt * (n_shed - 1 - m)
    c, s = np.cos(psi)[..., None], np.sin(psi)[..., None]                 # (n_shed, blades, 1)
    pos = np.stack([r * c, r * s, np.broadcast_to(z[:, None, None], c.shape[:2] + (stations,))],
                   axis=-1)
    mag = (g * r * dpsi)[None, None, :]
    vort = np.stack([-mag * s, mag * c, np.zeros(c.shape[:2] + (stations,))], axis=-1)
    n = n_shed * blades * stations
    return {
        "pos": np.ascontiguousarray(pos.reshape(n, 3)),
        "vort": np.ascontiguousarray(vort.reshape(n, 3)),
        "radius": np.full(n, core), "volume": np.full(n, 4.0 / 3.0 * math.pi * core ** 3),
        "n": n, "core": core, "inflow": inflow, "dt": dt,
    }


def rotor_emission(step: int, blades: int = 4, stations: int = 64, R: float = 1.0,
                   gamma0: float = 1.0, dpsi: float = 0.1, core: float = 0.02) -> dict:
    """The particles every blade sheds at shedding step ``step`` (z = 0): the
    newest row of ``rotor_wake``."""
    w = rotor_wake(1, blades, stations, R, gamma0, dpsi, 0.0, 0.0, core)
    psi0 = dpsi * step
    c, s = math.cos(psi0), math.sin(psi0)
    rot = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    w["pos"] = np.ascontiguousarray(w["pos"] @ rot.T)
    w["vort"] = np.ascontiguousarray(w["vort"] @ rot.T)
    return w


# ---------------------------------------------------------------- workloads
def config_single_ring(seed: int = 0) -> dict:
    """Single vortex ring, plain FP64: R=1, core 0.1, circulation 1, 128 x 16 =
    2,048 particles, s = 1.3 h, +-1% jitter, nu = 1e-3, RK4 dt = 0.01 x 100."""
    w = vortex_ring(128, 16, R=1.0, core=0.1, gamma=1.0, seed=seed)
    w.update(nu=1e-3, dt=0.01, steps=100, wakes=[w["n"]])
    return w


def config_leapfrog(seed: int = 0, n_azi: int = 256, n_cs: int = 128) -> dict:
    """Two coaxial rings R = 1 at axial separation 0.5, circulation 1 each,
    core 0.08, 256 x 128 per ring (65,536 particles), nu = 5e-4, RK4
    dt = 0.005 x 400.  Two wakes, as __interaction(W1, W2)."""
    a = vortex_ring(n_azi, n_cs, R=1.0, core=0.08, gamma=1.0, seed=seed)
    b = vortex_ring(n_azi, n_cs, R=1.0, core=0.08, gamma=1.0, seed=seed + 1, center=(0, 0, 0.5))
    out = {k: np.concatenate([a[k], b[k]]) for k in ("pos", "vort", "radius", "volume")}
    out.update(n=a["n"] + b["n"], wakes=[a["n"], b["n"]], nu=5e-4, dt=0.005, steps=400)
    return out


def config_rotor_wake(n_shed: int = 1024) -> dict:
    """Synthetic 4-blade rotor wake frozen at n_shed steps (1,024 steps gives
    N = 262,144), core 0.02, nu = 1e-4, freestream -0.05 along z."""
    w = rotor_wake(n_shed)
    w.update(nu=1e-4, uinf=(0.0, 0.0, -0.05), wakes=[w["n"]])
    return w


def config_ring_large(n_azi: int = 2048, n_cs: int = 128, seed: int = 0) -> dict:
    """One ring of 2,048 x 128 = 262,144 particles (R=1, core 0.1, circulation 1)."""
    w = vortex_ring(n_azi, n_cs, R=1.0, core=0.1, gamma=1.0, seed=seed)
    w.update(nu=1e-3, wakes=[w["n"]])
    return w


def ring_cs_for(n: int, R: float = 1.0, core: float = 0.1) -> int:
    """Cross-section particle count for a ring of about n particles whose
    azimuthal spacing 2 pi R n_cs / n matches the cross-section spacing
    core sqrt(pi / n_cs): n_cs = (n core / (2 sqrt(pi) R))^(2/3), at least 8.
    (The single-ring config, 2,048 = 128 x 16, is balanced the same way.)
    A ring with far more stations than cross-section particles overlaps its
    neighbours along the ring many times over and is unstable: with 64
    cross-section particles a 229,568-particle ring diverges within 30 RK4
    steps of dt 0.005 in binary64."""
    n_cs = int(round((n * core / (2.0 * math.sqrt(math.pi) * R)) ** (2.0 / 3.0)))
    return max(8, min(n_cs, max(8, n // 8)))


def batch_jobs(count: int = 256, seed: int = 0, n_min: int = 1024, n_max: int = 65536,
               nu: float = 5e-4, dt: float = 0.005) -> list:
    """Independent rings: N log-uniform in [n_min, n_max], R ~ U[0.5, 2],
    core/R ~ U[0.05, 0.15], circulation ~ U[0.5, 2], uniformly random
    orientation.  N is rounded to n_azi x n_cs (n_cs from ``ring_cs_for`` for
    the job's R and core)."""
    jobs = []
    u = uniform(seed, 7 * count).reshape(count, 7)
    for k in range(count):
        n_target = n_min * (n_max / n_min) ** u[k, 0]
        R = 0.5 + 1.5 * u[k, 1]
        core = R * (0.05 + 0.10 * u[k, 2])
        n_cs = ring_cs_for(int(n_target), R, core)
        n_azi = max(8, int(round(n_target / n_cs)))
        gamma = 0.5 + 1.5 * u[k, 3]
        rot = rotation_from_uniforms(u[k, 4], u[k, 5], u[k, 6])
        w = vortex_ring(n_azi, n_cs, R=R, core=core, gamma=gamma, seed=seed * 1000003 + k + 1,
                        rotation=rot)
        w.update(nu=nu, dt=dt)
        jobs.append(w)
    return jobs


def checksum(w: dict) -> dict:
    """Size-independent fingerprints of a generated particle set."""
    out = {}
    for k in ("pos", "vort", "radius", "volume"):
        a = np.asarray(w[k], dtype=np.float64)
        out[k] = [float(a.sum()), float(np.abs(a).sum()), float((a * a).sum())]
    return out


def serving_stream(count: int, seed: int = 1, rate_hz: float = 1.0, n_lo: int = 2048,
                   n_hi: int = 262144, s_lo: int = 50, s_hi: int = 500,
                   rotor_dt: float = 1e-5) -> list:
    """Jobs of the serving workload, in arrival order: Poisson arrivals at
    rate_hz (arrival_ms), each a vortex ring (R = 1, core 0.1, circulation 1,
    nu = 1e-3, dt = 0.005) or a frozen rotor wake (nu = 1e-4, freestream 0.05
    along -z, dt = rotor_dt), 50/50; N log-uniform in [n_lo, n_hi] (a ring
    is rounded to n_azi x n_cs, a rotor wake to whole shedding rows of 256);
    steps uniform in [s_lo, s_hi]; precision binary64 or compensated binary32
    50/50 (0 / 1).  Job k draws five uniforms from the SplitMix64 stream of
    ``seed``."""
    u = uniform(seed, 5 * count).reshape(count, 5)
    jobs, t = [], 0.0
    for k in range(count):
        a = u[k]
        n = int(n_lo * (n_hi / n_lo) ** a[0])
        steps = int(s_lo + (s_hi - s_lo) * a[1])
        prec = 0 if a[2] < 0.5 else 1
        if a[3] < 0.5:
            n_cs = ring_cs_for(n, 1.0, 0.1)
            w = vortex_ring(max(8, n // n_cs), n_cs, R=1.0, core=0.1, seed=seed + k)
            job = dict(kind="ring", w=w, nu=1e-3, dt=0.005, uinf=(0.0, 0.0, 0.0))
        else:
            w = rotor_wake(max(1, n // 256))
            job = dict(kind="rotor", w=w, nu=1e-4, dt=rotor_dt, uinf=(0.0, 0.0, -0.05))
        job.update(n=int(w["n"]), steps=steps, precision=prec, arrival_ms=t)
        jobs.append(job)
        t += -math.log(1.0 - a[4]) / rate_hz * 1000.0
    return jobs
