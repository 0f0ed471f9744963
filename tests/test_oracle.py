"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests, fixtures or golden vectors, and its C++ cannot
be built here (it needs GSL), so every oracle function is checked against an
independent fact: a rearranged closed form, a textbook kernel, an asymptote,
a symmetry, a conservation law, a scaling law or a finite-difference
identity, each chosen so that a plausible slip (a dropped term, a wrong
exponent, sign, index or transposed operand) fails it.
"""
import math

import numpy as np
import pytest

import oracle as orc
from pawan_amd import generators as gen
from helpers import per_particle_rel

RB = np.array([1e-3, 1e-2, 0.1, 0.5, 1.0, 2.0, 5.0, 10.0, 100.0, 1e3])
SIG = [0.02, 0.065, 1.0]


# ------------------------------------------------------------------ kernels
@pytest.mark.parametrize("sigma", SIG)
def test_highorder_closed_form(sigma):
    """highorder.h:48-60 rearranged: 4 pi rho^3 q = (rb^2+1.5)/(rb^2+1)^2
    - 7.5 rb (rb^2+1)^-7.5 and sigma^3 Z = (15/8pi)(rb^2+1)^-7.5."""
    for rb in RB:
        rho = rb * sigma
        q, F, Z = orc.kernel(rho, sigma)
        u = rb * rb + 1.0
        g = (rb * rb + 1.5) / u ** 2 - 7.5 * rb * u ** -7.5
        assert math.isclose(4 * math.pi * rho ** 3 * q, g, rel_tol=1e-12)
        assert math.isclose(Z * sigma ** 3, 15.0 / (8 * math.pi) * u ** -7.5, rel_tol=1e-13)


def test_highorder_tabulated_values():
    """4 pi rho^3 q at sigma = 1 as tabulated from an independent 40-digit
    evaluation: 1.49 (rb=1e-3), 0.58 (1), 9.95e-3 (10), 1.0e-6 (1e3)."""
    for rb, val, digits in ((1e-3, 1.49, 2), (1.0, 0.58, 2), (10.0, 9.95e-3, 3), (1e3, 1.0e-6, 2)):
        q, _, _ = orc.kernel(rb, 1.0)
        got = 4 * math.pi * rb ** 3 * q
        assert abs(got - val) <= 0.5 * 10 ** (math.floor(math.log10(abs(val))) - digits + 1) + 1e-15


def test_highorder_limits():
    """rho -> 0: q -> 3/(8 pi rho^3) (next term -5 rb relative);
    rho -> inf: q -> sigma^2/(4 pi rho^5)."""
    s = 0.065
    q, _, _ = orc.kernel(1e-6 * s, s)
    assert math.isclose(q * 8 * math.pi * (1e-6 * s) ** 3 / 3, 1.0 - 5e-6, rel_tol=1e-9)
    rho = 1e4 * s
    q, _, _ = orc.kernel(rho, s)
    assert math.isclose(q * 4 * math.pi * rho ** 5 / s ** 2, 1.0, rel_tol=1e-7)


@pytest.mark.parametrize("kind", [orc.HIGHORDER, orc.GAUSSIAN])
def test_kernel_dimensional_scaling(kind):
    """Lengths x lam: q ~ lam^-3, Z ~ lam^-3, F ~ lam^-5 (F divides by rho^2,
    not by (rho/sigma)^2)."""
    for rb in (0.3, 1.0, 4.0):
        q1, F1, Z1 = orc.kernel(rb * 0.1, 0.1, kind)
        for lam in (0.37, 3.0):
            q2, F2, Z2 = orc.kernel(rb * 0.1 * lam, 0.1 * lam, kind)
            assert math.isclose(q2, q1 * lam ** -3, rel_tol=1e-13)
            assert math.isclose(Z2, Z1 * lam ** -3, rel_tol=1e-13)
            assert math.isclose(F2, F1 * lam ** -5, rel_tol=1e-12)


def _gauss_q(rho, s):
    x = rho / s
    g = math.erf(x / math.sqrt(2)) - math.sqrt(2 / math.pi) * x * math.exp(-x * x / 2)
    return g / (4 * math.pi * rho ** 3)


def _gauss_dq(rho, s):
    x = rho / s
    g = math.erf(x / math.sqrt(2)) - math.sqrt(2 / math.pi) * x * math.exp(-x * x / 2)
    gp = math.sqrt(2 / math.pi) * x * x * math.exp(-x * x / 2)
    return gp / s / (4 * math.pi * rho ** 3) - 3 * g / (4 * math.pi * rho ** 4)


@pytest.mark.parametrize("sigma", SIG)
def test_gaussian_textbook(sigma):
    """gaussian.h:20-58 is the Gaussian-core Biot-Savart kernel:
    q = [erf(rb/sqrt2) - sqrt(2/pi) rb e^{-rb^2/2}]/(4 pi rho^3),
    Z = (2 pi)^-1.5 sigma^-3 e^{-rb^2/2}, and F = q'(rho)/rho."""
    for rb in (0.2, 0.5, 1.0, 2.0, 5.0, 20.0):
        rho = rb * sigma
        q, F, Z = orc.kernel(rho, sigma, orc.GAUSSIAN)
        assert math.isclose(q, _gauss_q(rho, sigma), rel_tol=1e-10)
        assert math.isclose(Z, (2 * math.pi) ** -1.5 * sigma ** -3 * math.exp(-rb * rb / 2),
                            rel_tol=1e-13)
        assert math.isclose(F, _gauss_dq(rho, sigma) / rho, rel_tol=1e-8, abs_tol=1e-14 / rho ** 5)
    # finite core: q(0) = (2 pi)^-1.5 / (3 sigma^3)
    q, _, _ = orc.kernel(1e-4 * sigma, sigma, orc.GAUSSIAN)
    assert math.isclose(q, (2 * math.pi) ** -1.5 / (3 * sigma ** 3), rel_tol=1e-6)


def test_long_double_kernel_agrees():
    """The binary64 and long-double evaluations agree wherever the formula is
    well conditioned.  (The Gaussian q and F cancel at small rb: phi/rb and Z
    agree to O(rb^2), and Z - 3q vanishes at leading order, so there the
    reference arithmetic itself keeps only ~4-8 digits.)"""
    for kind, rbs in ((orc.HIGHORDER, RB[1:]), (orc.GAUSSIAN, RB[3:])):
        for rb in rbs:
            a = orc.kernel(rb * 0.05, 0.05, kind)
            L = orc.load()
            import ctypes as C
            q, F, Z = C.c_longdouble(), C.c_longdouble(), C.c_longdouble()
            L.orc_kernel_ld.argtypes = [C.c_int, C.c_longdouble, C.c_longdouble] + [C.POINTER(C.c_longdouble)] * 3
            L.orc_kernel_ld(kind, rb * 0.05, 0.05, C.byref(q), C.byref(F), C.byref(Z))
            for x, y in zip(a, (q.value, F.value, Z.value)):
                assert math.isclose(x, y, rel_tol=1e-11, abs_tol=1e-300)


def test_dnrm2():
    """gsl_blas_dnrm2 (reference BLAS): exact on Pythagorean triples, no
    overflow or underflow on extreme magnitudes."""
    assert orc.dnrm2([3.0, 4.0, 12.0]) == 13.0
    assert orc.dnrm2([0.0, 0.0, 0.0]) == 0.0
    assert math.isclose(orc.dnrm2([1e200, 1e200, 0.0]), math.sqrt(2) * 1e200, rel_tol=1e-15)
    assert math.isclose(orc.dnrm2([1e-200, 0.0, 1e-200]), math.sqrt(2) * 1e-200, rel_tol=1e-15)
    rng = np.random.default_rng(0)
    for _ in range(100):
        x = rng.normal(size=3)
        assert math.isclose(orc.dnrm2(x), float(np.sqrt(np.sum(x * x))), rel_tol=4e-16)


# ------------------------------------------------------------- pair terms
def _pair(kind, nu, s_s, s_t, r_s, r_t, a_s, a_t, v_s, v_t):
    import ctypes as C
    L = orc.load()
    arr = [np.ascontiguousarray(np.asarray(x, dtype=float)) for x in (r_s, r_t, a_s, a_t)]
    drt, dat, drs, das = (np.zeros(3) for _ in range(4))
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    L.orc_pair(kind, nu, s_s, s_t, *[p(a) for a in arr], v_s, v_t, p(drt), p(dat), p(drs), p(das))
    return drt, dat, drs, das


def test_velocity_is_biot_savart_far_field():
    """Gaussian far field: u(x_t) = (1/4pi) a_s x d / rho^3, d = x_t - x_s
    (sign and cross-product order of VELOCITY, interaction_utils.h:52-58)."""
    rng = np.random.default_rng(1)
    s = 0.01
    for _ in range(5):
        r_s, a_s = rng.normal(size=3), rng.normal(size=3)
        r_t = r_s + 20.0 * rng.normal(size=3)
        d = r_t - r_s
        u, _, _, _ = _pair(orc.GAUSSIAN, 0.0, s, s, r_s, r_t, a_s, rng.normal(size=3), 1.0, 1.0)
        ref = np.cross(a_s, d) / (4 * math.pi * np.linalg.norm(d) ** 3)
        assert np.allclose(u, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_stretching_is_transpose_scheme():
    """With a self-consistent kernel (Gaussian: F = q'/rho) the stretching
    term q (a_t x a_s) + F (d.(a_t x a_s)) d equals (grad u_s(x_t))^T a_t,
    where u_s is the velocity the source induces; checked against central
    differences of the oracle's own velocity.  Swapping the operands (the
    classical scheme grad(u) a_t) fails this at O(1)."""
    rng = np.random.default_rng(2)
    s = 0.3
    for _ in range(6):
        r_s, a_s, a_t = rng.normal(size=3), rng.normal(size=3), rng.normal(size=3)
        r_t = r_s + rng.normal(size=3) * 0.4
        _, da, _, _ = _pair(orc.GAUSSIAN, 0.0, s, s, r_s, r_t, a_s, a_t, 1.0, 1.0)
        h = 1e-5
        G = np.zeros((3, 3))  # G[k, m] = d u_k / d x_m
        for m in range(3):
            e = np.zeros(3)
            e[m] = h
            up, _, _, _ = _pair(orc.GAUSSIAN, 0.0, s, s, r_s, r_t + e, a_s, a_t, 1.0, 1.0)
            um, _, _, _ = _pair(orc.GAUSSIAN, 0.0, s, s, r_s, r_t - e, a_s, a_t, 1.0, 1.0)
            G[:, m] = (up - um) / (2 * h)
        ref = G.T @ a_t
        assert np.allclose(da, ref, rtol=1e-6, atol=1e-8 * np.abs(ref).max())
        assert not np.allclose(da, G @ a_t, rtol=1e-3)


def test_pair_antisymmetry_and_mirror():
    """Exchanging target and source negates the vorticity rate; the source's
    velocity part is the mirrored formula (interaction_utils.h:388-401)."""
    rng = np.random.default_rng(3)
    for kind in (orc.HIGHORDER, orc.GAUSSIAN):
        r_s, r_t, a_s, a_t = (rng.normal(size=3) for _ in range(4))
        s_s, s_t, v_s, v_t = 0.3, 0.4, 0.02, 0.05
        u_t, da_t, u_s, da_s = _pair(kind, 0.01, s_s, s_t, r_s, r_t, a_s, a_t, v_s, v_t)
        u_t2, da_t2, u_s2, da_s2 = _pair(kind, 0.01, s_t, s_s, r_t, r_s, a_t, a_s, v_t, v_s)
        assert np.allclose(da_t, -da_t2, rtol=1e-13, atol=1e-16)
        assert np.allclose(da_s, da_t2, rtol=1e-13, atol=1e-16)
        assert np.allclose(u_s, u_t2, rtol=1e-13, atol=1e-16)


def test_pse_diffusion_is_laplacian():
    """DIFFUSION (interaction_utils.h:253-286) with the normalised Gaussian
    core is particle-strength exchange: on a lattice of spacing h with
    alpha = omega h^3, v = h^3 and omega = z cos(k x), the rate at an interior
    particle is nu v (-(2/sigma^2)(1 - exp(-k^2 sigma^2/2))) omega, the exact
    PSE of cos (-> nu v laplacian(omega) as sigma -> 0)."""
    h, sigma, nu, k = 0.5, 1.0, 0.7, 0.6
    m = 14  # lattice half-width in cells (7 sigma)
    g = np.arange(-m, m + 1) * h
    X, Y, Zc = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Zc.ravel()], axis=1)
    v = h ** 3
    om = np.cos(k * P[:, 0])
    A = np.zeros_like(P)
    A[:, 2] = om * v
    c = len(P) // 2  # the centre particle
    tot = np.zeros(3)
    for j in range(len(P)):
        if j == c:
            continue
        _, da1, _, _ = _pair(orc.GAUSSIAN, nu, sigma, sigma, P[j], P[c], A[j], A[c], v, v)
        _, da0, _, _ = _pair(orc.GAUSSIAN, 0.0, sigma, sigma, P[j], P[c], A[j], A[c], v, v)
        tot += da1 - da0
    expect = nu * v * (-(2 / sigma ** 2) * (1 - math.exp(-k * k * sigma * sigma / 2))) * om[c]
    assert abs(tot[0]) < 1e-14 and abs(tot[1]) < 1e-14
    assert math.isclose(tot[2], expect, rel_tol=1e-6)


# --------------------------------------------------------- all-pairs loops
def _random_system(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3)) * spread
    vort = rng.normal(size=(n, 3)) * 1e-2
    rad = 0.2 + 0.1 * rng.random(n)
    vol = 1e-3 * (1 + rng.random(n))
    return pos, vort, rad, vol


@pytest.mark.parametrize("kind", [orc.HIGHORDER, orc.GAUSSIAN])
def test_loop_modes_agree(kind):
    """Serial symmetric (interaction.cpp), OpenMP faithful and clean
    (parallel.cpp) and an independent one-sided long-double sum."""
    pos, vort, rad, vol = _random_system(97, 4)
    ref = orc.rates(pos, vort, rad, vol, 0.01, orc.ONESIDED_LD, kind)
    for mode in (orc.SERIAL, orc.OMP_FAITHFUL, orc.OMP_CLEAN):
        v, d = orc.rates(pos, vort, rad, vol, 0.01, mode, kind)
        assert per_particle_rel(v, ref[0]) < 1e-12
        assert per_particle_rel(d, ref[1]) < 1e-12


def test_conservation_and_single_particle():
    pos, vort, rad, vol = _random_system(64, 5)
    v, d = orc.rates(pos, vort, rad, vol, 0.01)
    assert np.abs(d.sum(axis=0)).max() < 1e-14 * np.abs(d).sum()
    v1, d1 = orc.rates(pos[:1], vort[:1], rad[:1], vol[:1], 0.01)
    assert np.all(v1 == 0) and np.all(d1 == 0)


def test_translation_rotation_permutation():
    pos, vort, rad, vol = _random_system(50, 6)
    v, d = orc.rates(pos, vort, rad, vol, 0.01)
    vt, dt = orc.rates(pos + np.array([3.0, -2.0, 5.0]), vort, rad, vol, 0.01)
    assert per_particle_rel(vt, v) < 1e-11 and per_particle_rel(dt, d) < 1e-11
    R = gen.rotation_from_uniforms(0.3, 0.7, 0.1)
    vr, dr = orc.rates(pos @ R.T, vort @ R.T, rad, vol, 0.01)
    assert per_particle_rel(vr, v @ R.T) < 1e-12 and per_particle_rel(dr, d @ R.T) < 1e-12
    p = np.random.default_rng(7).permutation(50)
    vp, dp = orc.rates(pos[p], vort[p], rad[p], vol[p], 0.01)
    assert per_particle_rel(vp, v[p]) < 1e-12 and per_particle_rel(dp, d[p]) < 1e-12


def test_scaling_and_linearity():
    """Lengths x lam, volumes x lam^3, nu / lam: u -> lam^-2 u, da -> lam^-3 da.
    Strengths x c: u -> c u, stretching -> c^2, diffusion -> c."""
    pos, vort, rad, vol = _random_system(40, 8)
    nu = 0.02
    v, d = orc.rates(pos, vort, rad, vol, nu)
    lam = 2.5
    v2, d2 = orc.rates(pos * lam, vort, rad * lam, vol * lam ** 3, nu / lam)
    assert per_particle_rel(v2, v * lam ** -2) < 1e-12 and per_particle_rel(d2, d * lam ** -3) < 1e-12
    c = 3.0
    v0, s0 = orc.rates(pos, vort, rad, vol, 0.0)
    df = d - s0
    vc, dc = orc.rates(pos, vort * c, rad, vol, nu)
    assert per_particle_rel(vc, c * v) < 1e-12
    assert per_particle_rel(dc, c * c * s0 + c * df) < 1e-12


def test_two_wakes_equal_concatenation():
    pos, vort, rad, vol = _random_system(60, 9)
    v, d = orc.rates(pos, vort, rad, vol, 0.01, orc.SERIAL, wake_n=[60])
    for mode in (orc.SERIAL, orc.OMP_CLEAN):
        v2, d2 = orc.rates(pos, vort, rad, vol, 0.01, mode, wake_n=[25, 35])
        assert per_particle_rel(v2, v) < 1e-13 and per_particle_rel(d2, d) < 1e-13


def test_freestream_adds_constant():
    pos, vort, rad, vol = _random_system(20, 10)
    v, d = orc.rates(pos, vort, rad, vol, 0.01)
    u = np.array([0.0, 0.1, -0.05])
    vu, du = orc.rates(pos, vort, rad, vol, 0.01, uinf=u)
    assert np.allclose(vu, v + u, rtol=0, atol=1e-15) and np.array_equal(du, d)


# ------------------------------------------------------------ state layout
def test_state_layout():
    """wake.cpp:87-120: per wake [positions 3n | vorticities 3n]."""
    pos = np.arange(15, dtype=float).reshape(5, 3)
    vort = 100 + np.arange(15, dtype=float).reshape(5, 3)
    s = orc.pack_state(pos, vort, [2, 3])
    expect = np.concatenate([pos[:2].ravel(), vort[:2].ravel(), pos[2:].ravel(), vort[2:].ravel()])
    assert np.array_equal(s, expect)
    p2, v2 = orc.unpack_state(s, [2, 3])
    assert np.array_equal(p2, pos) and np.array_equal(v2, vort)


# --------------------------------------------------------------------- RK4
def test_rk4_on_linear_system():
    """Particles on the z axis with vorticity along z: a_j x d = 0 and
    a_i x a_j = 0, so only diffusion acts and da/dt = M a is linear.  One
    RK4 step must equal (I + hM + (hM)^2/2 + (hM)^3/6 + (hM)^4/24) a
    (rk4.cpp:12-77), one Euler step (I + hM) a (integration.cpp:37-45)."""
    n = 5
    pos = np.zeros((n, 3))
    pos[:, 2] = np.array([0.0, 0.13, 0.29, 0.4, 0.61])
    rad = np.full(n, 0.2)
    vol = np.array([1.0, 1.5, 0.7, 1.2, 0.9]) * 1e-3
    nu = 0.3
    M = np.zeros((n, n))
    for j in range(n):
        a = np.zeros((n, 3))
        a[j, 2] = 1.0
        _, d = orc.rates(pos, a, rad, vol, nu)
        M[:, j] = d[:, 2]
    a0 = np.zeros((n, 3))
    a0[:, 2] = np.array([1.0, -0.5, 0.3, 0.8, -0.2]) * 1e-2
    h = 0.05
    hM = h * M
    P = np.eye(n) + hM + hM @ hM / 2 + hM @ hM @ hM / 6 + hM @ hM @ hM @ hM / 24
    for euler, poly in ((False, P), (True, np.eye(n) + hM)):
        st = orc.rk4(orc.pack_state(pos, a0), rad, vol, nu, h, 1, mode=orc.SERIAL, euler=euler)
        p1, a1 = orc.unpack_state(st)
        assert np.array_equal(p1, pos)
        assert np.allclose(a1[:, 2], poly @ a0[:, 2], rtol=1e-12, atol=1e-18)
    # two steps = polynomial squared
    st = orc.rk4(orc.pack_state(pos, a0), rad, vol, nu, h, 2, mode=orc.SERIAL)
    _, a2 = orc.unpack_state(st)
    assert np.allclose(a2[:, 2], P @ P @ a0[:, 2], rtol=1e-12, atol=1e-18)


# ------------------------------------------------------------- diagnostics
def test_ring_linear_diagnostics_closed_form():
    """Points on a circle of radius R with alpha = Gamma R dpsi e_psi:
    total vorticity 0, linear impulse pi R^2 Gamma z."""
    w = gen.vortex_ring(64, 1, R=1.3, core=0.1, gamma=0.8, jitter=0.0)
    dg = orc.diagnostics(w["pos"], w["vort"], w["radius"], ke=False)
    assert np.abs(dg["total_vorticity"]).max() < 1e-15
    assert np.allclose(dg["linear_impulse"], [0, 0, math.pi * 1.3 ** 2 * 0.8], rtol=1e-14, atol=1e-15)


def test_kinetic_energy_pair_limits():
    """KINETICENERGY (interaction_utils.h:197-240): the self term is the pair
    term at zero separation with equal cores; the far field is the singular
    vortex-particle energy (1/16 pi)[a1.a2 + (a1.e)(a2.e)]/rho."""
    rng = np.random.default_rng(11)
    a1, a2 = rng.normal(size=3), rng.normal(size=3)
    s = 0.2
    # two coincident particles with equal strength: KE = 2[2 self + 2 pair(0)] = 8 self
    pos = np.zeros((2, 3))
    e2 = orc.diagnostics(pos, np.stack([a1, a1]), np.array([s, s]))["kinetic_energy"]
    self_ = a1 @ a1 / (8 * math.pi * s)
    assert math.isclose(e2, 8 * self_, rel_tol=1e-14)
    rho = 1e4 * s
    d = rng.normal(size=3)
    d *= rho / np.linalg.norm(d)
    pos = np.stack([np.zeros(3), d])
    e = orc.diagnostics(pos, np.stack([a1, a2]), np.array([s, s]))["kinetic_energy"]
    pair = (e - 2 * (a1 @ a1 + a2 @ a2) / (8 * math.pi * s)) / 4
    dh = d / rho
    sing = (a1 @ a2 + (a1 @ dh) * (a2 @ dh)) / (16 * math.pi * rho)
    assert math.isclose(pair, sing, rel_tol=1e-6)


def test_kinetic_energy_multiwake_consistent():
    pos, vort, rad, _ = _random_system(30, 12)
    e1 = orc.diagnostics(pos, vort, rad, wake_n=[30])["kinetic_energy"]
    e2 = orc.diagnostics(pos, vort, rad, wake_n=[12, 18])["kinetic_energy"]
    assert math.isclose(e1, e2, rel_tol=1e-13)
    # the reference's routines: KE(W1) + KE(W2) + 2 KE(W1, W2) is the union's KE
    ka = orc.diagnostics(pos[:12], vort[:12], rad[:12])["kinetic_energy"]
    kb = orc.diagnostics(pos[12:], vort[12:], rad[12:])["kinetic_energy"]
    kc = orc.kinetic_energy_cross(pos, vort, rad, 12)
    assert math.isclose(ka + kb + 2 * kc, e1, rel_tol=1e-13)


def test_influence_matches_kernel_sum():
    """influence() (interaction.cpp:110-155): k_i = sum_j Z(rho_ij) a_j with
    the self term Z(0, s_i) a_i; checked against a direct one-sided sum."""
    pos, vort, rad, _ = _random_system(25, 13)
    f = orc.influence(pos, vort, rad)
    ref = np.zeros_like(f)
    for i in range(25):
        for j in range(25):
            sig = rad[i] if i == j else math.sqrt(0.5 * (rad[i] ** 2 + rad[j] ** 2))
            rho = float(np.linalg.norm(pos[i] - pos[j]))
            ref[i] += 15 / (8 * math.pi) * ((rho / sig) ** 2 + 1) ** -7.5 / sig ** 3 * vort[j]
    assert per_particle_rel(f, ref) < 1e-12


def test_enstrophy_helicity_symmetry():
    """Helicity is a pseudoscalar (odd under reflection), enstrophy a scalar
    (even); both are invariant under rotation and wake splitting."""
    pos, vort, rad, _ = _random_system(30, 14)
    H = orc.helicity(pos, vort, rad)
    E = orc.enstrophy(pos, vort, rad)
    Rf = np.diag([1.0, 1.0, -1.0])
    # vorticity is an axial vector: under a reflection it picks up det = -1
    Hm = orc.helicity(pos @ Rf, -(vort @ Rf), rad)
    Em = orc.enstrophy(pos @ Rf, -(vort @ Rf), rad)
    assert math.isclose(Hm, -H, rel_tol=1e-12, abs_tol=1e-18)
    assert math.isclose(Em, E, rel_tol=1e-12)
    R = gen.rotation_from_uniforms(0.2, 0.4, 0.9)
    assert math.isclose(orc.helicity(pos @ R.T, vort @ R.T, rad), H, rel_tol=1e-10, abs_tol=1e-18)
    assert math.isclose(orc.enstrophy(pos, vort, rad, wake_n=[10, 20]), E, rel_tol=1e-12)


def test_angular_impulse_offset_ring_closed_form():
    """calculateAngularImpulse (interaction.cpp:261-275), A = (1/3) sum x x (x x a),
    on a planar ring of N points at radius R about centre c with
    a = Gamma R dpsi e_psi: the sums over psi of the degree-2 trigonometric
    terms are exact for N >= 3, giving A = pi R^2 Gamma (c_y, -c_x, 0) = c x I
    (I = pi R^2 Gamma z, the linear impulse); c_z drops out."""
    for c in ((0.0, 0.0, 0.0), (0.4, -0.7, 0.25), (-1.5, 0.3, -2.0)):
        w = gen.vortex_ring(48, 1, R=1.3, core=0.1, gamma=0.8, jitter=0.0, center=c)
        dg = orc.diagnostics(w["pos"], w["vort"], w["radius"], ke=False)
        I = math.pi * 1.3 ** 2 * 0.8
        want = np.array([I * c[1], -I * c[0], 0.0])
        assert np.allclose(dg["angular_impulse"], want, rtol=1e-13, atol=1e-14)
        assert np.allclose(np.cross(c, dg["linear_impulse"]), want, rtol=1e-13, atol=1e-14)


def test_enstrophy_self_and_coincident_pair():
    """ENST (highorder.h:69-86) with ENSTROPHY (interaction_utils.h:107-156):
    one particle gives 5 |a|^2 / (4 pi s^3); at zero separation with equal
    cores the pair factor F1 is the same 5/(4 pi s^3) (F2 = 0), so two
    coincident particles have the enstrophy of one particle carrying a1 + a2
    (calculateEnstrophy, interaction.cpp:277-298, doubles each pair)."""
    rng = np.random.default_rng(21)
    a1, a2 = rng.normal(size=3), rng.normal(size=3)
    for s in (0.02, 0.3, 1.7):
        e1 = orc.enstrophy(np.zeros((1, 3)), a1[None, :], np.array([s]))
        assert math.isclose(e1, 5.0 * (a1 @ a1) / (4 * math.pi * s ** 3), rel_tol=1e-14)
        pos = np.full((2, 3), 0.37)
        e2 = orc.enstrophy(pos, np.stack([a1, a2]), np.array([s, s]))
        a = a1 + a2
        assert math.isclose(e2, 5.0 * (a @ a) / (4 * math.pi * s ** 3), rel_tol=1e-13)


def test_enstrophy_pair_far_field_and_unit_separation():
    """ENST pair factors: far field (rho >> sigma) F1 -> -1/(4 pi rho^3) and
    F2 -> -15/(8 pi rho^5) (relative corrections O(sigma^2/rho^2)); at
    rho = sigma exactly, F1 = 1/(8 pi 2^3.5 sigma^3) (10 - 2 - 7 = 1) and
    F2 = -15/(8 pi 2^3.5 sigma^5).  E(pair) - self terms = 2 (F1 a1.a2 +
    F2 (a1.x12)(a2.x12))."""
    rng = np.random.default_rng(22)
    a1, a2 = rng.normal(size=3), rng.normal(size=3)
    s1, s2 = 0.05, 0.09
    sig = math.sqrt(0.5 * (s1 * s1 + s2 * s2))
    selfs = 5.0 * (a1 @ a1 / s1 ** 3 + a2 @ a2 / s2 ** 3) / (4 * math.pi)
    for rho, F1, F2, tol in ((1e4 * sig, -1 / (4 * math.pi * (1e4 * sig) ** 3),
                              -15 / (8 * math.pi * (1e4 * sig) ** 5), 1e-6),
                             (sig, 1 / (8 * math.pi * 2 ** 3.5 * sig ** 3),
                              -15 / (8 * math.pi * 2 ** 3.5 * sig ** 5), 1e-12)):
        d = rng.normal(size=3)
        d *= rho / np.linalg.norm(d)
        x1 = np.array([0.2, -0.1, 0.3])
        pos = np.stack([x1, x1 + d])
        e = orc.enstrophy(pos, np.stack([a1, a2]), np.array([s1, s2]))
        want = 2.0 * (F1 * (a1 @ a2) + F2 * (a1 @ d) * (a2 @ d))
        assert math.isclose(e - selfs, want, rel_tol=tol, abs_tol=1e-12 * abs(selfs)), (rho, e - selfs, want)


def test_helicity_two_particles_and_velocity_identity():
    """HELICITY (interaction_utils.h:167-190): H = 2 sum_{I<J} q(rho, sigma)
    x_IJ.(a_I x a_J) with x_IJ = x_J - x_I (calculateHelicity,
    interaction.cpp:317-332).  Hand case: x12 = (0, 0, rho), a1 = x, a2 = y
    gives H = 2 q rho > 0.  In general, since u_i = sum_j q (a_j x (x_i - x_j))
    (VELOCITY, interaction_utils.h:52-58), sum_i a_i.u_i = 2 sum_{I<J} q
    (x_I - x_J).(a_I x a_J): the reference's helicity is minus the sum of
    a_i.u_i over its own (pinned) velocities."""
    s, rho = 0.1, 0.07
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, rho]])
    vort = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    q, _, _ = orc.kernel(rho, s)
    H = orc.helicity(pos, vort, np.array([s, s]))
    assert q > 0 and math.isclose(H, 2 * q * rho, rel_tol=1e-14)
    # exchanging the particles' labels leaves it unchanged
    assert math.isclose(orc.helicity(pos[::-1], vort[::-1], np.array([s, s])), H, rel_tol=1e-14)
    # mirror image (z -> -z, axial vectors pick up det = -1): sign flips
    assert math.isclose(orc.helicity(pos * [1, 1, -1], -(vort * [1, 1, -1]), np.array([s, s])),
                        -H, rel_tol=1e-14)
    for n, seed in ((2, 30), (17, 31), (60, 32)):
        p, a, rad, vol = _random_system(n, seed)
        v, _ = orc.rates(p, a, rad, vol, 0.0, orc.ONESIDED_LD)
        H = orc.helicity(p, a, rad)
        want = -float(np.sum(a * v))
        assert math.isclose(H, want, rel_tol=1e-11, abs_tol=1e-14 * float(np.sum(np.abs(a * v)))), (n, H, want)


# ------------------------------------------------------------ sampled rates
def test_rates_sample_matches_full_oracle():
    """orc_rates_sample (targets against every particle, the reference's
    binary64 pair arithmetic summed in source order) equals the symmetric
    serial sweep (interaction.cpp:95-108) and the OpenMP sweep to summation
    order, with several wakes concatenated; the long-double form equals the
    one-sided long-double sums exactly (same arithmetic, same order); the
    stretching and diffusion split adds up to the total; a subset returns the
    same rows bit for bit; the freestream is added to velocities only."""
    a = gen.vortex_ring(24, 16, R=1.0, core=0.1, seed=41)
    b = gen.vortex_ring(20, 9, R=0.8, core=0.1, seed=42, center=(0, 0, 0.3))
    w = {k: np.concatenate([a[k], b[k]]) for k in ("pos", "vort", "radius", "volume")}
    n = w["radius"].size
    nu = 1e-3
    args = (w["pos"], w["vort"], w["radius"], w["volume"])
    rv, rd = orc.rates(*args, nu, mode=orc.SERIAL, wake_n=[a["n"], b["n"]])
    ov, od = orc.rates(*args, nu, mode=orc.OMP_FAITHFUL, wake_n=[a["n"], b["n"]])
    allt = np.arange(n)
    v, st, df = orc.rates_sample(*args, nu, allt)
    assert per_particle_rel(v, rv) < 1e-13 and per_particle_rel(st + df, rd) < 1e-13
    assert per_particle_rel(v, ov) < 1e-13 and per_particle_rel(st + df, od) < 1e-13
    v2, d2 = orc.rates_sample(*args, nu, allt, split=False)
    assert np.array_equal(v2, v) and per_particle_rel(d2, st + df) < 1e-12
    _, st0, df0 = orc.rates_sample(*args, 0.0, allt)
    assert np.array_equal(st0, st) and np.all(df0 == 0.0)
    lv, ls, ld = orc.rates_sample(*args, nu, allt, ld=True)
    sv, ss, sd = orc.rates_split(*args, nu)
    assert np.array_equal(lv, sv) and np.array_equal(ls, ss) and np.array_equal(ld, sd)
    assert per_particle_rel(v, lv) < 1e-13 and per_particle_rel(st, ls) < 1e-11
    sub = np.array([n - 1, 0, 17, 300, 301, 5])
    vs, ss_, ds = orc.rates_sample(*args, nu, sub)
    assert np.array_equal(vs, v[sub]) and np.array_equal(ss_, st[sub]) and np.array_equal(ds, df[sub])
    u = np.array([0.01, -0.02, 0.05])
    vu, su, du = orc.rates_sample(*args, nu, sub, uinf=u)
    assert np.array_equal(vu, vs + u) and np.array_equal(su, ss_) and np.array_equal(du, ds)
    with pytest.raises(ValueError):
        orc.rates_sample(*args, nu, [n])


def test_rates_sample_gaussian_kernel():
    pos, vort, rad, vol = _random_system(40, 43)
    v, d = orc.rates(pos, vort, rad, vol, 0.01, orc.SERIAL, orc.GAUSSIAN)
    sv, sst, sdf = orc.rates_sample(pos, vort, rad, vol, 0.01, np.arange(40), kind=orc.GAUSSIAN)
    assert per_particle_rel(sv, v) < 1e-13 and per_particle_rel(sst + sdf, d) < 1e-13


def test_sample_targets_cover_edges():
    from helpers import sample_targets
    t = sample_targets(262144, 512)
    assert t[0] == 0 and t[-1] == 262143 and np.all(np.diff(t) > 0)
    for e in (63, 64, 255, 256, 511, 512, 16383, 16384):
        assert e in t
    # every stratum holds a pick
    assert np.all(np.histogram(t, bins=512, range=(0, 262144))[0] >= 1)
