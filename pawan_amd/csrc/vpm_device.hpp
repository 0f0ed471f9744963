// vpm_device.hpp -- gfx950 device code for the vortex-particle interaction path.
//
// One ordered pair (target i, source j != i) contributes, with d = x_i - x_j,
// rho = |d|, sigma^2 = (s_i^2 + s_j^2)/2 and t = (rho/sigma)^2 (reference:
// src/interaction/interaction_utils.h:429-481 with KERNEL, highorder.h:48-60):
//   u_i   += q (a_j x d)                                   VELOCITY  :52-58
//   da_i  += q (a_i x a_j) + F (d.(a_i x a_j)) d           VORSTRETCH :69-100
//          + (nu Z/(sigma^2/2)) (v_i a_j - v_j a_i)        DIFFUSION :253-286
// The kernel is evaluated in dimensionless form: with sigma^3 q = Q(t),
// sigma^5 F = F'(t), sigma^3 Z = Zs(t),
//   HIGHORDER: Zs = (1.875/pi) (t+1)^-7.5
//              Q  = [(0.25/pi)(t+1.5)(t+1)^-2 t^-1/2 - Zs] / t
//   GAUSSIAN : Zs = (2 pi)^-1.5 exp(-t/2)
//              Q  = [(0.25/pi) erf(sqrt(t/2)) t^-1/2 - Zs] / t
//   both     : F' = (Zs - 3Q) / t
// The PSE sum is regrouped, sum k (v_i a_j - v_j a_i) = v_i sum k a_j -
// a_i sum k v_j, which saves four operations per pair and loses little in
// binary64 (the two sums cancel by a few digits at most); the binary32 passes
// add close pairs' terms as k (a_j - a_i), k (v_j - v_i), the differences
// formed in binary64, which the same epilogue turns into the same sum.  Far
// pairs take the kernel in a far-field form (below).  The stretching cross
// product of close pairs is kept per pair: a_i x sum q a_j over all sources
// cancels when the vorticities are nearly parallel, as they are along a
// vortex filament (on a rotor wake in binary64 the stretching error grows
// from 1e-13 to 1.4e-11; in binary32 it would be far above 1e-6).  Their
// velocity stays per pair for the same reason (sum q a_j x x_i -
// sum q a_j x x_j cancels by |x| / |d|).  The binary64 one-sided pass
// regroups only far pairs, per 64-source group and about the group's first
// particle, where both ratios are bounded (far groups, by moments, below).
// Each work item (target tile x source chunk) writes its NCOMP partial sums to
// its own slot; a second kernel adds the chunks in a fixed order and applies
// the epilogue, so results never depend on scheduling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "vpm_types.hpp"

namespace pw {

// ------------------------------------------------------------------ math
// Reciprocal square root: hardware estimate (about 2^-24 relative on gfx950)
// refined by one Newton-Raphson step r <- r (1 + (1 - t r^2)/2), which
// squares the relative error (about 2^-47 after one step).
__device__ __forceinline__ double rsq_nr(double t) {
    double r = __builtin_amdgcn_rsq(t);
    double e = __builtin_fma(-t, r * r, 1.0);
    return __builtin_fma(r * 0.5, e, r);
}

// Twice the reciprocal square root, from the same estimate and Newton step
// written as 2r' = r (3 - t r^2): three operations instead of four.  The
// factor 2 is exact; the pair kernels fold it into their constants.
__device__ __forceinline__ double rsq2_nr(double t) {
    const double r = __builtin_amdgcn_rsq(t);
    return r * __builtin_fma(-t, r * r, 3.0);
}

template <int KIND>
__device__ __forceinline__ void kernel64(double t, double rt, double &Q, double &Fp, double &Zs) {
    const double it = rt * rt;
    if (KIND == KIND_HIGHORDER) {
        const double w = rsq_nr(t + 1.0);  // (t+1)^-1/2
        const double iu = w * w;
        const double iu2 = iu * iu;        // (t+1)^-2
        const double w5 = iu2 * w;
        Zs = K_CZ * (w5 * w5 * w5);        // (1.875/pi) (t+1)^-7.5
        // phi/rb - Z = (0.25/pi)(t+1.5)(t+1)^-2 t^-1/2 - Z
        const double P = __builtin_fma(__builtin_fma(t, K_CPHI, 1.5 * K_CPHI), iu2 * rt, -Zs);
        Q = P * it;
    } else {
        Zs = K_CG * exp(-0.5 * t);
        const double phir = K_CPHI * erf(K_SQRT1_2 * (t * rt)) * rt;
        Q = (phir - Zs) * it;
    }
    Fp = __builtin_fma(-3.0, Q, Zs) * it;
}

// HIGHORDER kernel for the binary64 pair pass, scaled: returns kq Q, kq F'
// and kz Zs with kz = 2^15 / (1.875/pi), kq = 4 kz (pair_scale), from
// R = 2 t^-1/2.  Every power-of-two factor is exact and the 1.875/pi of Zs
// moves into the epilogue, which saves three operations per pair.
__device__ __forceinline__ void kernel64_scaled(double t, double R, double &Q, double &Fp,
                                                double &Zs) {
    const double W = rsq2_nr(t + 1.0);  // 2 (t+1)^-1/2
    const double iu = W * W;
    const double iu2 = iu * iu;         // 16 (t+1)^-2
    const double w5 = iu2 * W;          // 32 (t+1)^-2.5
    Zs = w5 * w5 * w5;                  // 2^15 (t+1)^-7.5 = kz Zs
    // kz [(0.25/pi)(t+1.5)(t+1)^-2 t^-1/2 - Zs] with iu2 R = 32 (t+1)^-2 t^-1/2
    const double P = __builtin_fma(__builtin_fma(t, K_HC64, 1.5 * K_HC64), iu2 * R, -Zs);
    const double it = R * R;            // 4 / t
    Q = P * it;                         // kq Q
    Fp = __builtin_fma(-0.75, Q, Zs) * it;  // 4 kz (Zs - 3 Q) / t = kq F'
}

// ---------------------------------------------------------- far field
// For t >= T_FAR64 (rho >= 32 sigma) the same kernel values follow from one
// reciprocal square root and a polynomial, to binary64 rounding:
//   HIGHORDER: (t+1.5)(t+1)^-2 = u g(u), u = 1/t, with
//              g(u) = (1 + 1.5u)(1 + u)^-2 = sum_k (-1)^k (1 - k/2) u^k
//              = 1 - u/2 + u^3/2 - u^4 + 3u^5/2 - 2u^6 + ...
//              truncated after u^4 (the u^5 term is below 1.4e-15 relative);
//              Zs = (1.875/pi)(t+1)^-7.5 is dropped: it changes Q by less
//              than 8 t^-6 relative (7e-18 at t = 1024) and F' by less than
//              2.5 t^-5 (2.2e-15), and it is below 7e-15 of Q there.
//   GAUSSIAN : erf(sqrt(t/2)) = 1 and exp(-t/2) = 0 in binary64 (t/2 >= 512),
//              so Q = (0.25/pi) t^-1.5 and F' = -3 Q / t.
// Checked in multiple precision over t in [1024, 1e13] with the operation
// order below: Q and F' within 2.0e-15 and 1.6e-15 relative (the full form
// evaluated the same way: 1.2e-15 and 1.6e-15).  A pair
// uses the far form when the bounding boxes of its target's 64-particle group
// and its source's 64-particle group are at least sqrt(T_FAR64) sigma apart:
// the choice depends on the particles alone, never on the kernel variant.
constexpr float T_FAR64 = 1024.0f;
// K_HC64-scaled coefficients of g in powers of it = 4/t (u = it/4): G_k =
// 4 K_HC64 c_k / 4^k, c = 1, -1/2, 0, 1/2, -1.
constexpr double K_G0 = 4.0 * K_HC64;
constexpr double K_G1 = -0.5 * K_HC64;
constexpr double K_G3 = K_HC64 / 32.0;
constexpr double K_G4 = -K_HC64 / 64.0;

// Binary32 kernel.  HIGHORDER is scaled by kz = kq = pi/1.875 (pair_scale):
// the 1.875/pi of Zs moves into the epilogue.
template <int KIND>
__device__ __forceinline__ void kernel32(float t, float rt, float &Q, float &Fp, float &Zs) {
    const float it = rt * rt;
    if (KIND == KIND_HIGHORDER) {
        const float w = __builtin_amdgcn_rsqf(t + 1.0f);
        const float iu = w * w;
        const float iu2 = iu * iu;
        const float w5 = iu2 * w;
        Zs = w5 * w5 * w5;
        const float P =
            __builtin_fmaf(__builtin_fmaf(t, (float)K_HC32, (float)(1.5 * K_HC32)), iu2 * rt, -Zs);
        Q = P * it;
    } else {
        Zs = (float)K_CG * __expf(-0.5f * t);
        const float phir = (float)K_CPHI * erff((float)K_SQRT1_2 * (t * rt)) * rt;
        Q = (phir - Zs) * it;
    }
    Fp = __builtin_fmaf(-3.0f, Q, Zs) * it;
}

// Binary32 far field (t >= T_FAR32, rho >= 11.3 sigma), scaled as kernel32:
// HIGHORDER Q = K_HC32 g(u) t^-1/2 u^2 with g truncated after u^3 (the u^4
// term is below 4e-9 relative) and Zs dropped (below 2e-12 relative in Q and
// F'); GAUSSIAN Q = (0.25/pi) t^-1.5 (erf = 1 and exp(-t/2) = 0 in binary32).
// F' = -3 Q / t for both.  DEG4: t >= T_FAR32B (rho >= 8 sigma), g truncated
// after u^4 (the u^5 term is below 1.5e-9 relative, Zs below 1e-8 of Q);
// Gaussian as before (exp(-32) is below 1e-13 of erf there).
constexpr float T_FAR32 = 128.0f;
constexpr float T_FAR32B = 64.0f;
// kernel32_farq gives Q and itm = F'/Q = -3/t instead of F'.
template <int KIND, bool DEG4 = false>
__device__ __forceinline__ void kernel32_farq(float rt, float &Q, float &itm) {
    const float u = rt * rt;  // 1/t
    if (KIND == KIND_HIGHORDER) {
        const float u2 = u * u;
        float g;
        if (DEG4)  // K_HC32 (1 - u/2 + u^3/2 - u^4)
            g = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf((float)-K_HC32, u, (float)(0.5 * K_HC32)), u2,
                                              (float)(-0.5 * K_HC32)),
                               u, (float)K_HC32);
        else  // K_HC32 (1 - u/2 + u^3/2)
            g = __builtin_fmaf(__builtin_fmaf((float)(0.5 * K_HC32), u2, (float)(-0.5 * K_HC32)), u,
                               (float)K_HC32);
        Q = (g * rt) * u2;
    } else {
        Q = ((float)K_CPHI * rt) * u;
    }
    itm = -3.0f * u;
}
template <int KIND>
__device__ __forceinline__ void kernel32_far(float rt, float &Q, float &Fp) {
    float itm;
    kernel32_farq<KIND>(rt, Q, itm);
    Fp = Q * itm;
}

struct Acc64 {
    double v[NCOMP];
};

// One pair in binary64.  UNIFORM: every particle of the job has the same
// core radius; positions are then stored divided by sigma, so t = |d|^2 and
// the powers of sigma are applied once in the epilogue.
template <int KIND, bool UNIFORM>
__device__ __forceinline__ void pair64(double xi, double yi, double zi, double aix, double aiy,
                                       double aiz, double hi, const Rec64 &s, double inv_s2,
                                       Acc64 &a) {
    const double dx = xi - s.x, dy = yi - s.y, dz = zi - s.z;
    const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
    double Q, Fp, Zs;
    if (UNIFORM) {
        if (KIND == KIND_HIGHORDER)
            kernel64_scaled(r2, rsq2_nr(r2), Q, Fp, Zs);
        else
            kernel64<KIND>(r2, rsq_nr(r2), Q, Fp, Zs);
    } else {
        const double isg = rsq_nr(hi + s.h);  // 1/sigma
        const double is2 = isg * isg;
        const double t = r2 * is2;
        if (KIND == KIND_HIGHORDER)
            kernel64_scaled(t, rsq2_nr(t), Q, Fp, Zs);
        else
            kernel64<KIND>(t, rsq_nr(t), Q, Fp, Zs);
        const double s3 = isg * is2, s5 = s3 * is2;
        Q *= s3;
        Fp *= s5;
        Zs *= s5;
    }
    // c = a_j x d
    const double cx = __builtin_fma(s.ay, dz, -s.az * dy);
    const double cy = __builtin_fma(s.az, dx, -s.ax * dz);
    const double cz = __builtin_fma(s.ax, dy, -s.ay * dx);
    a.v[0] = __builtin_fma(Q, cx, a.v[0]);
    a.v[1] = __builtin_fma(Q, cy, a.v[1]);
    a.v[2] = __builtin_fma(Q, cz, a.v[2]);
    // stretching q e + F (d.e) d, e = a_i x a_j, into one sum per component
    const double ex = __builtin_fma(aiy, s.az, -aiz * s.ay);
    const double ey = __builtin_fma(aiz, s.ax, -aix * s.az);
    const double ez = __builtin_fma(aix, s.ay, -aiy * s.ax);
    const double fd = Fp * __builtin_fma(dz, ez, __builtin_fma(dy, ey, dx * ex));
    a.v[3] = __builtin_fma(fd, dx, __builtin_fma(Q, ex, a.v[3]));
    a.v[4] = __builtin_fma(fd, dy, __builtin_fma(Q, ey, a.v[4]));
    a.v[5] = __builtin_fma(fd, dz, __builtin_fma(Q, ez, a.v[5]));
    a.v[6] = __builtin_fma(Zs, s.ax, a.v[6]);
    a.v[7] = __builtin_fma(Zs, s.ay, a.v[7]);
    a.v[8] = __builtin_fma(Zs, s.az, a.v[8]);
    a.v[9] = __builtin_fma(Zs, s.v, a.v[9]);
}

// The far-field kernel at t = r2 (uniform core, t >= T_FAR64), scaled as the
// full form of each kernel is: Q, and itm = F'/Q = -3/t.
template <int KIND>
__device__ __forceinline__ void kernel64_far(double r2, double &Q, double &itm) {
    if (KIND == KIND_HIGHORDER) {
        const double R = rsq2_nr(r2);  // 2 t^-1/2
        const double it = R * R;       // 4 / t
        const double it2 = it * it;
        // G0 + G1 it + it^3 (G3 + G4 it)
        const double g = __builtin_fma(__builtin_fma(K_G4, it, K_G3) * it, it2,
                                       __builtin_fma(K_G1, it, K_G0));
        Q = (g * R) * it2;             // kq Q
        itm = -0.75 * it;              // F'/Q
    } else {
        const double rt = rsq_nr(r2);
        const double it = rt * rt;
        Q = (K_CPHI * rt) * it;
        itm = -3.0 * it;
    }
}

// ------------------------------------------------- far groups, by moments
// The far pairs of a target with one 64-source group G (first particle o,
// x' = x - o) in binary64, regrouped exactly:
//   velocity    sum q a_j x d       = S x x'_i - sum q m_j
//   stretching  sum q a_i x a_j     = a_i x S
//               sum F'(d.e) d,  d.e = d.(a_i x a_j) = a_j.p_i - a_i.m_j
// with S = sum_G q a_j, m_j = a_j x x'_j (per source, formed once while the
// tile is staged) and p_i = x'_i x a_i (per target and group).  Each term is
// at most |x'|/|d| larger than the sum it stands for; a group takes this
// form only when its extent is at most sqrt(T_EXT64) sigma and the pair's
// boxes are at least sqrt(T_FAR64) apart, so that the rounding of the far
// sums grows by a bounded factor (at most 9; 1.5 measured on the rotor wake),
// not with the distance from the origin.
constexpr float T_EXT64 = 16384.0f;  // (128 sigma)^2

// Moment m_j of a staged source about its group's first particle (32 bytes,
// read as 2 x 16 B).
struct __align__(16) Mom64 {
    double x, y, z, pad;
};

// One far pair of a uniform-core job (t >= T_FAR64) in the moment form: the
// far-field kernel and no PSE terms.  pv = p_i; S accumulates q a_j.  35
// arithmetic operations and one v_rsq_f64 against 54 and two for the full
// pair.
template <int KIND>
__device__ __forceinline__ void pair64_mom(double xi, double yi, double zi, double aix, double aiy,
                                           double aiz, const double *pv, const Rec64 &s,
                                           const Mom64 &m, Acc64 &a, double *S) {
    const double dx = xi - s.x, dy = yi - s.y, dz = zi - s.z;
    const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
    double Q, itm;
    kernel64_far<KIND>(r2, Q, itm);
    // d.(a_i x a_j) = a_j.p_i - a_i.m_j
    double tr = __builtin_fma(s.az, pv[2], __builtin_fma(s.ay, pv[1], s.ax * pv[0]));
    tr = __builtin_fma(-aiz, m.z, __builtin_fma(-aiy, m.y, __builtin_fma(-aix, m.x, tr)));
    const double qk = Q * (itm * tr);  // F' (d.e)
    a.v[0] = __builtin_fma(-Q, m.x, a.v[0]);
    a.v[1] = __builtin_fma(-Q, m.y, a.v[1]);
    a.v[2] = __builtin_fma(-Q, m.z, a.v[2]);
    a.v[3] = __builtin_fma(qk, dx, a.v[3]);
    a.v[4] = __builtin_fma(qk, dy, a.v[4]);
    a.v[5] = __builtin_fma(qk, dz, a.v[5]);
    S[0] = __builtin_fma(Q, s.ax, S[0]);
    S[1] = __builtin_fma(Q, s.ay, S[1]);
    S[2] = __builtin_fma(Q, s.az, S[2]);
}

// Smallest and largest of v over the wave's 64 lanes, in every lane.
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float uniform_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
// v of the wave's first active lane, in every lane.
__device__ __forceinline__ double first_lane(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readfirstlane((int)b);
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// Bounding box of the 64 points (x, y, z) held one per lane, in binary32, in
// wave-uniform registers: lo[0..2], hi[0..2].  The box holds the binary64
// points: rounding a coordinate to binary32 moves it by at most 2^-24 of its
// magnitude, and the box is widened by 2^-22 of its corners' (at 2^30 sigma
// from the origin a rounded box could otherwise open a 128-sigma gap between
// touching groups).
__device__ __forceinline__ void wave_box(double x, double y, double z, float *lo, float *hi) {
    const float q[3] = {(float)x, (float)y, (float)z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float l = uniform_f(wave_min(q[c])), h = uniform_f(wave_max(q[c]));
        lo[c] = __builtin_fmaf(-0x1p-22f, fabsf(l), l);
        hi[c] = __builtin_fmaf(0x1p-22f, fabsf(h), h);
    }
}

// Squared distance between two boxes (0 when they overlap).
__device__ __forceinline__ float box_gap2(const float *alo, const float *ahi, const float *blo,
                                          const float *bhi) {
    float d2 = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float g = fmaxf(0.0f, fmaxf(blo[c] - ahi[c], alo[c] - bhi[c]));
        d2 = __builtin_fmaf(g, g, d2);
    }
    return d2;
}

struct Acc32 {
    float v[NCOMP];
};

// One pair in binary32.  Formed in binary64 and rounded: the displacement,
// the stretching cross product a_i x a_j (nearly parallel strengths along a
// filament cancel in it; from binary32 strengths the error would be
// eps32 |a|^2 per pair, which the cancelling sums over sources then expose),
// and the PSE differences a_j - a_i, v_j - v_i (the epilogue's
// v_i S_a - a_i S_v is the same either way; the differences keep the binary32
// rounding of a_j and v_j out of the cancellation of the two sums in a smooth
// core).  vi: the target's volume.
template <int KIND, bool UNIFORM>
__device__ __forceinline__ void pair32(double xi, double yi, double zi, double aix, double aiy,
                                       double aiz, double vi, float hi, const Rec32 &s,
                                       float inv_s2, Acc32 &a) {
    const float dx = (float)(xi - s.x), dy = (float)(yi - s.y), dz = (float)(zi - s.z);
    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    float Q, Fp, Zs;
    if (UNIFORM) {
        kernel32<KIND>(r2, __builtin_amdgcn_rsqf(r2), Q, Fp, Zs);
    } else {
        const float isg = __builtin_amdgcn_rsqf(hi + (float)s.h);
        const float is2 = isg * isg;
        const float t = r2 * is2;
        kernel32<KIND>(t, __builtin_amdgcn_rsqf(t), Q, Fp, Zs);
        const float s3 = isg * is2, s5 = s3 * is2;
        Q *= s3;
        Fp *= s5;
        Zs *= s5;
    }
    const float cx = __builtin_fmaf(s.fy, dz, -s.fz * dy);
    const float cy = __builtin_fmaf(s.fz, dx, -s.fx * dz);
    const float cz = __builtin_fmaf(s.fx, dy, -s.fy * dx);
    a.v[0] = __builtin_fmaf(Q, cx, a.v[0]);
    a.v[1] = __builtin_fmaf(Q, cy, a.v[1]);
    a.v[2] = __builtin_fmaf(Q, cz, a.v[2]);
    const float ex = (float)__builtin_fma(aiy, s.az, -aiz * s.ay);
    const float ey = (float)__builtin_fma(aiz, s.ax, -aix * s.az);
    const float ez = (float)__builtin_fma(aix, s.ay, -aiy * s.ax);
    const float fd = Fp * __builtin_fmaf(dz, ez, __builtin_fmaf(dy, ey, dx * ex));
    a.v[3] = __builtin_fmaf(fd, dx, __builtin_fmaf(Q, ex, a.v[3]));
    a.v[4] = __builtin_fmaf(fd, dy, __builtin_fmaf(Q, ey, a.v[4]));
    a.v[5] = __builtin_fmaf(fd, dz, __builtin_fmaf(Q, ez, a.v[5]));
    a.v[6] = __builtin_fmaf(Zs, (float)(s.ax - aix), a.v[6]);
    a.v[7] = __builtin_fmaf(Zs, (float)(s.ay - aiy), a.v[7]);
    a.v[8] = __builtin_fmaf(Zs, (float)(s.az - aiz), a.v[8]);
    a.v[9] = __builtin_fmaf(Zs, (float)(s.vd - vi), a.v[9]);
}

// One far pair (t >= T_FAR32) of a uniform-core job in binary32: the far-
// field kernel and no PSE terms; the displacement and the stretching cross
// product as pair32 forms them.
template <int KIND>
__device__ __forceinline__ void pair32_far(double xi, double yi, double zi, double aix, double aiy,
                                           double aiz, const Rec32 &s, Acc32 &a) {
    const float dx = (float)(xi - s.x), dy = (float)(yi - s.y), dz = (float)(zi - s.z);
    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    float Q, Fp;
    kernel32_far<KIND>(__builtin_amdgcn_rsqf(r2), Q, Fp);
    const float cx = __builtin_fmaf(s.fy, dz, -s.fz * dy);
    const float cy = __builtin_fmaf(s.fz, dx, -s.fx * dz);
    const float cz = __builtin_fmaf(s.fx, dy, -s.fy * dx);
    a.v[0] = __builtin_fmaf(Q, cx, a.v[0]);
    a.v[1] = __builtin_fmaf(Q, cy, a.v[1]);
    a.v[2] = __builtin_fmaf(Q, cz, a.v[2]);
    const float ex = (float)__builtin_fma(aiy, s.az, -aiz * s.ay);
    const float ey = (float)__builtin_fma(aiz, s.ax, -aix * s.az);
    const float ez = (float)__builtin_fma(aix, s.ay, -aiy * s.ax);
    const float fd = Fp * __builtin_fmaf(dz, ez, __builtin_fmaf(dy, ey, dx * ex));
    a.v[3] = __builtin_fmaf(fd, dx, __builtin_fmaf(Q, ex, a.v[3]));
    a.v[4] = __builtin_fmaf(fd, dy, __builtin_fmaf(Q, ey, a.v[4]));
    a.v[5] = __builtin_fmaf(fd, dz, __builtin_fmaf(Q, ez, a.v[5]));
}

// Knuth's TwoSum: s + e == a + b exactly (round-to-nearest, no reassociation).
__device__ __forceinline__ void two_sum_acc(float &s, float &c, float b) {
    const float a = s;
    const float t = a + b;
    const float bb = t - a;
    const float e = (a - (t - bb)) + (b - bb);
    s = t;
    c += e;
}

// The k-th work item a block takes: its own block index first, then the
// queue's, which hands out gridDim.x, gridDim.x + 1, ...  A grid of no more
// blocks than items (the engines launch no more) takes one atomic per item
// past the first round only, and one with a block per item none.  Every
// thread of the block gets the same value, so the exit below is uniform.
__device__ __forceinline__ int next_item(int *counter, int *slot, int &k, int nitems) {
    if (k++ == 0) return (int)blockIdx.x;
    if ((int)gridDim.x >= nitems) return nitems;
    __syncthreads();
    if (threadIdx.x == 0) *slot = (int)gridDim.x + atomicAdd(counter, 1);
    __syncthreads();
    return *slot;
}

// Sources [j0, j0 + jn) of a staged tile (first source s0) against the
// thread's TPT targets of a uniform-core job; target slot p takes them in the
// far form by moments when bit p of FM is set (a full group: mom holds its
// moments, org its first particle).  DIAG: the tile holds some of the block's
// own targets, and a pair j == i is skipped (a far group is never the
// target's own).
template <int KIND, int TPT, int UNR, unsigned FM, bool DIAG>
__device__ __forceinline__ void group64(const Rec64 *tile, const Mom64 *mom, const double *org,
                                        int j0, int jn, const Rec64 *me, Acc64 *acc,
                                        double inv_s2, int s0, const int *idx) {
    double pv[TPT][3], S[TPT][3];
    if constexpr (FM != 0) {
#pragma unroll
        for (int p = 0; p < TPT; ++p)
            if ((FM >> p) & 1u) {
                const double ux = me[p].x - org[0], uy = me[p].y - org[1], uz = me[p].z - org[2];
                pv[p][0] = __builtin_fma(uy, me[p].az, -uz * me[p].ay);  // x'_i x a_i
                pv[p][1] = __builtin_fma(uz, me[p].ax, -ux * me[p].az);
                pv[p][2] = __builtin_fma(ux, me[p].ay, -uy * me[p].ax);
                S[p][0] = S[p][1] = S[p][2] = 0.0;
            }
    }
#pragma unroll UNR
    for (int j = j0; j < j0 + jn; ++j) {
        const Rec64 sj = tile[j];
        Mom64 mj;
        if constexpr (FM != 0) mj = mom[j];
#pragma unroll
        for (int p = 0; p < TPT; ++p) {
            if ((FM >> p) & 1u) {
                pair64_mom<KIND>(me[p].x, me[p].y, me[p].z, me[p].ax, me[p].ay, me[p].az, pv[p],
                                 sj, mj, acc[p], S[p]);
            } else {
                if (DIAG && s0 + j == idx[p]) continue;
                pair64<KIND, true>(me[p].x, me[p].y, me[p].z, me[p].ax, me[p].ay, me[p].az,
                                   me[p].h, sj, inv_s2, acc[p]);
            }
        }
    }
    if constexpr (FM != 0) {
        // velocity + S x x'_i, stretching + a_i x S
#pragma unroll
        for (int p = 0; p < TPT; ++p)
            if ((FM >> p) & 1u) {
                const double ux = me[p].x - org[0], uy = me[p].y - org[1], uz = me[p].z - org[2];
                Acc64 &a = acc[p];
                a.v[0] = __builtin_fma(S[p][1], uz, __builtin_fma(-S[p][2], uy, a.v[0]));
                a.v[1] = __builtin_fma(S[p][2], ux, __builtin_fma(-S[p][0], uz, a.v[1]));
                a.v[2] = __builtin_fma(S[p][0], uy, __builtin_fma(-S[p][1], ux, a.v[2]));
                a.v[3] = __builtin_fma(me[p].ay, S[p][2], __builtin_fma(-me[p].az, S[p][1], a.v[3]));
                a.v[4] = __builtin_fma(me[p].az, S[p][0], __builtin_fma(-me[p].ax, S[p][2], a.v[4]));
                a.v[5] = __builtin_fma(me[p].ax, S[p][1], __builtin_fma(-me[p].ay, S[p][0], a.v[5]));
            }
    }
}

// group64 for the far mask fm (one bit per target slot, wave-uniform).
template <int KIND, int TPT, int UNR, bool DIAG>
__device__ __forceinline__ void group64_fm(unsigned fm, const Rec64 *tile, const Mom64 *mom,
                                           const double *org, int j0, int jn, const Rec64 *me,
                                           Acc64 *acc, double inv_s2, int s0, const int *idx) {
    constexpr unsigned ALL = (1u << TPT) - 1;
    if (fm == 0)
        group64<KIND, TPT, UNR, 0, DIAG>(tile, mom, org, j0, jn, me, acc, inv_s2, s0, idx);
    else if (fm == ALL)
        group64<KIND, TPT, UNR, ALL, DIAG>(tile, mom, org, j0, jn, me, acc, inv_s2, s0, idx);
    else if constexpr (TPT == 2) {
        if (fm == 1)
            group64<KIND, TPT, UNR, 1, DIAG>(tile, mom, org, j0, jn, me, acc, inv_s2, s0, idx);
        else
            group64<KIND, TPT, UNR, 2, DIAG>(tile, mom, org, j0, jn, me, acc, inv_s2, s0, idx);
    }
}

// ------------------------------------------------------------- pair kernels
// Persistent: a fixed grid takes (job, target tile, source chunk) items from
// an atomic queue until it is empty.  A workgroup of TB threads owns a tile of
// TB x TPT targets (thread t holds targets t, t + TB, ...); the chunk's
// sources are staged through LDS, TB records at a time, with coalesced
// 16-byte loads, and read back as wave-uniform broadcasts, each source serving
// TPT targets per thread.  UNR is the unroll of the source loop.  Neither TPT
// nor UNR changes the order of any target's sums.  In a uniform-core job each
// (64-target group, 64-source group) whose bounding boxes are at least
// sqrt(T_FAR64) apart, the source group's no wider than sqrt(T_EXT64), runs in
// the far form by moments: the groups are aligned to 64 particles of the job,
// so the choice is the same for every variant.  Occupancy: one target per
// thread is held to three waves per SIMD (168 VGPRs; the moments took the
// 256 x 1, unroll 4 variant to 242, two waves, without the bound), two
// targets to two (256; at 312 the unroll-4 variant ran one).  Four waves for
// one target (128 VGPRs, before the moments) measured the same as three.
template <int KIND, int TB, int TPT, int UNR>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(TPT == 1 ? 3 : 2))) void pair64_kernel(const JobDev *__restrict__ jobs,
                                                    const Item *__restrict__ items,
                                                    const int *__restrict__ nitems_p,
                                                    int *__restrict__ counter) {
    constexpr int ST = TB;
    constexpr int TW = TB * TPT;  // targets per tile
    __shared__ Rec64 tile[ST];
    __shared__ Mom64 mom[ST];           // moments of full groups' sources
    __shared__ double sorg[ST / 64][3]; // first particle of each 64-source group
    __shared__ float sbox[ST / 64][6];  // boxes of the staged tile's 64-source groups
    __shared__ int slot;
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int nitems = *nitems_p;
    int taken = 0;
    for (;;) {
        const int it = next_item(counter, &slot, taken, nitems);
        if (it >= nitems) break;
        const Item item = items[it];
        const JobDev &J = jobs[item.job];
        const Rec64 *rec = (const Rec64 *)J.rec;
        const int n = J.n;
        const int t0 = item.tile * TW;
        int idx[TPT];
        Rec64 me[TPT];
        Acc64 acc[TPT];
        float tlo[TPT][3], thi[TPT][3];  // box of this wave's 64 targets of slot p
#pragma unroll
        for (int p = 0; p < TPT; ++p) {
            idx[p] = t0 + p * TB + (int)threadIdx.x;
            me[p] = rec[idx[p] < n ? idx[p] : n - 1];  // targets past n repeat particle n - 1
#pragma unroll
            for (int m = 0; m < NCOMP; ++m) acc[p].v[m] = 0.0;
            wave_box(me[p].x, me[p].y, me[p].z, tlo[p], thi[p]);
        }
        const double inv_s2 = J.inv_s2;
        const bool uniform = J.uniform != 0;
        const int c0 = item.chunk * J.chunk_len;
        const int c1 = min(c0 + J.chunk_len, n);
        for (int s0 = c0; s0 < c1; s0 += ST) {
            const int cnt = min(ST, c1 - s0);
            __syncthreads();
            {
                const double2 *src = (const double2 *)(rec + s0);
                double2 *dst = (double2 *)tile;
                for (int k = threadIdx.x; k < cnt * 4; k += TB) dst[k] = src[k];
            }
            // wave w: the box of source group w and its sources' moments
            // about its first particle, when full (read beside the copy,
            // shared through LDS)
            if (uniform && (w + 1) * 64 <= cnt) {
                const Rec64 q = rec[s0 + w * 64 + lane];
                float lo[3], hi[3];
                wave_box(q.x, q.y, q.z, lo, hi);
                const double ox = first_lane(q.x), oy = first_lane(q.y), oz = first_lane(q.z);
                const double ux = q.x - ox, uy = q.y - oy, uz = q.z - oz;
                Mom64 m;  // a_j x x'_j
                m.x = __builtin_fma(q.ay, uz, -q.az * uy);
                m.y = __builtin_fma(q.az, ux, -q.ax * uz);
                m.z = __builtin_fma(q.ax, uy, -q.ay * ux);
                m.pad = 0.0;
                mom[w * 64 + lane] = m;
                if (lane == 0) {
                    sorg[w][0] = ox;
                    sorg[w][1] = oy;
                    sorg[w][2] = oz;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        sbox[w][c] = lo[c];
                        sbox[w][3 + c] = hi[c];
                    }
                }
            }
            __syncthreads();
            // a tile holding some of the block's own targets skips j == i
            const bool diag = (s0 < t0 + TW) && (t0 < s0 + cnt);
            if (uniform) {
                // 64 sources at a time; target slot p takes them in the far
                // form when its group's box and theirs are far apart (never
                // its own group: the gap is 0) and theirs is narrow
                for (int g0 = 0; g0 < cnt; g0 += 64) {
                    const int gn = min(64, cnt - g0);
                    unsigned fm = 0;
                    if (gn == 64) {
                        float slo[3], shi[3], ext = 0.0f;
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            slo[c] = uniform_f(sbox[g0 >> 6][c]);
                            shi[c] = uniform_f(sbox[g0 >> 6][3 + c]);
                            ext = __builtin_fmaf(shi[c] - slo[c], shi[c] - slo[c], ext);
                        }
                        if (ext <= T_EXT64)
#pragma unroll
                            for (int p = 0; p < TPT; ++p)
                                if (box_gap2(tlo[p], thi[p], slo, shi) >= T_FAR64) fm |= 1u << p;
                        fm = __builtin_amdgcn_readfirstlane(fm);
                    }
                    const double *org = sorg[g0 >> 6];
                    if (diag)
                        group64_fm<KIND, TPT, UNR, true>(fm, tile, mom, org, g0, gn, me, acc,
                                                         inv_s2, s0, idx);
                    else
                        group64_fm<KIND, TPT, UNR, false>(fm, tile, mom, org, g0, gn, me, acc,
                                                          inv_s2, s0, idx);
                }
            } else if (!diag) {
#pragma unroll UNR
                for (int j = 0; j < cnt; ++j) {
                    const Rec64 sj = tile[j];
#pragma unroll
                    for (int p = 0; p < TPT; ++p)
                        pair64<KIND, false>(me[p].x, me[p].y, me[p].z, me[p].ax, me[p].ay,
                                            me[p].az, me[p].h, sj, inv_s2, acc[p]);
                }
            } else {
                for (int j = 0; j < cnt; ++j) {
                    const Rec64 sj = tile[j];
#pragma unroll
                    for (int p = 0; p < TPT; ++p) {
                        if (s0 + j == idx[p]) continue;
                        pair64<KIND, false>(me[p].x, me[p].y, me[p].z, me[p].ax, me[p].ay,
                                            me[p].az, me[p].h, sj, inv_s2, acc[p]);
                    }
                }
            }
        }
#pragma unroll
        for (int p = 0; p < TPT; ++p)
            if (idx[p] < n) {
                double *o = J.part + (size_t)item.chunk * NCOMP * J.npad + idx[p];
#pragma unroll
                for (int m = 0; m < NCOMP; ++m) o[(size_t)m * J.npad] = acc[p].v[m];
            }
    }
}

// One 32-source sub-block [j0, j0 + 32) of a staged binary32 tile (first
// source s0) against the thread's TPT targets of a uniform-core job; target
// slot p takes it in the far form when bit p of FM is set.  DIAG: skip j == i.
template <int KIND, int TPT, int UNR, unsigned FM, bool DIAG>
__device__ __forceinline__ void sub32(const Rec32 *tile, int j0, const Rec32 *me, const float *hi,
                                      float inv_s2, Acc32 *a, int s0, const int *idx) {
#pragma unroll UNR
    for (int j = j0; j < j0 + 32; ++j) {
        const Rec32 sj = tile[j];
#pragma unroll
        for (int p = 0; p < TPT; ++p) {
            if ((FM >> p) & 1u) {
                pair32_far<KIND>(me[p].x, me[p].y, me[p].z, me[p].ax, me[p].ay, me[p].az, sj, a[p]);
            } else {
                if (DIAG && s0 + j == idx[p]) continue;
                pair32<KIND, true>(me[p].x, me[p].y, me[p].z, me[p].ax, me[p].ay, me[p].az, me[p].vd,
                                   hi[p], sj, inv_s2, a[p]);
            }
        }
    }
}

template <int KIND, int TPT, int UNR, bool DIAG>
__device__ __forceinline__ void sub32_fm(unsigned fm, const Rec32 *tile, int j0, const Rec32 *me,
                                         const float *hi, float inv_s2, Acc32 *a, int s0,
                                         const int *idx) {
    constexpr unsigned ALL = (1u << TPT) - 1;
    if (fm == 0)
        sub32<KIND, TPT, UNR, 0, DIAG>(tile, j0, me, hi, inv_s2, a, s0, idx);
    else if (fm == ALL)
        sub32<KIND, TPT, UNR, ALL, DIAG>(tile, j0, me, hi, inv_s2, a, s0, idx);
    else if constexpr (TPT == 2) {
        if (fm == 1)
            sub32<KIND, TPT, UNR, 1, DIAG>(tile, j0, me, hi, inv_s2, a, s0, idx);
        else
            sub32<KIND, TPT, UNR, 2, DIAG>(tile, j0, me, hi, inv_s2, a, s0, idx);
    }
}

// Compensated binary32 pair pass: plain binary32 sums over SB sources, folded
// into a (sum, carry) pair per component with TwoSum; the item's result is
// sum + carry in binary64, and chunks are added in binary64 afterwards.  Sub-
// block boundaries are fixed by the chunk start, whatever TB, TPT and UNR are.
// A uniform-core job's pairs between a 64-target group and a 64-source group
// whose boxes are at least sqrt(T_FAR32) apart take the far form (chosen per
// aligned group pair, so the same for every variant).
template <int KIND, int TB, int TPT, int UNR>
__global__ __launch_bounds__(TB) void pair32_kernel(const JobDev *__restrict__ jobs,
                                                    const Item *__restrict__ items,
                                                    const int *__restrict__ nitems_p,
                                                    int *__restrict__ counter) {
    constexpr int ST = TB;
    constexpr int SB = 32;
    constexpr int TW = TB * TPT;
    static_assert(ST % 64 == 0, "64-source groups must divide the LDS tile");
    __shared__ Rec32 tile[ST];
    __shared__ float sbox[ST / 64][6];  // boxes of the staged tile's 64-source groups
    __shared__ int slot;
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int nitems = *nitems_p;
    int taken = 0;
    for (;;) {
        const int it = next_item(counter, &slot, taken, nitems);
        if (it >= nitems) break;
        const Item item = items[it];
        const JobDev &J = jobs[item.job];
        const Rec32 *rec = (const Rec32 *)J.rec;
        const int n = J.n;
        const int t0 = item.tile * TW;
        int idx[TPT];
        Rec32 me[TPT];
        float hi[TPT];
        Acc32 sum[TPT], car[TPT];
        float tlo[TPT][3], thi[TPT][3];  // box of this wave's 64 targets of slot p
#pragma unroll
        for (int p = 0; p < TPT; ++p) {
            idx[p] = t0 + p * TB + (int)threadIdx.x;
            me[p] = rec[idx[p] < n ? idx[p] : n - 1];  // targets past n repeat particle n - 1
            hi[p] = (float)me[p].h;
#pragma unroll
            for (int m = 0; m < NCOMP; ++m) sum[p].v[m] = car[p].v[m] = 0.0f;
            wave_box(me[p].x, me[p].y, me[p].z, tlo[p], thi[p]);
        }
        const float inv_s2 = (float)J.inv_s2;
        const bool uniform = J.uniform != 0;
        const int c0 = item.chunk * J.chunk_len;
        const int c1 = min(c0 + J.chunk_len, n);
        for (int s0 = c0; s0 < c1; s0 += ST) {
            const int cnt = min(ST, c1 - s0);
            __syncthreads();
            {
                const float4 *src = (const float4 *)(rec + s0);
                float4 *dst = (float4 *)tile;
                for (int k = threadIdx.x; k < cnt * 5; k += TB) dst[k] = src[k];
            }
            // wave w: the box of source group w, when full (read beside the
            // copy, shared through LDS)
            if (uniform && (w + 1) * 64 <= cnt) {
                const Rec32 &q = rec[s0 + w * 64 + lane];
                float lo[3], hi[3];
                wave_box(q.x, q.y, q.z, lo, hi);
                if (lane == 0)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        sbox[w][c] = lo[c];
                        sbox[w][3 + c] = hi[c];
                    }
            }
            __syncthreads();
            const bool diag = (s0 < t0 + TW) && (t0 < s0 + cnt);
            unsigned fm = 0;  // far mask of the 64-source group holding the sub-block
            for (int b0 = 0; b0 < cnt; b0 += SB) {
                const int bn = min(SB, cnt - b0);
                if ((b0 & 63) == 0) {
                    fm = 0;
                    if (uniform && b0 + 64 <= cnt) {
                        float slo[3], shi[3];
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            slo[c] = uniform_f(sbox[b0 >> 6][c]);
                            shi[c] = uniform_f(sbox[b0 >> 6][3 + c]);
                        }
#pragma unroll
                        for (int p = 0; p < TPT; ++p)
                            if (box_gap2(tlo[p], thi[p], slo, shi) >= T_FAR32) fm |= 1u << p;
                        fm = __builtin_amdgcn_readfirstlane(fm);
                    }
                }
                Acc32 a[TPT];
#pragma unroll
                for (int p = 0; p < TPT; ++p)
#pragma unroll
                    for (int m = 0; m < NCOMP; ++m) a[p].v[m] = 0.0f;
                if (uniform && bn == SB) {
                    if (diag)
                        sub32_fm<KIND, TPT, UNR, true>(fm, tile, b0, me, hi, inv_s2, a, s0, idx);
                    else
                        sub32_fm<KIND, TPT, UNR, false>(fm, tile, b0, me, hi, inv_s2, a, s0, idx);
                } else if (!diag && bn == SB) {
#pragma unroll UNR
                    for (int j = 0; j < SB; ++j) {
                        const Rec32 sj = tile[b0 + j];
#pragma unroll
                        for (int p = 0; p < TPT; ++p)
                            pair32<KIND, false>(me[p].x, me[p].y, me[p].z, me[p].ax, me[p].ay,
                                                me[p].az, me[p].vd, hi[p], sj, inv_s2, a[p]);
                    }
                } else {
                    for (int j = 0; j < bn; ++j) {
                        const Rec32 sj = tile[b0 + j];
#pragma unroll
                        for (int p = 0; p < TPT; ++p) {
                            if (s0 + b0 + j == idx[p]) continue;
                            if (uniform)
                                pair32<KIND, true>(me[p].x, me[p].y, me[p].z, me[p].ax, me[p].ay,
                                                   me[p].az, me[p].vd, hi[p], sj, inv_s2, a[p]);
                            else
                                pair32<KIND, false>(me[p].x, me[p].y, me[p].z, me[p].ax,
                                                    me[p].ay, me[p].az, me[p].vd, hi[p], sj, inv_s2,
                                                    a[p]);
                        }
                    }
                }
#pragma unroll
                for (int p = 0; p < TPT; ++p)
#pragma unroll
                    for (int m = 0; m < NCOMP; ++m) two_sum_acc(sum[p].v[m], car[p].v[m], a[p].v[m]);
            }
        }
#pragma unroll
        for (int p = 0; p < TPT; ++p)
            if (idx[p] < n) {
                double *o = J.part + (size_t)item.chunk * NCOMP * J.npad + idx[p];
#pragma unroll
                for (int m = 0; m < NCOMP; ++m)
                    o[(size_t)m * J.npad] = (double)sum[p].v[m] + (double)car[p].v[m];
            }
    }
}

// ------------------------------------------- symmetric binary32 pair pass
// Each unordered pair (i < j) is evaluated once: the kernel values, the
// displacement and the stretching terms are shared by both particles, which
// saves about a third of the arithmetic per ordered pair.  A workgroup owns a
// block of TB x TPT targets (thread t holds targets t, t + TB, ...) and walks
// the sources j > i of one chunk, 64 at a time, staged through LDS as
// arrays of 16-byte field pairs.  Within a 64-source tile, lane l takes source
// (l + s) mod 64 at step s, so the 64 lanes meet 64 different sources at
// every step and each source meets every lane once in 64 steps.  The source-
// side sums ride along: each step they move one lane down (DPP wave rotate,
// fused into the add where the compiler can), so lane l always holds the
// sums of the source it is working on.  (LDS float atomics, the other way to
// keep per-source sums, measured 12x slower on gfx950.)  Both sides are
// binary32 runs of PW_SYM_RUN terms added into binary64: the target side in
// registers, then the (target, chunk) slot; the source side in LDS, then
// added over the waves in wave order and written once to the (source, block)
// slot.  Every slot has one writer and every sum a fixed order, so results do
// not depend on scheduling.

// Terms per binary32 run before it is added into binary64 (both sides).
// Shorter runs bound the rounding of the near-field terms, which in this
// pass land on different sides for a particle's neighbours before and after
// it and cancel only in binary64: 32 terms cost 2x the one-sided pass's
// stretching error on the rotor wake, 16 terms match it.
#ifndef PW_SYM_RUN
#define PW_SYM_RUN 16
#endif


struct Acc32s {
    float v[NCOMP_S];
};

// x of lane (l + 1) mod 64, added to y (v_add_f32 with a wave_rol:1 operand).
__device__ __forceinline__ float rot_add(float x, float y) {
    return __builtin_bit_cast(float,
                              __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x134, 0xf, 0xf, true)) +
           y;
}

// A source tile in LDS as arrays of 16-byte pairs of fields, so that 64
// lanes reading 64 different sources read 64 consecutive 16-byte words with
// one ds_read_b128 per pair.  Every source is stored twice (entries k and
// k + 64), so lane l reads source (l + s) mod 64 at entry l + s: within an
// unrolled run of steps every load is one base register plus an immediate
// offset.  (With one array per field the compiler pairs the loads into
// ds_read2 forms whose 8-bit offsets need a new base register every step:
// 4 more VALU instructions per step, 65.3 against 61.3 ms per launch at
// N = 262,144.)
struct SrcTile {
    double2 xy[128], za[128], aa[128];  // (x, y), (z, ax), (ay, az)
    float4 f[128];                      // (fx, fy, fz, v)
    float4 p[128];                      // (x - c, y - c, z - c, h) in binary32, c the centre of the
                                        // block's box: far tiles
    float h[128];
    double vd[128];                     // binary64 volumes: PSE differences of close pairs
};

// Fields of one particle in registers.
struct Part32 {
    double x, y, z, ax, ay, az;
    float fx, fy, fz, v, h;
    double vd;
};
// The same for a far tile: every field in binary32.
struct Part32f {
    float x, y, z, fx, fy, fz, v, h;
};

// A source tile is far from a target block when the distance between their
// bounding boxes is at least K_FAR times the largest coordinate magnitude of
// the two boxes, measured from the centre of the block's box, relative to
// which a far tile's sources and the block's targets are rounded to binary32.
// Its pairs then run wholly in binary32: the displacement from those binary32
// positions has a relative error below 2 eps / K_FAR (1e-6),
// and the stretching cross product from binary32 strengths loses accuracy
// only in proportion to the pair's small weight.  Emulated against the
// oracle, the far pairs add 4e-8 relative RMS to the stretching rates of the
// 262,144-particle rotor wake (87% of its pairs far), 6e-9 on two leapfrogging
// rings (96% far), 2e-9 on one ring (69% far); velocities less.
constexpr float K_FAR = 0.125f;

__device__ __forceinline__ Part32f load_src(const SrcTile &t, int k, Part32f) {
    const float4 p = t.p[k], f = t.f[k];
    return Part32f{p.x, p.y, p.z, f.x, f.y, f.z, f.w, p.w};
}

// Displacement x_i - x_j in binary32: formed in binary64 and rounded, or (far
// tiles) from the binary32 positions.
__device__ __forceinline__ float3 disp32(const Part32 &i, const Part32 &j) {
    return make_float3((float)(i.x - j.x), (float)(i.y - j.y), (float)(i.z - j.z));
}
__device__ __forceinline__ float3 disp32(const Part32f &i, const Part32f &j) {
    return make_float3(i.x - j.x, i.y - j.y, i.z - j.z);
}
// Stretching cross product a_i x a_j: formed in binary64 from the binary64
// strengths and rounded, or (far tiles) in binary32.
__device__ __forceinline__ float3 cross32(const Part32 &i, const Part32 &j) {
    return make_float3((float)__builtin_fma(i.ay, j.az, -i.az * j.ay),
                       (float)__builtin_fma(i.az, j.ax, -i.ax * j.az),
                       (float)__builtin_fma(i.ax, j.ay, -i.ay * j.ax));
}
__device__ __forceinline__ float3 cross32(const Part32f &i, const Part32f &j) {
    return make_float3(__builtin_fmaf(i.fy, j.fz, -i.fz * j.fy),
                       __builtin_fmaf(i.fz, j.fx, -i.fx * j.fz),
                       __builtin_fmaf(i.fx, j.fy, -i.fy * j.fx));
}

__device__ __forceinline__ Part32 load_src(const SrcTile &t, int k, Part32) {
    Part32 p;
    const double2 a = t.xy[k], b = t.za[k], c = t.aa[k];
    const float4 f = t.f[k];
    p.x = a.x;
    p.y = a.y;
    p.z = b.x;
    p.ax = b.y;
    p.ay = c.x;
    p.az = c.y;
    p.fx = f.x;
    p.fy = f.y;
    p.fz = f.z;
    p.v = f.w;
    p.h = t.h[k];
    p.vd = t.vd[k];
    return p;
}

// One unordered pair (target me, source s): target-side terms added to a,
// source-side terms stored (FIRST) or added to b.  MASKED: a pair with
// valid == false contributes exactly zero (its operands are finite).
// DIFF == false: a far pair of a uniform-core job (t >= T_FAR32): the far-
// field kernel and no PSE terms (sums 6-9 untouched on both sides).
template <int KIND, bool UNIFORM, bool MASKED, bool FIRST, bool DIFF = true, bool DEG4 = false,
          class PT>
__device__ __forceinline__ void pair32s(const PT &me, const PT &s, bool valid, Acc32s &a,
                                        Acc32s &b) {
    static_assert(DIFF || (UNIFORM && !MASKED), "the far form is for full tiles of uniform jobs");
    const float3 d = disp32(me, s);
    const float dx = d.x, dy = d.y, dz = d.z;
    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    float Q, Fp, Zs = 0.0f;  // far-field form: Fp holds F'/Q
    if (!DIFF) {
        kernel32_farq<KIND, DEG4>(__builtin_amdgcn_rsqf(r2), Q, Fp);
    } else if (UNIFORM) {
        const float t = MASKED ? (valid ? r2 : 1.0f) : r2;
        kernel32<KIND>(t, __builtin_amdgcn_rsqf(t), Q, Fp, Zs);
    } else {
        const float isg = __builtin_amdgcn_rsqf(me.h + s.h);
        const float is2 = isg * isg;
        float t = r2 * is2;
        if (MASKED) t = valid ? t : 1.0f;
        kernel32<KIND>(t, __builtin_amdgcn_rsqf(t), Q, Fp, Zs);
        const float s3 = isg * is2, s5 = s3 * is2;
        Q *= s3;
        Fp *= s5;
        Zs *= s5;
    }
    if (MASKED) {
        Q = valid ? Q : 0.0f;
        Fp = valid ? Fp : 0.0f;
        Zs = valid ? Zs : 0.0f;
    }
    // velocity of i: q (a_j x d) = a_j x u; of j: q (a_i x (x_j - x_i)) =
    // u x a_i, with u = q d formed once (3 operations fewer than scaling
    // each side's cross product)
    const float ux = Q * dx, uy = Q * dy, uz = Q * dz;
    a.v[0] = __builtin_fmaf(s.fy, uz, __builtin_fmaf(-s.fz, uy, a.v[0]));
    a.v[1] = __builtin_fmaf(s.fz, ux, __builtin_fmaf(-s.fx, uz, a.v[1]));
    a.v[2] = __builtin_fmaf(s.fx, uy, __builtin_fmaf(-s.fy, ux, a.v[2]));
    // stretching of i: q e + F (d.e) d with e = a_i x a_j (cross32); of j:
    // the same with e and d negated, i.e. minus it
    const float3 e = cross32(me, s);
    const float ex = e.x, ey = e.y, ez = e.z;
    const float de = __builtin_fmaf(dz, ez, __builtin_fmaf(dy, ey, dx * ex));
    float sx, sy, sz;
    if (!DIFF) {  // q (e + (F'/q)(d.e) d): one operation fewer
        const float k = Fp * de;
        sx = Q * __builtin_fmaf(k, dx, ex);
        sy = Q * __builtin_fmaf(k, dy, ey);
        sz = Q * __builtin_fmaf(k, dz, ez);
    } else {
        const float fd = Fp * de;
        sx = __builtin_fmaf(fd, dx, Q * ex);
        sy = __builtin_fmaf(fd, dy, Q * ey);
        sz = __builtin_fmaf(fd, dz, Q * ez);
    }
    a.v[3] += sx;
    a.v[4] += sy;
    a.v[5] += sz;
    // PSE sums.  A far tile's pair (Part32f): i gets k a_j, k v_j and j gets
    // k a_i, k v_i.  A close pair (Part32) adds the differences, formed in
    // binary64 and rounded: i gets k (a_j - a_i), k (v_j - v_i), j their
    // negatives.  The epilogue's v_i S_a - a_i S_v is the same sum either way,
    // k (v_i a_j - v_j a_i) per pair; in a smooth core the two sums agree to a
    // digit or two, and the difference form keeps the rounding of a_j and v_j
    // to binary32 out of that cancellation.
    float tg[4], sg[4], zsg;
    if constexpr (std::is_same<PT, Part32>::value) {
        tg[0] = sg[0] = (float)(s.ax - me.ax);
        tg[1] = sg[1] = (float)(s.ay - me.ay);
        tg[2] = sg[2] = (float)(s.az - me.az);
        tg[3] = sg[3] = (float)(s.vd - me.vd);
        zsg = -Zs;
    } else {
        tg[0] = s.fx, tg[1] = s.fy, tg[2] = s.fz, tg[3] = s.v;
        sg[0] = me.fx, sg[1] = me.fy, sg[2] = me.fz, sg[3] = me.v;
        zsg = Zs;
    }
    if (DIFF) {
#pragma unroll
        for (int m = 0; m < 4; ++m) a.v[6 + m] = __builtin_fmaf(Zs, tg[m], a.v[6 + m]);
    }
    if (FIRST) {
        b.v[0] = __builtin_fmaf(uy, me.fz, -uz * me.fy);
        b.v[1] = __builtin_fmaf(uz, me.fx, -ux * me.fz);
        b.v[2] = __builtin_fmaf(ux, me.fy, -uy * me.fx);
        b.v[3] = sx;
        b.v[4] = sy;
        b.v[5] = sz;
        if (DIFF) {
#pragma unroll
            for (int m = 0; m < 4; ++m) b.v[6 + m] = zsg * sg[m];
        }
    } else {
        b.v[0] = __builtin_fmaf(uy, me.fz, __builtin_fmaf(-uz, me.fy, b.v[0]));
        b.v[1] = __builtin_fmaf(uz, me.fx, __builtin_fmaf(-ux, me.fz, b.v[1]));
        b.v[2] = __builtin_fmaf(ux, me.fy, __builtin_fmaf(-uy, me.fx, b.v[2]));
        b.v[3] += sx;
        b.v[4] += sy;
        b.v[5] += sz;
        if (DIFF) {
#pragma unroll
            for (int m = 0; m < 4; ++m) b.v[6 + m] = __builtin_fmaf(zsg, sg[m], b.v[6 + m]);
        }
    }
}

// Steps [s_begin, s_begin + SB) of the rotation over one staged tile.
// MASKED: per-target-slot state: live[p] (slot p pairs with this tile at
// all), nt[p] (targets of the slot's 64 that exist), diag[p] (the tile is
// the slot's own: only j > i); cnt sources exist in the tile.
// PT: Part32, or Part32f for a far tile (never masked).  DIFF == false: a
// far tile in the far-field form; sums 6-9 are neither added nor rotated (64
// rotations of a sum that receives nothing leave it where it was).
template <int KIND, bool UNIFORM, bool MASKED, int TPT, int SB, bool DIFF = true, bool DEG4 = false,
          class PT>
__device__ __forceinline__ void sym_steps(const SrcTile &tile, int s_begin, int lane, const PT *me,
                                          const bool *live, const int *nt, const bool *diag,
                                          int cnt, Acc32s *a, Acc32s &jacc) {
    const int e0 = lane + s_begin;  // entry of step s_begin
#pragma unroll
    for (int u = 0; u < SB; ++u) {
        const int s = s_begin + u;
        const int k = (lane + s) & 63;
        const PT sj = load_src(tile, e0 + u, PT{});
        Acc32s b;
        if (!MASKED) {
            pair32s<KIND, UNIFORM, false, true, DIFF, DEG4>(me[0], sj, true, a[0], b);
#pragma unroll
            for (int p = 1; p < TPT; ++p)
                pair32s<KIND, UNIFORM, false, false, DIFF, DEG4>(me[p], sj, true, a[p], b);
        } else {
            const bool src_ok = k < cnt;
#pragma unroll
            for (int m = 0; m < NCOMP_S; ++m) b.v[m] = 0.0f;
#pragma unroll
            for (int p = 0; p < TPT; ++p) {
                if (!live[p]) continue;
                const bool valid = src_ok && lane < nt[p] && (!diag[p] || lane + s < 64) &&
                                   !(diag[p] && s == 0);
                pair32s<KIND, UNIFORM, true, false>(me[p], sj, valid, a[p], b);
            }
        }
        // the caller starts jacc at zero: the first step takes b as it is
#pragma unroll
        for (int m = 0; m < (DIFF ? NCOMP_S : 6); ++m)
            jacc.v[m] = u == 0 ? b.v[m] : rot_add(jacc.v[m], b.v[m]);
    }
}

// Whether a tile whose box is [tlo, thi] is far from a block whose box is
// [blo, bhi] (K_FAR, coordinates measured from ctr); d2 receives the squared
// distance between the boxes.
__device__ __forceinline__ bool far_boxes(const float *blo, const float *bhi, const float *tlo,
                                          const float *thi, const float *ctr, float &d2) {
    float m = 0.0f;
    d2 = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float g = fmaxf(0.0f, fmaxf(tlo[c] - bhi[c], blo[c] - thi[c]));
        d2 = __builtin_fmaf(g, g, d2);
        m = fmaxf(m, fmaxf(fmaxf(fabsf(blo[c] - ctr[c]), fabsf(bhi[c] - ctr[c])),
                           fmaxf(fabsf(tlo[c] - ctr[c]), fabsf(thi[c] - ctr[c]))));
    }
    const float km = K_FAR * m;
    return d2 >= km * km;
}

// Occupancy: the HIGHORDER kernel with one target per thread is held to 128
// VGPRs, i.e. 4 waves per SIMD (the LDS of a 256-thread block, 37 KB, allows
// 4 blocks per CU).  Left to itself the compiler takes 148 (3 waves).  The
// unrolled steps are unchanged (the limit costs no spill inside them); at
// N = 262,144 the pass took 65.3 ms instead of 68.0 (before far tiles).  Five
// waves (96 VGPRs) spill inside the steps and take 90 ms; GAUSSIAN spills at
// 128.  Two targets per thread issue fewer instructions per pair (51.75 per
// far-field pair against 54.1: half the rotations and LDS reads) but need
// 252 VGPRs: held to two waves per SIMD they take 49.7 ms against 45.3 (at
// three, 74.4 against 72.0 before far tiles).  Other instantiations keep the
// compiler's choice.
template <int KIND, int TB, int TPT>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(
    (KIND == KIND_HIGHORDER && TPT == 1) ? 4 : 1))) void pair32s_kernel(const JobDev *__restrict__ jobs,
                                                                        const Item *__restrict__ items,
                                                                        const int *__restrict__ nitems_p,
                                                                        int *__restrict__ counter) {
    constexpr int NW = TB / 64;
    constexpr int TW = TB * TPT;
    constexpr int SB = PW_SYM_RUN;         // steps per target-side binary32 run
    constexpr int SBJ = PW_SYM_RUN / TPT;  // steps per source-side run (same term count)
    __shared__ SrcTile tile;
    __shared__ double jl[NW][NCOMP_S][64];  // per-source binary64 sums, per wave
    __shared__ double prev[NCOMP_S][64];    // the strip row's sums before this block
    __shared__ float bbw[NW][6];            // per wave: box of its targets (lo xyz, hi xyz)
    __shared__ int tfar;                    // the staged tile is far from the block (1),
                                            // and t >= T_FAR32 for all its pairs (2); not
                                            // far, but t >= T_FAR32 (3); far, with
                                            // T_FAR32B <= t (4: one series term more)
    __shared__ int slot;
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int nitems = *nitems_p;
    int taken = 0;
    for (;;) {
        // uniform: the item's fields and its job's are scalar loads
        const int it = __builtin_amdgcn_readfirstlane(next_item(counter, &slot, taken, nitems));
        if (it >= nitems) break;
        const Item item = items[it];
        const JobDev &J = jobs[item.job];
        const Rec32 *rec = (const Rec32 *)J.rec;
        const int n = J.n;
        // the item is one strip of blocks against one source chunk; the
        // strip's source-side sums go to one slot row, the blocks adding in
        // order into what the first block wrote
        const int bper = J.sw / TW, nblk = (n + TW - 1) / TW;
        const int b_lo = item.tile * bper, b_hi = min(b_lo + bper, nblk);
        const int S0 = item.tile * J.sw;
        const size_t len = (size_t)J.npad - (size_t)S0;  // this strip's slot row length
        double *jrow = J.jpart + jpart_offset((size_t)item.tile, (size_t)J.npad, (size_t)J.sw);
        for (int blk = b_lo; blk < b_hi; ++blk) {
        const int B0 = blk * TW;
        const int c0 = max(item.chunk * J.chunk_len, B0);
        const int c1 = min(item.chunk * J.chunk_len + J.chunk_len, n);
        if (c0 >= c1) break;  // later blocks start later still
        const bool first = blk == b_lo;
        int idx[TPT], wt[TPT], nt[TPT];
        Part32 me[TPT];
        double isum[TPT][NCOMP_S];  // target-side sums: 32-term binary32 runs added in binary64
#pragma unroll
        for (int p = 0; p < TPT; ++p) {
            idx[p] = B0 + p * TB + (int)threadIdx.x;
            wt[p] = (B0 + p * TB) / 64 + w;  // this wave's 64-target tile
            nt[p] = max(0, min(64, n - wt[p] * 64));
            const Rec32 r = rec[idx[p] < n ? idx[p] : n - 1];
            me[p] = Part32{r.x, r.y, r.z, r.ax, r.ay, r.az, r.fx, r.fy, r.fz, r.v, (float)r.h, r.vd};
#pragma unroll
            for (int m = 0; m < NCOMP_S; ++m) isum[p][m] = 0.0;
        }
        {  // this wave's share of the block's box (targets past n repeat particle n - 1)
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
            for (int p = 0; p < TPT; ++p) {
                float plo[3], phi[3];
                wave_box(me[p].x, me[p].y, me[p].z, plo, phi);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    lo[c] = fminf(lo[c], plo[c]);
                    hi[c] = fmaxf(hi[c], phi[c]);
                }
            }
            if (lane == 0)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    bbw[w][c] = lo[c];
                    bbw[w][3 + c] = hi[c];
                }
        }
        const bool uniform = J.uniform != 0;
        for (int s0 = c0; s0 < c1; s0 += 64) {
            const int cnt = min(64, c1 - s0);
            const int st = s0 >> 6;  // source tile
            __syncthreads();
            // the block's box (the waves' boxes were written before the
            // barrier above) and its centre, from which a far tile's binary32
            // positions are measured
            float blo[3], bhi[3], ctr[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                blo[c] = bbw[0][c];
                bhi[c] = bbw[0][3 + c];
            }
#pragma unroll
            for (int ww = 1; ww < NW; ++ww)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    blo[c] = fminf(blo[c], bbw[ww][c]);
                    bhi[c] = fmaxf(bhi[c], bbw[ww][3 + c]);
                }
#pragma unroll
            for (int c = 0; c < 3; ++c) ctr[c] = 0.5f * (blo[c] + bhi[c]);
            if (threadIdx.x < 64) {  // stage twice; missing sources read as zeros
                const int k = (int)threadIdx.x;
                Rec32 r;
                if (k < cnt) {
                    r = rec[s0 + k];
                } else {
                    r.x = r.y = r.z = r.h = r.ax = r.ay = r.az = r.vd = 0.0;
                    r.fx = r.fy = r.fz = r.v = 0.0f;
                }
                const float px = (float)(r.x - (double)ctr[0]), py = (float)(r.y - (double)ctr[1]),
                            pz = (float)(r.z - (double)ctr[2]);
#pragma unroll
                for (int c = 0; c < 128; c += 64) {
                    tile.xy[k + c] = make_double2(r.x, r.y);
                    tile.za[k + c] = make_double2(r.z, r.ax);
                    tile.aa[k + c] = make_double2(r.ay, r.az);
                    tile.f[k + c] = make_float4(r.fx, r.fy, r.fz, r.v);
                    tile.p[k + c] = make_float4(px, py, pz, (float)r.h);
                    tile.h[k + c] = (float)r.h;
                    tile.vd[k + c] = r.vd;
                }
                // far test (full tiles only): the tile's box against the
                // block's
                int f = 0;
                if (cnt == 64) {
                    float tlo[3], thi[3];
                    wave_box(r.x, r.y, r.z, tlo, thi);
                    float d2;
                    f = far_boxes(blo, bhi, tlo, thi, ctr, d2) ? 1 : 0;
                    // uniform core: coordinates are in units of sigma, t = rho^2
                    if (uniform && d2 >= T_FAR32)
                        f = f ? 2 : 3;
                    else if (uniform && f && d2 >= T_FAR32B)
                        f = 4;
                }
                if (k == 0) tfar = f;
            }
            if (!first)  // the row's sums so far, loaded beside the tile
                for (int e = (NW > 1 ? (int)threadIdx.x - 64 : (int)threadIdx.x); e >= 0 && e < NCOMP_S * 64;
                     e += (NW > 1 ? TB - 64 : TB)) {
                    const int m = e >> 6, j = e & 63;
                    if (j < cnt) prev[m][j] = jrow[(size_t)m * len + (size_t)(s0 - S0 + j)];
                }
            __syncthreads();
            const int farv = __builtin_amdgcn_readfirstlane(tfar);
            // far: binary32 positions and strengths; ffield: the far-field
            // kernel and no PSE terms (binary64 displacement and cross product
            // unless also far)
            const bool far = farv == 1 || farv == 2 || farv == 4, ffield = farv >= 2;
            bool live[TPT], diag[TPT], full = cnt == 64, any = false;
#pragma unroll
            for (int p = 0; p < TPT; ++p) {
                live[p] = st >= wt[p] && nt[p] > 0;
                diag[p] = st == wt[p];
                full = full && st > wt[p] && nt[p] == 64;
                any = any || live[p];
            }
#pragma unroll
            for (int m = 0; m < NCOMP_S; ++m) jl[w][m][lane] = 0.0;
            // other lanes of this wave add into these words next
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            if (any) {
#pragma unroll 1
                for (int sb = 0; sb < 64; sb += SB) {
                    Acc32s a[TPT];
#pragma unroll
                    for (int p = 0; p < TPT; ++p)
#pragma unroll
                        for (int m = 0; m < NCOMP_S; ++m) a[p].v[m] = 0.0f;
#pragma unroll 1
                    for (int sj = sb; sj < sb + SB; sj += SBJ) {
                        Acc32s jacc;
#pragma unroll
                        for (int m = 0; m < NCOMP_S; ++m) jacc.v[m] = 0.0f;
                        if (full && far) {
                            Part32f mf[TPT];
#pragma unroll
                            for (int p = 0; p < TPT; ++p)
                                mf[p] = Part32f{(float)(me[p].x - (double)ctr[0]),
                                                (float)(me[p].y - (double)ctr[1]),
                                                (float)(me[p].z - (double)ctr[2]), me[p].fx, me[p].fy,
                                                me[p].fz, me[p].v, me[p].h};
                            if (farv == 4)
                                sym_steps<KIND, true, false, TPT, SBJ, false, true>(
                                    tile, sj, lane, mf, live, nt, diag, cnt, a, jacc);
                            else if (ffield)
                                sym_steps<KIND, true, false, TPT, SBJ, false>(tile, sj, lane, mf, live,
                                                                              nt, diag, cnt, a, jacc);
                            else if (uniform)
                                sym_steps<KIND, true, false, TPT, SBJ>(tile, sj, lane, mf, live, nt,
                                                                       diag, cnt, a, jacc);
                            else
                                sym_steps<KIND, false, false, TPT, SBJ>(tile, sj, lane, mf, live,
                                                                        nt, diag, cnt, a, jacc);
                        } else if (full) {
                            if (ffield)
                                sym_steps<KIND, true, false, TPT, SBJ, false>(tile, sj, lane, me, live,
                                                                              nt, diag, cnt, a, jacc);
                            else if (uniform)
                                sym_steps<KIND, true, false, TPT, SBJ>(tile, sj, lane, me, live, nt,
                                                                       diag, cnt, a, jacc);
                            else
                                sym_steps<KIND, false, false, TPT, SBJ>(tile, sj, lane, me, live,
                                                                        nt, diag, cnt, a, jacc);
                        } else {
                            if (uniform)
                                sym_steps<KIND, true, true, TPT, SBJ>(tile, sj, lane, me, live, nt,
                                                                      diag, cnt, a, jacc);
                            else
                                sym_steps<KIND, false, true, TPT, SBJ>(tile, sj, lane, me, live, nt,
                                                                       diag, cnt, a, jacc);
                        }
                        // end of a source-side run: after step sj + SBJ - 1
                        // lane l holds the run sums of source
                        // (l + sj + SBJ - 1) mod 64, one lane per source
                        const int k = (lane + sj + SBJ - 1) & 63;
#pragma unroll
                        // ds_add_f64: the add happens in LDS (one instruction instead of
                        // read, add, write; each lane owns its entry, so the sums and
                        // their order are those of the plain update: bitwise the same
                        // rates, 0.35% faster at N = 262,144)
                        for (int m = 0; m < NCOMP_S; ++m)
                            if (m < 6 || !(full && ffield))  // far-field runs add no PSE sums
                                __hip_atomic_fetch_add(&jl[w][m][k], (double)jacc.v[m], __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WAVEFRONT);
                        // the next run's lanes update entries this run's
                        // lanes wrote
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                    }
#pragma unroll
                    for (int p = 0; p < TPT; ++p)
#pragma unroll
                        for (int m = 0; m < NCOMP_S; ++m) isum[p][m] += (double)a[p].v[m];
                }
            }
            __syncthreads();
            // source-side totals of this tile, waves added in order
            for (int e = threadIdx.x; e < NCOMP_S * 64; e += TB) {
                const int m = e >> 6, j = e & 63;
                if (j < cnt) {
                    double v = jl[0][m][j];
#pragma unroll
                    for (int ww = 1; ww < NW; ++ww) v += jl[ww][m][j];
                    jrow[(size_t)m * len + (size_t)(s0 - S0 + j)] = first ? v : prev[m][j] + v;
                }
            }
        }
        const size_t pl = sym_part_len((size_t)item.chunk, (size_t)J.npad, (size_t)J.chunk_len, (size_t)TW);
        double *orow = J.part + sym_part_offset((size_t)item.chunk, (size_t)J.npad, (size_t)J.chunk_len, (size_t)TW);
#pragma unroll
        for (int p = 0; p < TPT; ++p)
            if (idx[p] < n) {
#pragma unroll
                for (int m = 0; m < NCOMP_S; ++m) orow[(size_t)m * pl + idx[p]] = isum[p][m];
            }
        }  // blocks of the strip
    }
}

// ------------------------------------------- symmetric binary64 pair pass
// The same decomposition as the binary32 symmetric pass (upper-triangle
// items, lane l on source (l + s) mod 64 at step s, source-side sums riding
// along by DPP rotation, one writer per slot), in binary64 throughout: the
// rotation moves each binary64 sum as its two 32-bit halves, and no runs are
// needed.  Shares every per-pair term between the two particles: 1.14x the
// one-sided kernel's pairs/s on the rotor wake at N = 262,144, agreeing with
// it to ~5e-12 per particle.  An option (PW_SCHEME_SYMMETRIC); the one-sided
// kernel stays the plain binary64 method.

// Steps of the rotation unrolled per loop iteration in the binary64 pass.
// Measured at N = 262,144 (256 x 1): 4 steps 100 ms, 2 steps 103 ms, 8 steps
// 134 ms (register pressure); a 3-waves-per-SIMD bound spills and is slower.
#ifndef PW_SYM64_UNROLL
#define PW_SYM64_UNROLL 4
#endif

struct Acc64s {
    double v[NCOMP_S];
};

// x of lane (l + 1) mod 64 (both halves through DPP wave_rol:1).
__device__ __forceinline__ double rot64(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x134, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0x134, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// The binary64 tile, laid out as SrcTile for the same reason; v and h stay
// separate arrays (a uniform-core job reads v alone: half of a 16-byte pair
// would be read at a 16-byte lane stride, with 2-way bank conflicts).
struct SrcTile64 {
    double2 xy[128], za[128], aa[128];  // (x, y), (z, ax), (ay, az)
    double v[128], h[128];
};

struct Part64 {
    double x, y, z, ax, ay, az, v, h;
};

__device__ __forceinline__ Part64 load_src64(const SrcTile64 &t, int k) {
    const double2 a = t.xy[k], b = t.za[k], c = t.aa[k];
    return Part64{a.x, a.y, b.x, b.y, c.x, c.y, t.v[k], t.h[k]};
}

// One unordered pair in binary64: target-side terms added to a, source-side
// terms added to b (the source's running sums, just rotated in).  MASKED:
// valid == false contributes exactly zero.  DIFF == false: a far pair of a
// uniform-core job (t >= T_FAR64): the far-field kernel, no PSE terms.
template <int KIND, bool UNIFORM, bool MASKED, bool DIFF = true>
__device__ __forceinline__ void pair64s(const Part64 &me, const Part64 &s, bool valid, Acc64s &a,
                                        Acc64s &b) {
    static_assert(DIFF || (UNIFORM && !MASKED), "the far form is for full tiles of uniform jobs");
    const double dx = me.x - s.x, dy = me.y - s.y, dz = me.z - s.z;
    const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
    double Q, Fp, Zs = 0.0;
    if (!DIFF) {
        double itm;
        kernel64_far<KIND>(r2, Q, itm);
        Fp = Q * itm;
    } else if (UNIFORM) {
        const double t = MASKED ? (valid ? r2 : 1.0) : r2;
        if (KIND == KIND_HIGHORDER)
            kernel64_scaled(t, rsq2_nr(t), Q, Fp, Zs);
        else
            kernel64<KIND>(t, rsq_nr(t), Q, Fp, Zs);
    } else {
        const double isg = rsq_nr(me.h + s.h);
        const double is2 = isg * isg;
        double t = r2 * is2;
        if (MASKED) t = valid ? t : 1.0;
        if (KIND == KIND_HIGHORDER)
            kernel64_scaled(t, rsq2_nr(t), Q, Fp, Zs);
        else
            kernel64<KIND>(t, rsq_nr(t), Q, Fp, Zs);
        const double s3 = isg * is2, s5 = s3 * is2;
        Q *= s3;
        Fp *= s5;
        Zs *= s5;
    }
    if (MASKED) {
        Q = valid ? Q : 0.0;
        Fp = valid ? Fp : 0.0;
        Zs = valid ? Zs : 0.0;
    }
    // velocity of i: a_j x u, of j: u x a_i, with u = q d
    const double ux = Q * dx, uy = Q * dy, uz = Q * dz;
    a.v[0] = __builtin_fma(s.ay, uz, __builtin_fma(-s.az, uy, a.v[0]));
    a.v[1] = __builtin_fma(s.az, ux, __builtin_fma(-s.ax, uz, a.v[1]));
    a.v[2] = __builtin_fma(s.ax, uy, __builtin_fma(-s.ay, ux, a.v[2]));
    const double ex = __builtin_fma(me.ay, s.az, -me.az * s.ay);
    const double ey = __builtin_fma(me.az, s.ax, -me.ax * s.az);
    const double ez = __builtin_fma(me.ax, s.ay, -me.ay * s.ax);
    const double fd = Fp * __builtin_fma(dz, ez, __builtin_fma(dy, ey, dx * ex));
    const double sx = __builtin_fma(fd, dx, Q * ex);
    const double sy = __builtin_fma(fd, dy, Q * ey);
    const double sz = __builtin_fma(fd, dz, Q * ez);
    a.v[3] += sx;
    a.v[4] += sy;
    a.v[5] += sz;
    if (DIFF) {
        a.v[6] = __builtin_fma(Zs, s.ax, a.v[6]);
        a.v[7] = __builtin_fma(Zs, s.ay, a.v[7]);
        a.v[8] = __builtin_fma(Zs, s.az, a.v[8]);
        a.v[9] = __builtin_fma(Zs, s.v, a.v[9]);
    }
    b.v[0] = __builtin_fma(uy, me.az, __builtin_fma(-uz, me.ay, b.v[0]));
    b.v[1] = __builtin_fma(uz, me.ax, __builtin_fma(-ux, me.az, b.v[1]));
    b.v[2] = __builtin_fma(ux, me.ay, __builtin_fma(-uy, me.ax, b.v[2]));
    b.v[3] += sx;
    b.v[4] += sy;
    b.v[5] += sz;
    if (DIFF) {
        b.v[6] = __builtin_fma(Zs, me.ax, b.v[6]);
        b.v[7] = __builtin_fma(Zs, me.ay, b.v[7]);
        b.v[8] = __builtin_fma(Zs, me.az, b.v[8]);
        b.v[9] = __builtin_fma(Zs, me.v, b.v[9]);
    }
}

// DIFF == false: a far tile in the far-field form; sums 6-9 stay untouched
// (and unrotated: they receive nothing in this tile).
template <int KIND, bool UNIFORM, bool MASKED, int TPT, int SB, bool DIFF = true>
__device__ __forceinline__ void sym64_steps(const SrcTile64 &tile, int s_begin, int lane,
                                            const Part64 *me, const bool *live, const int *nt,
                                            const bool *diag, int cnt, Acc64s *a,
                                            Acc64s &jacc) {
    const int e0 = lane + s_begin;
#pragma unroll
    for (int u = 0; u < SB; ++u) {
        const int s = s_begin + u;
        const int k = (lane + s) & 63;
        const Part64 sj = load_src64(tile, e0 + u);
        Acc64s b;
#pragma unroll
        for (int m = 0; m < NCOMP_S; ++m) b.v[m] = (DIFF || m < 6) ? rot64(jacc.v[m]) : jacc.v[m];
        if (!MASKED) {
#pragma unroll
            for (int p = 0; p < TPT; ++p) pair64s<KIND, UNIFORM, false, DIFF>(me[p], sj, true, a[p], b);
        } else {
            const bool src_ok = k < cnt;
#pragma unroll
            for (int p = 0; p < TPT; ++p) {
                if (!live[p]) continue;
                const bool valid = src_ok && lane < nt[p] && (!diag[p] || lane + s < 64) &&
                                   !(diag[p] && s == 0);
                pair64s<KIND, UNIFORM, true>(me[p], sj, valid, a[p], b);
            }
        }
        jacc = b;
    }
}

// Held to two waves per SIMD (at most 256 VGPRs): the 64 x 2 variant the tuner
// picks at N = 262,144 otherwise takes 258 and runs one wave per SIMD.
template <int KIND, int TB, int TPT>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(2))) void pair64s_kernel(const JobDev *__restrict__ jobs,
                                                     const Item *__restrict__ items,
                                                     const int *__restrict__ nitems_p,
                                                     int *__restrict__ counter) {
    constexpr int NW = TB / 64;
    constexpr int TW = TB * TPT;
    constexpr int SB = PW_SYM64_UNROLL;  // unrolled steps
    __shared__ SrcTile64 tile;
    __shared__ double jl[NW][NCOMP_S][64];
    __shared__ double prev[NCOMP_S][64];  // the strip row's sums before this block
    __shared__ float bbw[NW][6];          // per wave: box of its targets (lo xyz, hi xyz)
    __shared__ int tfar;                  // the staged tile's pairs all have t >= T_FAR64
    __shared__ int slot;
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int nitems = *nitems_p;
    int taken = 0;
    for (;;) {
        // uniform: the item's fields and its job's are scalar loads
        const int it = __builtin_amdgcn_readfirstlane(next_item(counter, &slot, taken, nitems));
        if (it >= nitems) break;
        const Item item = items[it];
        const JobDev &J = jobs[item.job];
        const Rec64 *rec = (const Rec64 *)J.rec;
        const int n = J.n;
        // one strip of blocks against one source chunk (see pair32s_kernel)
        const int bper = J.sw / TW, nblk = (n + TW - 1) / TW;
        const int b_lo = item.tile * bper, b_hi = min(b_lo + bper, nblk);
        const int S0 = item.tile * J.sw;
        const size_t len = (size_t)J.npad - (size_t)S0;
        double *jrow = J.jpart + jpart_offset((size_t)item.tile, (size_t)J.npad, (size_t)J.sw);
        for (int blk = b_lo; blk < b_hi; ++blk) {
        const int B0 = blk * TW;
        const int c0 = max(item.chunk * J.chunk_len, B0);
        const int c1 = min(item.chunk * J.chunk_len + J.chunk_len, n);
        if (c0 >= c1) break;
        const bool first = blk == b_lo;
        int idx[TPT], wt[TPT], nt[TPT];
        Part64 me[TPT];
        Acc64s a[TPT];
#pragma unroll
        for (int p = 0; p < TPT; ++p) {
            idx[p] = B0 + p * TB + (int)threadIdx.x;
            wt[p] = (B0 + p * TB) / 64 + w;
            nt[p] = max(0, min(64, n - wt[p] * 64));
            const Rec64 r = rec[idx[p] < n ? idx[p] : n - 1];
            me[p] = Part64{r.x, r.y, r.z, r.ax, r.ay, r.az, r.v, r.h};
#pragma unroll
            for (int m = 0; m < NCOMP_S; ++m) a[p].v[m] = 0.0;
        }
        {  // this wave's share of the block's box (targets past n repeat particle n - 1)
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
            for (int p = 0; p < TPT; ++p) {
                float plo[3], phi[3];
                wave_box(me[p].x, me[p].y, me[p].z, plo, phi);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    lo[c] = fminf(lo[c], plo[c]);
                    hi[c] = fmaxf(hi[c], phi[c]);
                }
            }
            if (lane == 0)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    bbw[w][c] = lo[c];
                    bbw[w][3 + c] = hi[c];
                }
        }
        const bool uniform = J.uniform != 0;
        for (int s0 = c0; s0 < c1; s0 += 64) {
            const int cnt = min(64, c1 - s0);
            const int st = s0 >> 6;
            __syncthreads();
            if (threadIdx.x < 64) {  // stage twice; missing sources read as zeros
                const int k = (int)threadIdx.x;
                Rec64 r;
                if (k < cnt) {
                    r = rec[s0 + k];
                } else {
                    r.x = r.y = r.z = r.h = r.ax = r.ay = r.az = r.v = 0.0;
                }
#pragma unroll
                for (int c = 0; c < 128; c += 64) {
                    tile.xy[k + c] = make_double2(r.x, r.y);
                    tile.za[k + c] = make_double2(r.z, r.ax);
                    tile.aa[k + c] = make_double2(r.ay, r.az);
                    tile.v[k + c] = r.v;
                    tile.h[k + c] = r.h;
                }
                // far test (full tiles of uniform-core jobs, whose coordinates
                // are in units of sigma): the tile's box against the block's
                int f = 0;
                if (cnt == 64 && uniform) {
                    float tlo[3], thi[3], blo[3], bhi[3];
                    wave_box(r.x, r.y, r.z, tlo, thi);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        blo[c] = bbw[0][c];
                        bhi[c] = bbw[0][3 + c];
                    }
#pragma unroll
                    for (int ww = 1; ww < NW; ++ww)
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            blo[c] = fminf(blo[c], bbw[ww][c]);
                            bhi[c] = fmaxf(bhi[c], bbw[ww][3 + c]);
                        }
                    f = box_gap2(blo, bhi, tlo, thi) >= T_FAR64 ? 1 : 0;
                }
                if (k == 0) tfar = f;
            }
            if (!first)  // the row's sums so far, loaded beside the tile
                for (int e = (NW > 1 ? (int)threadIdx.x - 64 : (int)threadIdx.x); e >= 0 && e < NCOMP_S * 64;
                     e += (NW > 1 ? TB - 64 : TB)) {
                    const int m = e >> 6, j = e & 63;
                    if (j < cnt) prev[m][j] = jrow[(size_t)m * len + (size_t)(s0 - S0 + j)];
                }
            __syncthreads();
            const bool farz = __builtin_amdgcn_readfirstlane(tfar) != 0;
            bool live[TPT], diag[TPT], full = cnt == 64, any = false;
#pragma unroll
            for (int p = 0; p < TPT; ++p) {
                live[p] = st >= wt[p] && nt[p] > 0;
                diag[p] = st == wt[p];
                full = full && st > wt[p] && nt[p] == 64;
                any = any || live[p];
            }
            Acc64s jacc;
#pragma unroll
            for (int m = 0; m < NCOMP_S; ++m) jacc.v[m] = 0.0;
            if (any) {
#pragma unroll 1
                for (int sb = 0; sb < 64; sb += SB) {
                    if (full && farz) {
                        sym64_steps<KIND, true, false, TPT, SB, false>(tile, sb, lane, me, live, nt,
                                                                       diag, cnt, a, jacc);
                    } else if (full) {
                        if (uniform)
                            sym64_steps<KIND, true, false, TPT, SB>(tile, sb, lane, me, live, nt,
                                                                    diag, cnt, a, jacc);
                        else
                            sym64_steps<KIND, false, false, TPT, SB>(tile, sb, lane, me, live, nt,
                                                                     diag, cnt, a, jacc);
                    } else {
                        if (uniform)
                            sym64_steps<KIND, true, true, TPT, SB>(tile, sb, lane, me, live, nt,
                                                                   diag, cnt, a, jacc);
                        else
                            sym64_steps<KIND, false, true, TPT, SB>(tile, sb, lane, me, live, nt,
                                                                    diag, cnt, a, jacc);
                    }
                }
            }
            // after 64 steps lane l holds the sums of source (l + 63) mod 64
#pragma unroll
            for (int m = 0; m < NCOMP_S; ++m) jl[w][m][(lane + 63) & 63] = jacc.v[m];
            __syncthreads();
            for (int e = threadIdx.x; e < NCOMP_S * 64; e += TB) {
                const int m = e >> 6, j = e & 63;
                if (j < cnt) {
                    double v = jl[0][m][j];
#pragma unroll
                    for (int ww = 1; ww < NW; ++ww) v += jl[ww][m][j];
                    jrow[(size_t)m * len + (size_t)(s0 - S0 + j)] = first ? v : prev[m][j] + v;
                }
            }
        }
        const size_t pl = sym_part_len((size_t)item.chunk, (size_t)J.npad, (size_t)J.chunk_len, (size_t)TW);
        double *orow = J.part + sym_part_offset((size_t)item.chunk, (size_t)J.npad, (size_t)J.chunk_len, (size_t)TW);
#pragma unroll
        for (int p = 0; p < TPT; ++p)
            if (idx[p] < n) {
#pragma unroll
                for (int m = 0; m < NCOMP_S; ++m) orow[(size_t)m * pl + idx[p]] = a[p].v[m];
            }
        }  // blocks of the strip
    }
}

// --------------------------------------------------------- combine kernel
// Per particle: add the chunk partials in chunk order (binary64), form the
// rates
//   u   = su * S[0:3] + uinf
//   da  = sq * S[3:6] + sd * (v_i S[6:9] - a_i S[9])
// (uniform core, positions prescaled by 1/sigma: su = sigma^-2 / kq,
// sq = sigma^-3 / kq, sd = 2 nu sigma^-5 / kz; otherwise the per-pair
// powers of sigma are already inside and su = sq = 1/kq, sd = 2 nu / kz;
// kq, kz from pair_scale), then apply the RK4 / Euler stage combination of rk4.cpp:21-66
// with unfused multiplies and adds in the reference's order, and write the
// next stage's record.  Block 0 also re-arms the pair kernels' queues.  With
// mode == MODE_PER_JOB each job runs the stage its JobDev names (a server pass
// in which jobs are at different RK4 stages).
template <typename Rec>
__device__ __forceinline__ void store_rec(Rec *r, int i, const double *xs, double scale);
template <>
__device__ __forceinline__ void store_rec<Rec64>(Rec64 *r, int i, const double *xs, double scale) {
    r[i].x = xs[0] * scale;
    r[i].y = xs[1] * scale;
    r[i].z = xs[2] * scale;
    r[i].ax = xs[3];
    r[i].ay = xs[4];
    r[i].az = xs[5];
}
template <>
__device__ __forceinline__ void store_rec<Rec32>(Rec32 *r, int i, const double *xs, double scale) {
    r[i].x = xs[0] * scale;
    r[i].y = xs[1] * scale;
    r[i].z = xs[2] * scale;
    r[i].ax = xs[3];
    r[i].ay = xs[4];
    r[i].az = xs[5];
    r[i].fx = (float)xs[3];
    r[i].fy = (float)xs[4];
    r[i].fz = (float)xs[5];
}

// Global-memory view of a binary64 array (lets the compiler address it as a
// wave-uniform base plus a per-lane offset).
typedef __attribute__((address_space(1))) const double gdouble;

// Rates k of particle i: written out, checked, then the stage combination
// and the next stage's record.
__device__ __forceinline__ void finish_particle(const JobDev &J, int i, const double *k, int mode) {
#pragma clang fp contract(off)
    const size_t np = (size_t)J.npad;
    bool finite = true;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        J.rate[m * np + i] = k[m];
        finite = finite && __builtin_isfinite(k[m]);
    }
    if (!finite) atomicOr(J.flag, 1);
    if (mode == MODE_SOLVE) return;
    const double dt = J.dt;
    double nx[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const size_t o = m * np + i;
        const double x0 = J.x0[o];
        if (mode == MODE_RK4_1) {
            J.k1[o] = k[m];
            nx[m] = k[m] * (0.5 * dt) + x0;
        } else if (mode == MODE_RK4_2) {
            J.k23[o] = k[m];
            nx[m] = k[m] * (0.5 * dt) + x0;
        } else if (mode == MODE_RK4_3) {
            J.k23[o] = J.k23[o] + k[m];
            nx[m] = k[m] * dt + x0;
        } else if (mode == MODE_RK4_4) {
            const double a = (J.k1[o] + k[m]) * (dt / 6.);
            const double b = J.k23[o] * (dt / 3.);
            nx[m] = x0 + (a + b);
            J.x0[o] = nx[m];
        } else {  // MODE_EULER
            nx[m] = x0 + k[m] * dt;
            J.x0[o] = nx[m];
        }
        J.xs[o] = nx[m];
        finite = finite && __builtin_isfinite(nx[m]);
    }
    if (!finite) atomicOr(J.flag, 1);
    if (J.prec == PREC_FP64)
        store_rec<Rec64>((Rec64 *)J.rec, i, nx, J.rec_scale);
    else
        store_rec<Rec32>((Rec32 *)J.rec, i, nx, J.rec_scale);
}

// Items: 256 particles of a one-sided job (one thread each), or 64 particles
// of a symmetric job.  A symmetric item's particles read up to N / tw
// source-side slot rows each (10 GB per pass at N = 262,144): its four waves
// take every fourth row and every fourth target-side chunk, and the four
// wave sums are added in wave order, so that the item with the most rows is
// not left to one wave (one thread per particle over all rows took 2.2 ms
// per pass at N = 262,144, the last block's rows alone).
__global__ __launch_bounds__(256) void combine_kernel(const JobDev *__restrict__ jobs,
                                                      const CItem *__restrict__ citems,
                                                      const int *__restrict__ ncitems_p,
                                                      int *__restrict__ counters, int mode) {
#pragma clang fp contract(off)
    __shared__ double red[3][NCOMP_S][64];  // waves 1-3 of a symmetric item
    const int nci = *ncitems_p;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[0] = 0;
        counters[1] = 0;
    }
    for (int ci = blockIdx.x; ci < nci; ci += gridDim.x) {
        // the item is the block's: its fields (and the job's) are uniform
        CItem c = citems[ci];
        c.job = __builtin_amdgcn_readfirstlane(c.job);
        c.tile = __builtin_amdgcn_readfirstlane(c.tile);
        const JobDev &J = jobs[c.job];
        const size_t np = (size_t)J.npad;
        const int jmode = mode == MODE_PER_JOB ? J.mode : mode;
        if (J.sym) {
            // target side: the chunks from the one holding this particle's
            // block start; source side: the strips up to this particle's
            // own.  Every particle of the item is in the same block and
            // strip (64 divides tw, tw divides sw), so the loops and row
            // addresses are uniform per wave; lanes past n read padding
            // (inside every row) and are dropped at the end.
            const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);  // the wave
            const int lane = (int)threadIdx.x & 63;
            const int i0 = c.tile * 64, i = i0 + lane;
            const bool live = i < J.n;
            const size_t tw = (size_t)J.tw, sw = (size_t)J.sw, L = (size_t)J.chunk_len;
            const int blk = i0 / J.tw;   // this item's block
            const int sblk = i0 / J.sw;  // and strip
            const int cfirst = (int)(((size_t)blk * tw) / L);
            double S[NCOMP_S], T[NCOMP_S];
#pragma unroll
            for (int m = 0; m < NCOMP_S; ++m) S[m] = T[m] = 0.0;
            for (int ch = cfirst + w; ch < J.nchunk; ch += 4) {
                const size_t pl = sym_part_len((size_t)ch, np, L, tw);
                const gdouble *p =
                    (const gdouble *)(J.part + sym_part_offset((size_t)ch, np, L, tw) + i0) + lane;
#pragma unroll
                for (int m = 0; m < NCOMP_S; ++m) S[m] += p[m * pl];
            }
            // source side: the rows of the strips up to this item's; two rows
            // in flight per step (added in row order)
            int b = w;
            for (; b + 4 <= sblk; b += 8) {
                const size_t len0 = np - (size_t)b * sw, len1 = len0 - 4 * sw;
                const gdouble *q0 =
                    (const gdouble *)(J.jpart + jpart_offset((size_t)b, np, sw) + (i0 - (size_t)b * sw)) +
                    lane;
                const gdouble *q1 = (const gdouble *)(J.jpart + jpart_offset((size_t)b + 4, np, sw) +
                                                      (i0 - ((size_t)b + 4) * sw)) +
                                    lane;
                double v0[NCOMP_S], v1[NCOMP_S];
#pragma unroll
                for (int m = 0; m < NCOMP_S; ++m) {
                    v0[m] = q0[m * len0];
                    v1[m] = q1[m * len1];
                }
#pragma unroll
                for (int m = 0; m < NCOMP_S; ++m) T[m] = (T[m] + v0[m]) + v1[m];
            }
            if (b <= sblk) {
                const size_t len = np - (size_t)b * sw;
                const gdouble *q =
                    (const gdouble *)(J.jpart + jpart_offset((size_t)b, np, sw) + (i0 - (size_t)b * sw)) +
                    lane;
#pragma unroll
                for (int m = 0; m < NCOMP_S; ++m) T[m] += q[m * len];
            }
            // source-side stretching enters with a minus sign
            double U[NCOMP_S];
#pragma unroll
            for (int m = 0; m < NCOMP_S; ++m) U[m] = (m >= 3 && m < 6) ? S[m] - T[m] : S[m] + T[m];
            if (w > 0) {
#pragma unroll
                for (int m = 0; m < NCOMP_S; ++m) red[w - 1][m][lane] = U[m];
            }
            __syncthreads();
            if (w == 0 && live) {
#pragma unroll
                for (int m = 0; m < NCOMP_S; ++m)
                    U[m] = ((U[m] + red[0][m][lane]) + red[1][m][lane]) + red[2][m][lane];
                const double vi = J.vol[i];
                const double ax = J.xs[3 * np + i], ay = J.xs[4 * np + i], az = J.xs[5 * np + i];
                double k[6];
                k[0] = J.su * U[0] + J.uinf[0];
                k[1] = J.su * U[1] + J.uinf[1];
                k[2] = J.su * U[2] + J.uinf[2];
                k[3] = J.sq * U[3] + J.sd * (vi * U[6] - ax * U[9]);
                k[4] = J.sq * U[4] + J.sd * (vi * U[7] - ay * U[9]);
                k[5] = J.sq * U[5] + J.sd * (vi * U[8] - az * U[9]);
                finish_particle(J, i, k, jmode);
            }
            __syncthreads();  // red is rewritten by the next symmetric item
            continue;
        }
        const int i = c.tile * 256 + (int)threadIdx.x;
        if (i >= J.n) continue;
        const double vi = J.vol[i];
        const double ax = J.xs[3 * np + i], ay = J.xs[4 * np + i], az = J.xs[5 * np + i];
        double S[NCOMP];
#pragma unroll
        for (int m = 0; m < NCOMP; ++m) S[m] = 0.0;
        for (int ch = 0; ch < J.nchunk; ++ch) {
            const double *p = J.part + (size_t)ch * NCOMP * np + i;
#pragma unroll
            for (int m = 0; m < NCOMP; ++m) S[m] += p[m * np];
        }
        double k[6];
        k[0] = J.su * S[0] + J.uinf[0];
        k[1] = J.su * S[1] + J.uinf[1];
        k[2] = J.su * S[2] + J.uinf[2];
        k[3] = J.sq * S[3] + J.sd * (vi * S[6] - ax * S[9]);
        k[4] = J.sq * S[4] + J.sd * (vi * S[7] - ay * S[9]);
        k[5] = J.sq * S[5] + J.sd * (vi * S[8] - az * S[9]);
        finish_particle(J, i, k, jmode);
    }
}

// Write the records of the current stage state xs (after a state upload).
__global__ __launch_bounds__(256) void refresh_rec_kernel(JobDev J) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= J.n) return;
    const size_t np = (size_t)J.npad;
    double xs[6];
    for (int m = 0; m < 6; ++m) xs[m] = J.xs[m * np + i];
    if (J.prec == PREC_FP64)
        store_rec<Rec64>((Rec64 *)J.rec, i, xs, J.rec_scale);
    else
        store_rec<Rec32>((Rec32 *)J.rec, i, xs, J.rec_scale);
}

// Copy x0 -> xs (a new step or a new state starts from x0).
__global__ __launch_bounds__(256) void copy_soa_kernel(const double *__restrict__ src,
                                                       double *__restrict__ dst, size_t count) {
    for (size_t k = blockIdx.x * 256 + threadIdx.x; k < count; k += (size_t)gridDim.x * 256)
        dst[k] = src[k];
}

// ------------------------------------------------ state layout conversion
// Reference state layout (wake.cpp:87-120): for each wake w, starting at
// 6*off[w]: positions 3*n_w (row-major N x 3) then vorticities 3*n_w.
// SoA: [6][npad].  dir = 0: state -> SoA, dir = 1: SoA -> state.
__global__ __launch_bounds__(256) void state_layout_kernel(double *__restrict__ state,
                                                           double *__restrict__ soa, int npad,
                                                           const int64_t *__restrict__ woff,
                                                           int nwake, int dir) {
    const int64_t ntot = woff[nwake];
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < ntot; i += (int64_t)gridDim.x * 256) {
        int w = 0;
        while (w + 1 < nwake && i >= woff[w + 1]) ++w;
        const int64_t k = i - woff[w];
        const int64_t nw = woff[w + 1] - woff[w];
        const int64_t base = 6 * woff[w];
        for (int c = 0; c < 3; ++c) {
            const int64_t ip = base + 3 * k + c, iv = base + 3 * nw + 3 * k + c;
            if (dir == 0) {
                soa[(size_t)c * npad + i] = state[ip];
                soa[(size_t)(3 + c) * npad + i] = state[iv];
            } else {
                state[ip] = soa[(size_t)c * npad + i];
                state[iv] = soa[(size_t)(3 + c) * npad + i];
            }
        }
    }
}

// N x 3 row-major <-> SoA component block ([3][npad]).
__global__ __launch_bounds__(256) void aos_soa_kernel(double *__restrict__ aos,
                                                      double *__restrict__ soa, int n, int npad,
                                                      int dir) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        for (int c = 0; c < 3; ++c) {
            if (dir == 0)
                soa[(size_t)c * npad + i] = aos[3 * (size_t)i + c];
            else
                aos[3 * (size_t)i + c] = soa[(size_t)c * npad + i];
        }
}

// ------------------------------------------------------------- diagnostics
// Per 256-particle tile: sum a (3), sum x x a (3) and sum x x (x x a) (3),
// written to out[tile][9] (interaction.cpp:240-275 before scaling).
__global__ __launch_bounds__(256) void linear_diag_kernel(const double *__restrict__ xs, int n,
                                                          int npad, double *__restrict__ out) {
    __shared__ double red[9][256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < n) {
        const double x = xs[i], y = xs[npad + i], z = xs[2 * (size_t)npad + i];
        const double ax = xs[3 * (size_t)npad + i], ay = xs[4 * (size_t)npad + i],
                     az = xs[5 * (size_t)npad + i];
        v[0] = ax;
        v[1] = ay;
        v[2] = az;
        v[3] = y * az - z * ay;
        v[4] = z * ax - x * az;
        v[5] = x * ay - y * ax;
        v[6] = y * v[5] - z * v[4];
        v[7] = z * v[3] - x * v[5];
        v[8] = x * v[4] - y * v[3];
    }
    for (int m = 0; m < 9; ++m) red[m][threadIdx.x] = v[m];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int m = 0; m < 9; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 9) out[blockIdx.x * 9 + threadIdx.x] = red[threadIdx.x][0];
}

// Quadratic diagnostics per target, summed over one source chunk
// (blockIdx.y); out is [nchunk][3][npad]:
//   0  kinetic energy   e_i = a_i.a_i/(8 pi s_i) + sum_{j!=i} ((r^2 + 2 sig^2)(a_i.a_j)
//                             + (d.a_i)(d.a_j)) / (16 pi (r^2 + sig^2)^1.5)
//                       (KINETICENERGY, interaction_utils.h:197-240); KE = 2 sum_i e_i
//   1  enstrophy        (5/4 pi) a_i.a_i / s_i^3 + sum_{j!=i} F1 (a_i.a_j) + F2 (a_i.d)(a_j.d),
//                       F1 = (10 sig^4 - 2 r^4 - 7 sig^2 r^2)/(8 pi (r^2+sig^2)^3.5),
//                       F2 = -15 r^2/(8 pi (r^2+sig^2)^3.5)   (ENST, highorder.h:69-86;
//                       ENSTROPHY, interaction_utils.h:107-156); zero for GAUSSIAN (gaussian.h:62-81)
//   2  helicity         sum_{j!=i} q(r, sig) d.(a_i x a_j)   (HELICITY, interaction_utils.h:167-190)
// with d = x_j - x_i, sig^2 = (s_i^2 + s_j^2)/2.  Chunk 0 carries the self terms.
template <int KIND>
__global__ __launch_bounds__(256) void diag_pair_kernel(const double *__restrict__ xs,
                                                        const double *__restrict__ rad, int n,
                                                        int npad, int chunk_len,
                                                        double *__restrict__ out) {
    __shared__ double sh[7][256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int ii = i < n ? i : n - 1;
    const double x = xs[ii], y = xs[npad + ii], z = xs[2 * (size_t)npad + ii];
    const double ax = xs[3 * (size_t)npad + ii], ay = xs[4 * (size_t)npad + ii],
                 az = xs[5 * (size_t)npad + ii];
    const double si = rad[ii];
    const double hs = 0.5 * si * si;
    const int c0 = blockIdx.y * chunk_len, c1 = min(c0 + chunk_len, n);
    double e = 0.0, en = 0.0, he = 0.0;
    if (blockIdx.y == 0) {
        const double a2 = ax * ax + ay * ay + az * az;
        e = a2 / (8.0 * K_PI * si);
        if (KIND == KIND_HIGHORDER) en = 5.0 / (4.0 * K_PI) * a2 / (si * si * si);
    }
    for (int s0 = c0; s0 < c1; s0 += 256) {
        const int cnt = min(256, c1 - s0);
        __syncthreads();
        const int j = s0 + (int)threadIdx.x;
        if ((int)threadIdx.x < cnt) {
            for (int m = 0; m < 6; ++m) sh[m][threadIdx.x] = xs[(size_t)m * npad + j];
            sh[6][threadIdx.x] = 0.5 * rad[j] * rad[j];
        }
        __syncthreads();
        for (int k = 0; k < cnt; ++k) {
            if (s0 + k == i) continue;
            const double dx = sh[0][k] - x, dy = sh[1][k] - y, dz = sh[2][k] - z;
            const double bx = sh[3][k], by = sh[4][k], bz = sh[5][k];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double sg2 = hs + sh[6][k];
            const double aa = ax * bx + ay * by + az * bz;
            const double da = dx * ax + dy * ay + dz * az;
            const double db = dx * bx + dy * by + dz * bz;
            const double w = rsq_nr(r2 + sg2);
            const double w3 = w * w * w;
            e += ((r2 + 2.0 * sg2) * aa + da * db) * w3 * (1.0 / (16.0 * K_PI));
            if (KIND == KIND_HIGHORDER) {
                const double f = w3 * w3 * w * (1.0 / (8.0 * K_PI));
                en += f * ((10.0 * sg2 * sg2 - 2.0 * r2 * r2 - 7.0 * sg2 * r2) * aa -
                           15.0 * r2 * da * db);
            }
            const double isg = rsq_nr(sg2);
            const double is2 = isg * isg;
            const double t = r2 * is2;
            double Q, Fp, Zs;
            kernel64<KIND>(t, rsq_nr(t), Q, Fp, Zs);
            const double ex = ay * bz - az * by, ey = az * bx - ax * bz, ez = ax * by - ay * bx;
            he += Q * (isg * is2) * (dx * ex + dy * ey + dz * ez);
        }
    }
    if (i < n) {
        double *o = out + (size_t)blockIdx.y * 3 * npad + i;
        o[0] = e;
        o[npad] = en;
        o[2 * (size_t)npad] = he;
    }
}

// Vorticity field (influence(), interaction.cpp:110-155): for every target
//   k_i = sum_j Z(rho_ij, sigma_ij) a_j,  sigma_ij^2 = (s_i^2 + s_j^2)/2,
// the j = i term being SELFINFLUENCE's Z(0, s_i) a_i (interaction_utils.h:592-605).
// out is [nchunk][3][npad].
template <int KIND>
__global__ __launch_bounds__(256) void influence_kernel(const double *__restrict__ xs,
                                                        const double *__restrict__ rad, int n,
                                                        int npad, int chunk_len,
                                                        double *__restrict__ out) {
    __shared__ double sh[7][256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int ii = i < n ? i : n - 1;
    const double x = xs[ii], y = xs[npad + ii], z = xs[2 * (size_t)npad + ii];
    const double hs = 0.5 * rad[ii] * rad[ii];
    const int c0 = blockIdx.y * chunk_len, c1 = min(c0 + chunk_len, n);
    double kx = 0.0, ky = 0.0, kz = 0.0;
    for (int s0 = c0; s0 < c1; s0 += 256) {
        const int cnt = min(256, c1 - s0);
        __syncthreads();
        const int j = s0 + (int)threadIdx.x;
        if ((int)threadIdx.x < cnt) {
            for (int m = 0; m < 6; ++m) sh[m][threadIdx.x] = xs[(size_t)m * npad + j];
            sh[6][threadIdx.x] = 0.5 * rad[j] * rad[j];
        }
        __syncthreads();
        for (int k = 0; k < cnt; ++k) {
            const double dx = sh[0][k] - x, dy = sh[1][k] - y, dz = sh[2][k] - z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double sg2 = hs + sh[6][k];
            const double isg = rsq_nr(sg2);
            const double is2 = isg * isg;
            const double t = r2 * is2;
            double Zs;
            if (KIND == KIND_HIGHORDER) {
                const double w = rsq_nr(t + 1.0);
                const double iu = w * w, iu2 = iu * iu, iu4 = iu2 * iu2;
                Zs = K_CZ * (iu4 * iu2 * iu * w);
            } else {
                Zs = K_CG * exp(-0.5 * t);
            }
            const double Z = Zs * isg * is2;
            kx += Z * sh[3][k];
            ky += Z * sh[4][k];
            kz += Z * sh[5][k];
        }
    }
    if (i < n) {
        double *o = out + (size_t)blockIdx.y * 3 * npad + i;
        o[0] = kx;
        o[npad] = ky;
        o[2 * (size_t)npad] = kz;
    }
}

// out[m][i] = sum over chunks c (in order) of in[c][m][i], m < ncomp, i < n.
__global__ __launch_bounds__(256) void chunk_sum_kernel(const double *__restrict__ in,
                                                        double *__restrict__ out, int nchunk,
                                                        int ncomp, int n, int npad) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int m = 0; m < ncomp; ++m) {
        double s = 0.0;
        for (int c = 0; c < nchunk; ++c) s += in[((size_t)c * ncomp + m) * npad + i];
        out[(size_t)m * npad + i] = s;
    }
}

}  // namespace pw
