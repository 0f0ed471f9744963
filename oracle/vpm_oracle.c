/*
 * vpm_oracle.c -- CPU oracle for the vortex-particle interaction hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker and the timed CPU baseline, never the
 * product.  Every function cites the reference file:line whose arithmetic it
 * restates.  Operation order follows the reference so that rounding agrees
 * with it; the reference's GSL calls are restated by the small helpers below
 * (vector add/sub/scale/copy, ddot, the reference-BLAS dnrm2 and gsl_cross).
 *
 * Third-party boundary: GSL (version not pinned by the reference,
 * CMakeLists.txt:20-21).  Only these elementary GSL routines are on the path:
 *   gsl_blas_dnrm2  -> orc_dnrm2   (scaled sum of squares, reference BLAS)
 *   gsl_blas_ddot   -> dot3        (left-to-right sum from 0.0)
 *   gsl_blas_dscal, gsl_vector_scale/add/sub/memcpy -> elementwise loops
 *   gsl_cross (src/utils/gsl_utils.h:140-152) -> cross3
 */
#include "vpm_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_1_PI
#define M_1_PI 0.31830988618379067154
#endif
#ifndef M_SQRT1_2
#define M_SQRT1_2 0.70710678118654752440
#endif

/* ------------------------------------------------------------------ helpers */

/* gsl_cross, src/utils/gsl_utils.h:140-152: c = a x b, components in the
 * order (ay*bz - az*by, az*bx - ax*bz, ax*by - ay*bx). */
static inline void cross3(const double *a, const double *b, double *c) {
    double ax = a[0], ay = a[1], az = a[2];
    double bx = b[0], by = b[1], bz = b[2];
    c[0] = ay * bz - az * by;
    c[1] = az * bx - ax * bz;
    c[2] = ax * by - ay * bx;
}

/* gsl_blas_ddot on 3-vectors: r = 0; r += x[i]*y[i] for i = 0,1,2 */
static inline double dot3(const double *x, const double *y) {
    double r = 0.0;
    for (int i = 0; i < 3; ++i) r += x[i] * y[i];
    return r;
}

/* gsl_blas_dnrm2 (GSL cblas source_nrm2_r.h, the reference-BLAS algorithm):
 * keep a running scale = max |x_i| and ssq = sum (x_i/scale)^2, rescaling ssq
 * whenever a larger element appears; return scale*sqrt(ssq).  A one-element
 * vector returns |x_0|. */
double orc_dnrm2(size_t n, const double *x) {
    if (n == 0) return 0.0;
    if (n == 1) return fabs(x[0]);
    double scale = 0.0, ssq = 1.0;
    for (size_t i = 0; i < n; ++i) {
        if (x[i] == 0.0) continue;
        double ax = fabs(x[i]);
        if (scale < ax) {
            double t = scale / ax;
            ssq = 1.0 + ssq * t * t;
            scale = ax;
        } else {
            double t = ax / scale;
            ssq += t * t;
        }
    }
    return scale * sqrt(ssq);
}

/* ------------------------------------------------------------------ kernels */

/* ZETASIG, highorder.h:20-23 and gaussian.h:20-23 */
static double zeta_sig(int kind, double rho, double sigma) {
    double rb = rho / sigma;
    if (kind == ORC_KERNEL_GAUSSIAN)
        return exp(-0.5 * rb * rb) / pow(sigma, 3.0) / pow(2.0 * M_PI, 1.5);
    return 1.875 * M_1_PI / pow(rb * rb + 1.0, 7.5) / pow(sigma, 3);
}

/* KERNEL, highorder.h:48-60 (HIGHORDER) and gaussian.h:48-58 (GAUSSIAN):
 *   Z   = zeta(rho/sigma) / sigma^3
 *   phi = HIGHORDER: (0.25/pi)(rb^2 + 1.5)/(rb^2 + 1)^2 / sigma^3
 *         GAUSSIAN : (0.25/pi) erf(rb/sqrt 2) / sigma^3
 *   q   = (phi/rb - Z)/rb^2
 *   F   = (Z - 3q)/rho^2                                                  */
void orc_kernel(int kind, double rho, double sigma, double *q, double *F, double *Z) {
    double rho_bar = rho / sigma;
    double z = zeta_sig(kind, rho, sigma);
    double rho_bar2 = rho_bar * rho_bar;
    double sig3 = sigma * sigma * sigma;
    double phi;
    if (kind == ORC_KERNEL_GAUSSIAN)
        phi = 0.25 * M_1_PI * erf(M_SQRT1_2 * rho_bar) / sig3;
    else
        phi = 0.25 * M_1_PI * (rho_bar2 + 1.5) / pow(rho_bar2 + 1.0, 2.0) / sig3;
    *Z = z;
    *q = (phi / rho_bar - z) / (rho_bar * rho_bar);
    *F = (z - 3 * (*q)) / (rho * rho);
}

/* Independent long-double evaluation of the same closed forms. */
void orc_kernel_ld(int kind, long double rho, long double sigma, long double *q,
                   long double *F, long double *Z) {
    const long double pi = 3.141592653589793238462643383279502884L;
    long double rb = rho / sigma;
    long double rb2 = rb * rb;
    long double s3 = sigma * sigma * sigma;
    long double z, phi;
    if (kind == ORC_KERNEL_GAUSSIAN) {
        z = expl(-0.5L * rb2) / s3 / powl(2.0L * pi, 1.5L);
        phi = 0.25L / pi * erfl(rb / sqrtl(2.0L)) / s3;
    } else {
        z = 1.875L / pi / powl(rb2 + 1.0L, 7.5L) / s3;
        phi = 0.25L / pi * (rb2 + 1.5L) / ((rb2 + 1.0L) * (rb2 + 1.0L)) / s3;
    }
    *Z = z;
    *q = (phi / rb - z) / rb2;
    *F = (z - 3.0L * (*q)) / (rho * rho);
}

/* -------------------------------------------------------------- pair terms */

/* Everything one unordered pair contributes, in the reference's order.
 * disp = r_trg - r_src; dr_t = VELOCITY(q, a_src, disp);
 * dr_s = VELOCITY(-q, a_trg, disp); da = VORSTRETCH + DIFFUSION.
 * interaction_utils.h:52-58 (VELOCITY), :69-100 (VORSTRETCH),
 * :253-286 (DIFFUSION), :451-469 (displacement, rho, sigma, kernel). */
typedef struct {
    double dr_t[3], dr_s[3], da[3];
    /* da split into its stretching (VORSTRETCH) and diffusion (DIFFUSION)
     * parts; st + df rounds to da exactly, since da is formed as
     * ((0 + crossed) + stretch) + dva */
    double st[3], df[3];
} pair_out;

static inline void pair_eval(int kind, double nu, double s_src, double s_trg,
                             const double *r_src, const double *r_trg, const double *a_src,
                             const double *a_trg, double v_src, double v_trg, pair_out *o) {
    double disp[3];
    for (int i = 0; i < 3; ++i) disp[i] = r_trg[i] - r_src[i];
    double rho = orc_dnrm2(3, disp);
    double q, F, Z;
    double sigma = sqrt(0.5 * (s_src * s_src + s_trg * s_trg));
    orc_kernel(kind, rho, sigma, &q, &F, &Z);
    /* VELOCITY: cross then scale */
    cross3(a_src, disp, o->dr_t);
    for (int i = 0; i < 3; ++i) o->dr_t[i] = q * o->dr_t[i];
    cross3(a_trg, disp, o->dr_s);
    for (int i = 0; i < 3; ++i) o->dr_s[i] = -q * o->dr_s[i];
    /* VORSTRETCH: da = 0 + q*(a_trg x a_src) + (F*(disp.(a_trg x a_src)))*disp */
    double txs[3], crossed[3], stretch[3];
    cross3(a_trg, a_src, txs);
    for (int i = 0; i < 3; ++i) crossed[i] = q * txs[i];
    double roaxa = dot3(disp, txs);
    double fr = F * roaxa;
    for (int i = 0; i < 3; ++i) stretch[i] = fr * disp[i];
    for (int i = 0; i < 3; ++i) o->da[i] = 0.0;
    for (int i = 0; i < 3; ++i) o->da[i] += crossed[i];
    for (int i = 0; i < 3; ++i) o->da[i] += stretch[i];
    for (int i = 0; i < 3; ++i) o->st[i] = o->da[i];
    /* DIFFUSION: dva = (v_trg*a_src - v_src*a_trg) * (Z*nu/(0.5*sigma*sigma)) */
    double sig12 = 0.5 * sigma * sigma;
    double k = Z * nu / sig12;
    double dva[3];
    for (int i = 0; i < 3; ++i) {
        double va12 = v_trg * a_src[i];
        double va21 = v_src * a_trg[i];
        dva[i] = va12;
        dva[i] -= va21;
        dva[i] = k * dva[i];
    }
    for (int i = 0; i < 3; ++i) o->da[i] += dva[i];
    for (int i = 0; i < 3; ++i) o->df[i] = dva[i];
}

/* --- heap-temporary emulation of the reference's per-pair gsl_vector use ---
 * gsl_vector_calloc(3) allocates a vector header, a block header and the data
 * array, and zeroes the data; gsl_vector_free releases all three.  The
 * reference's OpenMP INTERACT (interaction_utils.h:429-481) makes 3 such
 * vectors and its VORSTRETCH and DIFFUSION calls make 3 each, so 9 per pair.
 * The faithful CPU baseline repeats that allocation traffic and routes the
 * arithmetic through non-inlined helpers, as calls into libgsl are. */
typedef struct {
    size_t size;
    double *data;
} hblock;
typedef struct {
    size_t size, stride;
    double *data;
    hblock *block;
    int owner;
} hvec;

__attribute__((noinline)) static hvec *hv_calloc(size_t n) {
    hblock *b = (hblock *)malloc(sizeof(hblock));
    b->data = (double *)malloc(n * sizeof(double));
    b->size = n;
    hvec *v = (hvec *)malloc(sizeof(hvec));
    v->size = n;
    v->stride = 1;
    v->data = b->data;
    v->block = b;
    v->owner = 1;
    for (size_t i = 0; i < n; ++i) v->data[i] = 0.0;
    return v;
}
__attribute__((noinline)) static void hv_free(hvec *v) {
    free(v->block->data);
    free(v->block);
    free(v);
}
__attribute__((noinline)) static void hv_copy(hvec *d, const double *s) {
    for (int i = 0; i < 3; ++i) d->data[i] = s[i];
}
__attribute__((noinline)) static void hv_scale(double a, hvec *x) {
    for (size_t i = 0; i < x->size; ++i) x->data[i] = a * x->data[i];
}
__attribute__((noinline)) static void hv_add(hvec *a, const hvec *b) {
    for (size_t i = 0; i < a->size; ++i) a->data[i] += b->data[i];
}
__attribute__((noinline)) static void hv_sub(hvec *a, const hvec *b) {
    for (size_t i = 0; i < a->size; ++i) a->data[i] -= b->data[i];
}
__attribute__((noinline)) static void hv_cross(const double *a, const double *b, hvec *c) {
    if (c->size == 3) cross3(a, b, c->data);
}
__attribute__((noinline)) static double hv_dot(const hvec *x, const hvec *y) {
    return dot3(x->data, y->data);
}

/* The OpenMP INTERACT of interaction_utils.h:429-481 with its heap traffic:
 * VELOCITY/VORSTRETCH/DIFFUSION through temporaries, same arithmetic order as
 * pair_eval. */
static void pair_eval_heap(int kind, double nu, double s_src, double s_trg,
                           const double *r_src, const double *r_trg, const double *a_src,
                           const double *a_trg, double v_src, double v_trg, pair_out *o) {
    hvec *disp = hv_calloc(3);
    hv_copy(disp, r_trg);
    for (int i = 0; i < 3; ++i) disp->data[i] -= r_src[i];
    double rho = orc_dnrm2(3, disp->data);
    double q, F, Z;
    double sigma = sqrt(0.5 * (s_src * s_src + s_trg * s_trg));
    hvec *dr = hv_calloc(3);
    orc_kernel(kind, rho, sigma, &q, &F, &Z);
    hv_cross(a_src, disp->data, dr);
    hv_scale(q, dr);
    memcpy(o->dr_t, dr->data, sizeof o->dr_t);
    hv_cross(a_trg, disp->data, dr);
    hv_scale(-q, dr);
    memcpy(o->dr_s, dr->data, sizeof o->dr_s);
    hvec *da = hv_calloc(3);
    { /* VORSTRETCH */
        hvec *txs = hv_calloc(3);
        hv_cross(a_trg, a_src, txs);
        hvec *crossed = hv_calloc(3);
        hvec *stretch = hv_calloc(3);
        hv_copy(crossed, txs->data);
        hv_scale(q, crossed);
        double roaxa = hv_dot(disp, txs);
        hv_copy(stretch, disp->data);
        hv_scale(F * roaxa, stretch);
        hv_add(da, crossed);
        hv_add(da, stretch);
        hv_free(txs);
        hv_free(crossed);
        hv_free(stretch);
    }
    { /* DIFFUSION */
        hvec *va12 = hv_calloc(3);
        hvec *va21 = hv_calloc(3);
        hvec *dva = hv_calloc(3);
        hv_copy(va12, a_src);
        hv_scale(v_trg, va12);
        hv_copy(va21, a_trg);
        hv_scale(v_src, va21);
        double sig12 = 0.5 * sigma * sigma;
        hv_copy(dva, va12->data);
        hv_sub(dva, va21);
        hv_scale(Z * nu / sig12, dva);
        hv_add(da, dva);
        hv_free(va12);
        hv_free(va21);
        hv_free(dva);
    }
    memcpy(o->da, da->data, sizeof o->da);
    hv_free(dr);
    hv_free(da);
    hv_free(disp);
}

void orc_pair(int kind, double nu, double s_src, double s_trg, const double r_src[3],
              const double r_trg[3], const double a_src[3], const double a_trg[3],
              double v_src, double v_trg, double dr_trg[3], double da_trg[3],
              double dr_src_out[3], double da_src_out[3]) {
    pair_out o;
    pair_eval(kind, nu, s_src, s_trg, r_src, r_trg, a_src, a_trg, v_src, v_trg, &o);
    for (int i = 0; i < 3; ++i) {
        dr_trg[i] += o.dr_t[i];
        da_trg[i] += o.da[i];
        dr_src_out[i] = o.dr_s[i];
        da_src_out[i] = -o.da[i];
    }
}

/* -------------------------------------------------------- all-pairs loops */

/* Serial symmetric sweep over a source range [s0,s1) and target range
 * [t0,t1); self == 1 restricts targets to i_trg > i_src (interact(W),
 * interaction.cpp:157-175), self == 0 takes every target (interact(W1,W2),
 * interaction.cpp:177-195).  Source rows accumulate in place, as the
 * two-sided INTERACT does (interaction_utils.h:363-407). */
static void serial_block(int kind, double nu, size_t s0, size_t s1, size_t t0, size_t t1,
                         int self, const double *pos, const double *vort, const double *rad,
                         const double *vol, double *vel, double *dvort) {
    for (size_t is = s0; is < s1; ++is) {
        size_t tb = self ? is + 1 : t0;
        for (size_t it = tb; it < t1; ++it) {
            pair_out o;
            pair_eval(kind, nu, rad[is], rad[it], pos + 3 * is, pos + 3 * it, vort + 3 * is,
                      vort + 3 * it, vol[is], vol[it], &o);
            for (int c = 0; c < 3; ++c) vel[3 * it + c] += o.dr_t[c];
            for (int c = 0; c < 3; ++c) vel[3 * is + c] += o.dr_s[c];
            for (int c = 0; c < 3; ++c) dvort[3 * it + c] += o.da[c];
            for (int c = 0; c < 3; ++c) dvort[3 * is + c] -= o.da[c];
        }
    }
}

/* OpenMP sweep of parallel.cpp:13-48 (self) / :50-85 (cross): per source row,
 * a parallel-for over its targets with a reduction of the source's parts,
 * which are added to the source row after the loop. */
static void omp_row(int faithful, int kind, double nu, size_t is, size_t tb, size_t t1,
                    const double *pos, const double *vort, const double *rad,
                    const double *vol, double *vel, double *dvort) {
    double vx = 0.0, vy = 0.0, vz = 0.0, qx = 0.0, qy = 0.0, qz = 0.0;
    const double *r_src = pos + 3 * is, *a_src = vort + 3 * is;
    double s_src = rad[is], v_src = vol[is];
#pragma omp parallel for reduction(+ : vx, vy, vz, qx, qy, qz)
    for (size_t it = tb; it < t1; ++it) {
        pair_out o;
        if (faithful)
            pair_eval_heap(kind, nu, s_src, rad[it], r_src, pos + 3 * it, a_src, vort + 3 * it,
                           v_src, vol[it], &o);
        else
            pair_eval(kind, nu, s_src, rad[it], r_src, pos + 3 * it, a_src, vort + 3 * it,
                      v_src, vol[it], &o);
        for (int c = 0; c < 3; ++c) vel[3 * it + c] += o.dr_t[c];
        for (int c = 0; c < 3; ++c) dvort[3 * it + c] += o.da[c];
        vx += o.dr_s[0];
        vy += o.dr_s[1];
        vz += o.dr_s[2];
        qx += -o.da[0];
        qy += -o.da[1];
        qz += -o.da[2];
    }
    vel[3 * is + 0] = vx + vel[3 * is + 0];
    vel[3 * is + 1] = vy + vel[3 * is + 1];
    vel[3 * is + 2] = vz + vel[3 * is + 2];
    dvort[3 * is + 0] = qx + dvort[3 * is + 0];
    dvort[3 * is + 1] = qy + dvort[3 * is + 1];
    dvort[3 * is + 2] = qz + dvort[3 * is + 2];
}

/* One-sided per-target sum in long double: for target i,
 *   u_i  = sum_{j!=i} q (a_j x d)
 *   da_i = sum_{j!=i} q (a_i x a_j) + F (d.(a_i x a_j)) d
 *                     + (nu Z/(sigma^2/2)) (v_i a_j - v_j a_i),  d = x_i - x_j,
 * which is what the symmetric sweep adds up (the pair terms are odd under
 * exchange of target and source). */
static void onesided_ld_target(int kind, double nu, size_t n, const double *pos,
                               const double *vort, const double *rad, const double *vol,
                               size_t i, long double u[3], long double st[3],
                               long double df[3]) {
    for (int m = 0; m < 3; ++m) u[m] = st[m] = df[m] = 0.0L;
    const double *xi = pos + 3 * i, *ai = vort + 3 * i;
    for (size_t j = 0; j < n; ++j) {
        if (j == i) continue;
        const double *xj = pos + 3 * j, *aj = vort + 3 * j;
        long double d[3] = {(long double)xi[0] - xj[0], (long double)xi[1] - xj[1],
                            (long double)xi[2] - xj[2]};
        long double rho = sqrtl(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        long double sig2 = 0.5L * ((long double)rad[i] * rad[i] + (long double)rad[j] * rad[j]);
        long double q, F, Z;
        orc_kernel_ld(kind, rho, sqrtl(sig2), &q, &F, &Z);
        long double c[3] = {aj[1] * d[2] - aj[2] * d[1], aj[2] * d[0] - aj[0] * d[2],
                            aj[0] * d[1] - aj[1] * d[0]};
        long double e[3] = {(long double)ai[1] * aj[2] - (long double)ai[2] * aj[1],
                            (long double)ai[2] * aj[0] - (long double)ai[0] * aj[2],
                            (long double)ai[0] * aj[1] - (long double)ai[1] * aj[0]};
        long double de = d[0] * e[0] + d[1] * e[1] + d[2] * e[2];
        long double k = (long double)nu * Z / (0.5L * sig2);
        for (int m = 0; m < 3; ++m) {
            u[m] += q * c[m];
            st[m] += q * e[m] + F * de * d[m];
            df[m] += k * ((long double)vol[i] * aj[m] - (long double)vol[j] * ai[m]);
        }
    }
}

static void onesided_ld(int kind, double nu, size_t n, const double *pos, const double *vort,
                        const double *rad, const double *vol, double *vel, double *stretch,
                        double *diff) {
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t i = 0; i < n; ++i) {
        long double u[3], st[3], df[3];
        onesided_ld_target(kind, nu, n, pos, vort, rad, vol, i, u, st, df);
        for (int m = 0; m < 3; ++m) {
            vel[3 * i + m] = (double)u[m];
            if (diff) {
                stretch[3 * i + m] = (double)st[m];
                diff[3 * i + m] = (double)df[m];
            } else {
                stretch[3 * i + m] = (double)(st[m] + df[m]);
            }
        }
    }
}

static size_t total_n(size_t nwake, const size_t *wake_n) {
    size_t n = 0;
    for (size_t w = 0; w < nwake; ++w) n += wake_n[w];
    return n;
}

int orc_rates(int mode, int kind, double nu, size_t nwake, const size_t *wake_n,
              const double *pos, const double *vort, const double *radius,
              const double *volume, const double *uinf, double *vel, double *dvort) {
    if (!wake_n || (kind != ORC_KERNEL_HIGHORDER && kind != ORC_KERNEL_GAUSSIAN)) return -1;
    size_t n = total_n(nwake, wake_n);
    /* interact(): zero the rates (interaction.cpp:96-99) */
    memset(vel, 0, 3 * n * sizeof(double));
    memset(dvort, 0, 3 * n * sizeof(double));
    if (mode == ORC_MODE_ONESIDED_LD) {
        onesided_ld(kind, nu, n, pos, vort, radius, volume, vel, dvort, NULL);
    } else if (mode == ORC_MODE_SERIAL || mode == ORC_MODE_OMP_FAITHFUL ||
               mode == ORC_MODE_OMP_CLEAN) {
        size_t *off = (size_t *)malloc((nwake + 1) * sizeof(size_t));
        off[0] = 0;
        for (size_t w = 0; w < nwake; ++w) off[w + 1] = off[w] + wake_n[w];
        /* each wake with itself, then each wake pair i<j (interaction.cpp:100-107) */
        for (size_t w = 0; w < nwake; ++w) {
            if (mode == ORC_MODE_SERIAL)
                serial_block(kind, nu, off[w], off[w + 1], off[w], off[w + 1], 1, pos, vort,
                             radius, volume, vel, dvort);
            else
                for (size_t is = off[w]; is < off[w + 1]; ++is)
                    omp_row(mode == ORC_MODE_OMP_FAITHFUL, kind, nu, is, is + 1, off[w + 1], pos,
                            vort, radius, volume, vel, dvort);
        }
        for (size_t a = 0; a < nwake; ++a)
            for (size_t b = a + 1; b < nwake; ++b) {
                if (mode == ORC_MODE_SERIAL)
                    serial_block(kind, nu, off[a], off[a + 1], off[b], off[b + 1], 0, pos, vort,
                                 radius, volume, vel, dvort);
                else
                    for (size_t is = off[a]; is < off[a + 1]; ++is)
                        omp_row(mode == ORC_MODE_OMP_FAITHFUL, kind, nu, is, off[b], off[b + 1],
                                pos, vort, radius, volume, vel, dvort);
            }
        free(off);
    } else {
        return -1;
    }
    if (uinf)
        for (size_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) vel[3 * i + c] += uinf[c];
    return 0;
}

int orc_rates_split_ld(int kind, double nu, size_t n, const double *pos, const double *vort,
                       const double *radius, const double *volume, double *vel,
                       double *stretch, double *diffusion) {
    if (kind != ORC_KERNEL_HIGHORDER && kind != ORC_KERNEL_GAUSSIAN) return -1;
    onesided_ld(kind, nu, n, pos, vort, radius, volume, vel, stretch, diffusion);
    return 0;
}

/* Rates of a sample of targets against every particle.  interact() pairs
 * every particle with every other once (interaction.cpp:95-108: each wake with
 * itself, then each pair of wakes) and adds each pair's terms to both
 * particles (interaction_utils.h:429-481, parallel.cpp:13-85); the terms a
 * particle receives as the pair's source are the exact negations of those it
 * would receive as its target (every term is odd under exchange: the
 * displacement and a_t x a_s change sign, sigma and the kernel do not), so a
 * particle's rate is the sum over j != i of its target-side terms.  Here that
 * sum runs over j in index order, with pair_eval's binary64 arithmetic (ld ==
 * 0), or in long double (ld != 0, onesided_ld_target). */
int orc_rates_sample(int ld, int kind, double nu, size_t n, const double *pos,
                     const double *vort, const double *radius, const double *volume,
                     const double *uinf, size_t ns, const size_t *targets, double *vel,
                     double *stretch, double *diffusion) {
    if (kind != ORC_KERNEL_HIGHORDER && kind != ORC_KERNEL_GAUSSIAN) return -1;
    if (!targets || !vel || !stretch) return -1;
    for (size_t k = 0; k < ns; ++k)
        if (targets[k] >= n) return -1;
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t k = 0; k < ns; ++k) {
        const size_t i = targets[k];
        double u[3] = {0.0, 0.0, 0.0}, st[3] = {0.0, 0.0, 0.0}, df[3] = {0.0, 0.0, 0.0};
        if (ld) {
            long double lu[3], ls[3], lf[3];
            onesided_ld_target(kind, nu, n, pos, vort, radius, volume, i, lu, ls, lf);
            for (int m = 0; m < 3; ++m) {
                u[m] = (double)lu[m];
                if (diffusion) {
                    st[m] = (double)ls[m];
                    df[m] = (double)lf[m];
                } else {
                    st[m] = (double)(ls[m] + lf[m]);
                }
            }
        } else {
            for (size_t j = 0; j < n; ++j) {
                if (j == i) continue;
                pair_out o;
                pair_eval(kind, nu, radius[j], radius[i], pos + 3 * j, pos + 3 * i,
                          vort + 3 * j, vort + 3 * i, volume[j], volume[i], &o);
                for (int m = 0; m < 3; ++m) {
                    u[m] += o.dr_t[m];
                    if (diffusion) {
                        st[m] += o.st[m];
                        df[m] += o.df[m];
                    } else {
                        st[m] += o.da[m];
                    }
                }
            }
        }
        for (int m = 0; m < 3; ++m) {
            vel[3 * k + m] = uinf ? u[m] + uinf[m] : u[m];
            stretch[3 * k + m] = st[m];
            if (diffusion) diffusion[3 * k + m] = df[m];
        }
    }
    return 0;
}

double orc_omp_rows(int faithful, int kind, double nu, size_t n, const double *pos,
                    const double *vort, const double *radius, const double *volume,
                    size_t row_begin, size_t stride, size_t nrows, double *vel, double *dvort) {
    double pairs = 0.0;
    if (stride == 0) stride = 1;
    for (size_t k = 0; k < nrows; ++k) {
        size_t is = row_begin + k * stride;
        if (is >= n) break;
        omp_row(faithful, kind, nu, is, is + 1, n, pos, vort, radius, volume, vel, dvort);
        pairs += (double)(n - 1 - is);
    }
    return pairs;
}

/* ------------------------------------------------------------- state layout */

/* __interaction::getStates (interaction.cpp:222-229) over
 * __wake::getStates (wake.cpp:110-120): per wake [pos 3n | vort 3n]. */
void orc_pack_state(size_t nwake, const size_t *wake_n, const double *pos,
                    const double *vort, double *state) {
    size_t p = 0, off = 0;
    for (size_t w = 0; w < nwake; ++w) {
        size_t m = 3 * wake_n[w];
        memcpy(state + off, pos + p, m * sizeof(double));
        memcpy(state + off + m, vort + p, m * sizeof(double));
        off += 2 * m;
        p += m;
    }
}

/* __interaction::setStates (interaction.cpp:204-211) over __wake::setStates
 * (wake.cpp:87-96). */
void orc_unpack_state(size_t nwake, const size_t *wake_n, const double *state, double *pos,
                      double *vort) {
    size_t p = 0, off = 0;
    for (size_t w = 0; w < nwake; ++w) {
        size_t m = 3 * wake_n[w];
        memcpy(pos + p, state + off, m * sizeof(double));
        memcpy(vort + p, state + off + m, m * sizeof(double));
        off += 2 * m;
        p += m;
    }
}

/* setStates(x); solve(); getRates(k) */
static int rhs(int mode, int kind, double nu, size_t nwake, const size_t *wake_n,
               const double *radius, const double *volume, const double *uinf,
               const double *x, double *k, double *pos, double *vort, double *vel,
               double *dvort) {
    orc_unpack_state(nwake, wake_n, x, pos, vort);
    int rc = orc_rates(mode, kind, nu, nwake, wake_n, pos, vort, radius, volume, uinf, vel, dvort);
    orc_pack_state(nwake, wake_n, vel, dvort, k);
    return rc;
}

/* __rk4::step, rk4.cpp:12-77:
 *   k1 = f(x);  x1 = k1*(0.5 dt) + x
 *   k2 = f(x1); x2 = k2*(0.5 dt) + x
 *   k3 = f(x2); x3 = k3*dt + x
 *   k4 = f(x3); x += (k1 + k4)*(dt/6) + (k2 + k3)*(dt/3)               */
int orc_rk4(int mode, int kind, double nu, size_t nwake, const size_t *wake_n,
            const double *radius, const double *volume, const double *uinf, double *state,
            double dt, size_t nsteps) {
    size_t n = total_n(nwake, wake_n), m = 6 * n;
    double *buf = (double *)calloc(7 * m + 12 * n, sizeof(double));
    if (!buf) return -1;
    double *x1 = buf, *x2 = x1 + m, *x3 = x2 + m, *k1 = x3 + m, *k2 = k1 + m, *k3 = k2 + m,
           *k4 = k3 + m;
    double *pos = k4 + m, *vort = pos + 3 * n, *vel = vort + 3 * n, *dvort = vel + 3 * n;
    int rc = 0;
    for (size_t s = 0; s < nsteps && rc == 0; ++s) {
        rc |= rhs(mode, kind, nu, nwake, wake_n, radius, volume, uinf, state, k1, pos, vort, vel, dvort);
        for (size_t i = 0; i < m; ++i) {
            x1[i] = k1[i] * (0.5 * dt);
            x1[i] += state[i];
        }
        rc |= rhs(mode, kind, nu, nwake, wake_n, radius, volume, uinf, x1, k2, pos, vort, vel, dvort);
        for (size_t i = 0; i < m; ++i) {
            x2[i] = k2[i] * (0.5 * dt);
            x2[i] += state[i];
        }
        rc |= rhs(mode, kind, nu, nwake, wake_n, radius, volume, uinf, x2, k3, pos, vort, vel, dvort);
        for (size_t i = 0; i < m; ++i) {
            x3[i] = k3[i] * dt;
            x3[i] += state[i];
        }
        rc |= rhs(mode, kind, nu, nwake, wake_n, radius, volume, uinf, x3, k4, pos, vort, vel, dvort);
        for (size_t i = 0; i < m; ++i) {
            k1[i] += k4[i];
            k1[i] *= dt / 6.;
            k2[i] += k3[i];
            k2[i] *= dt / 3.;
            k1[i] += k2[i];
            state[i] += k1[i];
        }
    }
    free(buf);
    return rc;
}

/* __integration::step, integration.cpp:37-45: x += f(x)*dt */
int orc_euler(int mode, int kind, double nu, size_t nwake, const size_t *wake_n,
              const double *radius, const double *volume, const double *uinf, double *state,
              double dt, size_t nsteps) {
    size_t n = total_n(nwake, wake_n), m = 6 * n;
    double *buf = (double *)calloc(m + 12 * n, sizeof(double));
    if (!buf) return -1;
    double *k = buf, *pos = k + m, *vort = pos + 3 * n, *vel = vort + 3 * n, *dvort = vel + 3 * n;
    int rc = 0;
    for (size_t s = 0; s < nsteps && rc == 0; ++s) {
        rc |= rhs(mode, kind, nu, nwake, wake_n, radius, volume, uinf, state, k, pos, vort, vel, dvort);
        for (size_t i = 0; i < m; ++i) {
            k[i] *= dt;
            state[i] += k[i];
        }
    }
    free(buf);
    return rc;
}

/* ------------------------------------------------------------- diagnostics */

/* calculateTotalVorticity, interaction.cpp:240-246, summed over wakes as in
 * diagnose() (:58-68) */
void orc_total_vorticity(size_t nwake, const size_t *wake_n, const double *vort, double out[3]) {
    size_t p = 0;
    out[0] = out[1] = out[2] = 0.0;
    for (size_t w = 0; w < nwake; ++w) {
        double O[3] = {0.0, 0.0, 0.0};
        for (size_t i = 0; i < wake_n[w]; ++i, ++p)
            for (int c = 0; c < 3; ++c) O[c] += vort[3 * p + c];
        for (int c = 0; c < 3; ++c) out[c] += O[c];
    }
}

/* calculateLinearImpulse, interaction.cpp:248-259: I = 0.5 * sum x x a */
void orc_linear_impulse(size_t nwake, const size_t *wake_n, const double *pos,
                        const double *vort, double out[3]) {
    size_t p = 0;
    out[0] = out[1] = out[2] = 0.0;
    for (size_t w = 0; w < nwake; ++w) {
        double I[3] = {0.0, 0.0, 0.0}, rxa[3];
        for (size_t i = 0; i < wake_n[w]; ++i, ++p) {
            cross3(pos + 3 * p, vort + 3 * p, rxa);
            for (int c = 0; c < 3; ++c) I[c] += rxa[c];
        }
        for (int c = 0; c < 3; ++c) out[c] += I[c] * 0.5;
    }
}

/* calculateAngularImpulse, interaction.cpp:261-275: A = (1/3) sum x x (x x a)
 * (serial version; the OpenMP one reduces into uninitialised seeds) */
void orc_angular_impulse(size_t nwake, const size_t *wake_n, const double *pos,
                         const double *vort, double out[3]) {
    size_t p = 0;
    out[0] = out[1] = out[2] = 0.0;
    for (size_t w = 0; w < nwake; ++w) {
        double A[3] = {0.0, 0.0, 0.0}, rxa[3], rxrxa[3];
        for (size_t i = 0; i < wake_n[w]; ++i, ++p) {
            cross3(pos + 3 * p, vort + 3 * p, rxa);
            cross3(pos + 3 * p, rxa, rxrxa);
            for (int c = 0; c < 3; ++c) A[c] += rxrxa[c];
        }
        for (int c = 0; c < 3; ++c) out[c] += A[c] * (1.0 / 3.0);
    }
}

/* KINETICENERGY, interaction_utils.h:197-204 (self) and :215-240 (pair):
 *   self = a.a/(8 pi s)
 *   pair = ((rho^2 + 2 sig^2)(a1.a2) + (x12.a1)(x12.a2)) / (rho^2+sig^2)^1.5/(16 pi)
 * with x12 = x2 - x1, sig^2 = (s1^2 + s2^2)/2. */
static double ke_self(double s, const double *a) { return M_1_PI * dot3(a, a) / s / 8.0; }
static double ke_pair(double s1, double s2, const double *x1, const double *x2,
                      const double *a1, const double *a2) {
    double x12[3];
    for (int i = 0; i < 3; ++i) x12[i] = x2[i] - x1[i];
    double rho2 = dot3(x12, x12);
    double sigma2 = 0.5 * (s1 * s1 + s2 * s2);
    double a1a2 = dot3(a1, a2);
    double x12a1 = dot3(x12, a1);
    double x12a2 = dot3(x12, a2);
    return ((rho2 + 2.0 * sigma2) * a1a2 + (x12a1 * x12a2)) / pow(rho2 + sigma2, 1.5) / 16.0 / M_PI;
}

/* ENST, highorder.h:69-86 and ENSTROPHY, interaction_utils.h:107-156 */
static double gpow7(double x) {
    double x2 = x * x, x4 = x2 * x2;
    return x4 * x2 * x;
}
static double enst_self(double s, const double *a) {
    return 5.0 * M_1_PI / 4.0 / (s * s * s) * dot3(a, a);
}
static double enst_pair(double s1, double s2, const double *x1, const double *x2,
                        const double *a1, const double *a2) {
    double x12[3];
    for (int i = 0; i < 3; ++i) x12[i] = x2[i] - x1[i];
    double rho = orc_dnrm2(3, x12);
    double sigma = sqrt(0.5 * (s1 * s1 + s2 * s2));
    double rho2 = rho * rho, sig2 = sigma * sigma;
    double factor = (M_1_PI / 8.0) / sqrt(gpow7(rho2 + sig2));
    double F1 = factor * (10.0 * sig2 * sig2 - 2.0 * rho2 * rho2 - 7.0 * sig2 * rho2);
    double F2 = factor * (-15.0 * rho2);
    return F1 * dot3(a1, a2) + F2 * dot3(a1, x12) * dot3(a2, x12);
}

/* HELICITY, interaction_utils.h:167-190 with QSIG (highorder.h:30-38) */
static double helicity_pair(double s1, double s2, const double *x1, const double *x2,
                            const double *a1, const double *a2) {
    double x12[3], axa[3];
    for (int i = 0; i < 3; ++i) x12[i] = x2[i] - x1[i];
    double rho = orc_dnrm2(3, x12);
    double sigma = sqrt(0.5 * (s1 * s1 + s2 * s2));
    double q, F, Z;
    orc_kernel(ORC_KERNEL_HIGHORDER, rho, sigma, &q, &F, &Z);
    cross3(a1, a2, axa);
    return q * dot3(x12, axa);
}

typedef double (*pairfn)(double, double, const double *, const double *, const double *,
                         const double *);

/* Generic quadratic diagnostic, following calculateKineticEnergy /
 * calculateEnstrophy / calculateHelicity (interaction.cpp:277-383):
 *   per wake:  s = sum_I [ self(I) + sum_{J>I} within * pair(I,J) ]; s *= wout
 *   per pair of wakes: s = cross_scale * sum_I sum_J pair(I,J)               */
static double quad_diag(size_t nwake, const size_t *wake_n, const double *pos,
                        const double *vort, const double *rad,
                        double (*self)(double, const double *), pairfn pair, double within,
                        double wout, double cross_scale) {
    size_t *off = (size_t *)malloc((nwake + 1) * sizeof(size_t));
    off[0] = 0;
    for (size_t w = 0; w < nwake; ++w) off[w + 1] = off[w] + wake_n[w];
    double total = 0.0;
    for (size_t w = 0; w < nwake; ++w) {
        double s = 0.0;
        for (size_t I = off[w]; I < off[w + 1]; ++I) {
            if (self) s += self(rad[I], vort + 3 * I);
            for (size_t J = I + 1; J < off[w + 1]; ++J)
                s += within * pair(rad[I], rad[J], pos + 3 * I, pos + 3 * J, vort + 3 * I,
                                   vort + 3 * J);
        }
        total += wout * s;
        for (size_t v = w + 1; v < nwake; ++v) {
            double c = 0.0;
            for (size_t I = off[w]; I < off[w + 1]; ++I)
                for (size_t J = off[v]; J < off[v + 1]; ++J)
                    c += pair(rad[I], rad[J], pos + 3 * I, pos + 3 * J, vort + 3 * I,
                              vort + 3 * J);
            total += cross_scale * c;
        }
    }
    free(off);
    return total;
}

/* calculateKineticEnergy(W) = 2[sum self + 2 sum_{I<J} pair] (:351-366),
 * applied to all particles of all wakes.  The reference's cross-wake routine
 * calculateKineticEnergy(W1,W2) = 2 sum pair (:368-383) weights a cross pair
 * half as much as the single-wake routine weights a pair, so a sum of the two
 * would depend on how particles are split into wakes (diagnose() calls
 * neither); cross pairs get the single-wake weight 4 here. */
double orc_kinetic_energy(size_t nwake, const size_t *wake_n, const double *pos,
                          const double *vort, const double *radius) {
    return quad_diag(nwake, wake_n, pos, vort, radius, ke_self, ke_pair, 2.0, 2.0, 4.0);
}

/* The reference's cross-wake routine as written (interaction.cpp:368-383). */
double orc_kinetic_energy_cross(size_t n1, size_t n2, const double *pos, const double *vort,
                                const double *radius) {
    double ke = 0.0;
    for (size_t I = 0; I < n1; ++I)
        for (size_t J = n1; J < n1 + n2; ++J)
            ke += ke_pair(radius[I], radius[J], pos + 3 * I, pos + 3 * J, vort + 3 * I,
                          vort + 3 * J);
    return 2.0 * ke;
}

/* calculateEnstrophy(W) = sum self + 2 sum_{I<J} pair (:277-298);
 * calculateEnstrophy(W1,W2) = 2 sum pair (:300-315) */
double orc_enstrophy(size_t nwake, const size_t *wake_n, const double *pos, const double *vort,
                     const double *radius) {
    return quad_diag(nwake, wake_n, pos, vort, radius, enst_self, enst_pair, 2.0, 1.0, 2.0);
}

/* calculateHelicity(W) = 2 sum_{I<J} pair (:317-332);
 * calculateHelicity(W1,W2) = 2 sum pair (:334-349) */
double orc_helicity(size_t nwake, const size_t *wake_n, const double *pos, const double *vort,
                    const double *radius) {
    return quad_diag(nwake, wake_n, pos, vort, radius, NULL, helicity_pair, 1.0, 2.0, 2.0);
}

/* influence(), interaction.cpp:110-155: k_i = zeta(0, s_i) a_i (SELFINFLUENCE,
 * interaction_utils.h:592-605) + sum_{j!=i} zeta(rho_ij, sigma_ij) a_j
 * (INFLUENCE, :530-556), symmetric sweep. */
int orc_influence(int kind, size_t nwake, const size_t *wake_n, const double *pos,
                  const double *vort, const double *radius, double *field) {
    if (kind != ORC_KERNEL_HIGHORDER && kind != ORC_KERNEL_GAUSSIAN) return -1;
    size_t n = total_n(nwake, wake_n);
    memset(field, 0, 3 * n * sizeof(double));
    size_t *off = (size_t *)malloc((nwake + 1) * sizeof(size_t));
    off[0] = 0;
    for (size_t w = 0; w < nwake; ++w) off[w + 1] = off[w] + wake_n[w];
    for (size_t pass = 0; pass < 2; ++pass)
        for (size_t a = 0; a < nwake; ++a)
            for (size_t b = (pass == 0 ? a : a + 1); b < (pass == 0 ? a + 1 : nwake); ++b)
                for (size_t is = off[a]; is < off[a + 1]; ++is) {
                    if (pass == 0) {
                        double Z = zeta_sig(kind, 0.0, radius[is]);
                        for (int c = 0; c < 3; ++c) field[3 * is + c] += Z * vort[3 * is + c];
                    }
                    size_t tb = pass == 0 ? is + 1 : off[b];
                    for (size_t it = tb; it < off[b + 1]; ++it) {
                        double sig = sqrt(0.5 * (radius[is] * radius[is] + radius[it] * radius[it]));
                        double disp[3];
                        for (int c = 0; c < 3; ++c) disp[c] = pos[3 * it + c] - pos[3 * is + c];
                        double Z = zeta_sig(kind, orc_dnrm2(3, disp), sig);
                        for (int c = 0; c < 3; ++c) field[3 * it + c] += Z * vort[3 * is + c];
                        for (int c = 0; c < 3; ++c) field[3 * is + c] += Z * vort[3 * it + c];
                    }
                }
    free(off);
    return 0;
}

int orc_max_threads(void) { return omp_get_max_threads(); }
double orc_wtime(void) { return omp_get_wtime(); }
