/*
 * vpm_oracle.h -- CPU oracle for the vortex-particle interaction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against, and the timed CPU baseline.  Only tests/, the smoke
 * check and the cpu_baseline leg of bench.py may load it.  The product library
 * (libpawan_hip.so) never links, loads or calls it.
 *
 * It restates, in plain C, the arithmetic of pawan's reference method:
 *   - smoothing kernel      src/interaction/highorder.h:20-60, gaussian.h:20-58
 *   - pair terms            src/interaction/interaction_utils.h:52-100,253-286
 *   - pair drivers          interaction_utils.h:363-407 (two-sided, serial)
 *                           interaction_utils.h:429-481 (source parts, OpenMP)
 *   - all-pairs loops       src/interaction/interaction.cpp:95-108,157-195
 *                           src/interaction/parallel.cpp:13-85
 *   - state layout          src/wake/wake.cpp:87-120, interaction.cpp:204-229
 *   - RK4 / Euler           src/integration/rk4.cpp:12-77, integration.cpp:37-45
 *   - diagnostics           interaction.cpp:240-383, interaction_utils.h:107-240,
 *                           highorder.h:69-86
 *   - vorticity field       interaction.cpp:110-155, interaction_utils.h:493-605
 *
 * Particle data is passed the way pawan stores it: positions and vorticities
 * as N x 3 row-major doubles (gsl_matrix with tda = 3), radius and volume as N
 * doubles.  Several wakes are concatenated in order; wake_n[w] gives the
 * particle count of wake w (the reference's std::vector<__wake*> _W).
 *
 * Rounding: built with -O2 -ffp-contract=off, so every product and sum is an
 * unfused IEEE binary64 operation in the order the reference writes it, as the
 * reference's interaction library is built (-O3 without fast-math, no FMA on
 * the default x86-64 target).
 */
#ifndef VPM_ORACLE_H_
#define VPM_ORACLE_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* smoothing kernel selector */
#define ORC_KERNEL_HIGHORDER 0 /* highorder.h (compiled default)           */
#define ORC_KERNEL_GAUSSIAN 1  /* gaussian.h  (compiled out in reference)  */

/* all-pairs evaluation mode */
#define ORC_MODE_SERIAL 0        /* interaction.cpp:95-108,157-195 symmetric i<j      */
#define ORC_MODE_OMP_FAITHFUL 1  /* parallel.cpp:13-85 with per-pair heap temporaries */
#define ORC_MODE_OMP_CLEAN 2     /* parallel.cpp:13-85 with stack temporaries         */
#define ORC_MODE_ONESIDED_LD 3   /* one-sided per-target sum in long double            */

/* q, F, Z of the smoothing kernel at distance rho and core sigma */
void orc_kernel(int kind, double rho, double sigma, double *q, double *F, double *Z);
/* the same in long double (independent evaluation used to pin the double one) */
void orc_kernel_ld(int kind, long double rho, long double sigma, long double *q,
                   long double *F, long double *Z);
/* Euclidean norm by the reference-BLAS scaled sum of squares (GSL cblas_dnrm2) */
double orc_dnrm2(size_t n, const double *x);

/* One pair (target, source) as the reference's OpenMP INTERACT computes it
 * (interaction_utils.h:429-481): adds the target's velocity and vorticity rate
 * contributions into dr_trg/da_trg and returns the source's parts. */
void orc_pair(int kind, double nu, double s_src, double s_trg, const double r_src[3],
              const double r_trg[3], const double a_src[3], const double a_trg[3],
              double v_src, double v_trg, double dr_trg[3], double da_trg[3],
              double dr_src_out[3], double da_src_out[3]);

/* Velocity and vorticity rate of every particle (solve() + getRates split into
 * the velocity and retvorcity matrices).  uinf, when not NULL, is a uniform
 * freestream added to every velocity (not part of the reference; the rotor-wake
 * workload needs it).  Returns 0, or -1 on bad arguments. */
int orc_rates(int mode, int kind, double nu, size_t nwake, const size_t *wake_n,
              const double *pos, const double *vort, const double *radius,
              const double *volume, const double *uinf, double *vel, double *dvort);

/* Per-term split of the one-sided long-double rates: velocity, stretching
 * (q and F terms) and PSE diffusion, each N x 3. */
int orc_rates_split_ld(int kind, double nu, size_t n, const double *pos, const double *vort,
                       const double *radius, const double *volume, double *vel,
                       double *stretch, double *diffusion);

/* Rates of the targets targets[0..ns) against all n particles (every wake
 * concatenated: interact() pairs all of them), N x 3 rows k = 0..ns-1 in
 * vel / stretch / diffusion; diffusion == NULL puts stretching + diffusion in
 * stretch.  ld == 0: the reference's binary64 pair arithmetic, summed over
 * sources in index order; ld != 0: long double.  uinf (may be NULL) is added
 * to vel.  For checking a GPU pass at sizes where all N rates would take too
 * long on the CPU.  Returns 0, or -1 on bad arguments. */
int orc_rates_sample(int ld, int kind, double nu, size_t n, const double *pos,
                     const double *vort, const double *radius, const double *volume,
                     const double *uinf, size_t ns, const size_t *targets, double *vel,
                     double *stretch, double *diffusion);

/* State vector <-> particle arrays in the reference layout: for each wake in
 * turn, [positions 3n | vorticities 3n] (wake.cpp:87-120). */
void orc_pack_state(size_t nwake, const size_t *wake_n, const double *pos,
                    const double *vort, double *state);
void orc_unpack_state(size_t nwake, const size_t *wake_n, const double *state, double *pos,
                      double *vort);

/* nsteps of classic RK4 (rk4.cpp:12-77) or forward Euler (integration.cpp:37-45)
 * on the state vector, in place. */
int orc_rk4(int mode, int kind, double nu, size_t nwake, const size_t *wake_n,
            const double *radius, const double *volume, const double *uinf, double *state,
            double dt, size_t nsteps);
int orc_euler(int mode, int kind, double nu, size_t nwake, const size_t *wake_n,
              const double *radius, const double *volume, const double *uinf, double *state,
              double dt, size_t nsteps);

/* Diagnostics summed over wakes as __interaction::diagnose() does. */
void orc_total_vorticity(size_t nwake, const size_t *wake_n, const double *vort, double out[3]);
void orc_linear_impulse(size_t nwake, const size_t *wake_n, const double *pos,
                        const double *vort, double out[3]);
void orc_angular_impulse(size_t nwake, const size_t *wake_n, const double *pos,
                         const double *vort, double out[3]);
/* kinetic energy of all particles: calculateKineticEnergy(W)
 * (interaction.cpp:351-366) over the union of the wakes */
double orc_kinetic_energy(size_t nwake, const size_t *wake_n, const double *pos,
                          const double *vort, const double *radius);
/* calculateKineticEnergy(W1, W2) as written (interaction.cpp:368-383), with
 * W1 the first n1 particles and W2 the next n2 */
double orc_kinetic_energy_cross(size_t n1, size_t n2, const double *pos, const double *vort,
                                const double *radius);
double orc_enstrophy(size_t nwake, const size_t *wake_n, const double *pos, const double *vort,
                     const double *radius);
double orc_helicity(size_t nwake, const size_t *wake_n, const double *pos, const double *vort,
                    const double *radius);

/* Vorticity field at every particle (influence(): interaction.cpp:110-155),
 * N x 3 output. */
int orc_influence(int kind, size_t nwake, const size_t *wake_n, const double *pos,
                  const double *vort, const double *radius, double *field);

/* CPU baseline sample: the OpenMP all-pairs loop of parallel.cpp:13-48 over a
 * single wake, restricted to the source rows row_begin, row_begin+stride, ...
 * (nrows rows).  Rates are accumulated into vel/dvort (N x 3, caller zeroes).
 * Returns the number of unordered pairs evaluated. */
double orc_omp_rows(int faithful, int kind, double nu, size_t n, const double *pos,
                    const double *vort, const double *radius, const double *volume,
                    size_t row_begin, size_t stride, size_t nrows, double *vel, double *dvort);

/* threads OpenMP will use, and wall clock (omp_get_wtime, as timing_utils.h:20) */
int orc_max_threads(void);
double orc_wtime(void);

#ifdef __cplusplus
}
#endif
#endif
