// vpm_types.hpp -- data layout shared by the gfx950 kernels and the host engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pw {

constexpr int KIND_HIGHORDER = 0;
constexpr int KIND_GAUSSIAN = 1;
constexpr int PREC_FP64 = 0;
constexpr int PREC_FP32C = 1;

// Partial-sum components per (target, chunk) of the one-sided passes:
//   0-2 sum q (a_j x d)   3-5 sum [q e + F (d.e) d], e = a_i x a_j (the pair's
//   stretching vector, interaction_utils.h:69-100)   6-8 sum k a_j   9 sum k v_j
constexpr int NCOMP = 10;
// Symmetric binary32 pass (each unordered pair once): per (particle, slot)
//   0-2 velocity   3-5 stretching q e + F (d.e) d   6-8 sum k a   9 sum k v
// for the target side (slots per source chunk) and for the source side
// (slots per target block; the stretching there enters with a minus sign).
constexpr int NCOMP_S = 10;

// Combine-kernel modes (MODE_PER_JOB: each job's JobDev::mode)
constexpr int MODE_PER_JOB = -1;
constexpr int MODE_SOLVE = 0;
constexpr int MODE_RK4_1 = 1;
constexpr int MODE_RK4_2 = 2;
constexpr int MODE_RK4_3 = 3;
constexpr int MODE_RK4_4 = 4;
constexpr int MODE_EULER = 5;

constexpr double K_PI = 3.14159265358979323846;
constexpr double K_CZ = 1.875 / K_PI;   // HIGHORDER zeta(0)
constexpr double K_CPHI = 0.25 / K_PI;  // phi prefactor
constexpr double K_CG = 0.063493635934240969;  // (2 pi)^-1.5
// phi prefactor of the scaled HIGHORDER kernels: kz K_CPHI / 32 for binary64
// (kz = 2^15 / K_CZ), kz K_CPHI for binary32 (kz = 1 / K_CZ)
constexpr double K_HC64 = 2048.0 / 15.0;
constexpr double K_HC32 = 2.0 / 15.0;

// Factors the pair kernels leave on their sums: kq on Q and F' (velocity,
// stretching, d(d.e) terms), kz on Zs (PSE sums).  The epilogue divides them
// out.  Host and device code.
struct PairScale {
    double q, z;
};
__host__ __device__ inline PairScale pair_scale(int kind, int precision) {
    if (kind != KIND_HIGHORDER) return PairScale{1.0, 1.0};
    if (precision == PREC_FP64) {
        const double kz = 32768.0 / K_CZ;
        return PairScale{4.0 * kz, kz};
    }
    return PairScale{1.0 / K_CZ, 1.0 / K_CZ};
}
constexpr double K_SQRT1_2 = 0.70710678118654752440;

// Source/target record of the binary64 path: 64 bytes, read as 4 x 16 B.
// h = s^2 / 2, so that sigma^2 = h_i + h_j.
struct __align__(16) Rec64 {
    double x, y, z, h;
    double ax, ay, az, v;
};
// Record of the compensated binary32 path: binary64 positions (the
// displacement is formed in binary64 and then rounded) and binary64
// strengths (the stretching cross product a_i x a_j is formed in binary64:
// nearly parallel strengths would lose it in binary32), a binary32 copy of
// the strengths and the volume for the velocity and diffusion terms, and the
// volume in binary64 (close pairs form the PSE differences v_j - v_i in
// binary64).  80 bytes, 5 x 16 B.
struct __align__(16) Rec32 {
    double x, y, z, h;
    double ax, ay, az;
    float fx, fy, fz, v;
    double vd;
};

static_assert(sizeof(Rec64) == 64, "Rec64 is four 16-byte vectors");
static_assert(sizeof(Rec32) == 80, "Rec32 is five 16-byte vectors");

// One simulation as the kernels see it.  Arrays are structure-of-arrays with
// a padded stride npad (multiple of 256): component m of particle i is at
// [m * npad + i].
struct JobDev {
    void *rec;      // Rec64[npad] or Rec32[npad]: the stage being evaluated
    double *x0;     // [6][npad] state at the start of the step (x y z ax ay az)
    double *xs;     // [6][npad] state of the stage being evaluated
    double *k1;     // [6][npad] RK4 k1
    double *k23;    // [6][npad] RK4 k2 (+ k3)
    double *rate;   // [6][npad] last rates (u, da/dt)
    double *part;   // per-chunk target-side partial sums: one-sided [nchunk][NCOMP][npad];
                    // symmetric: chunk c's row [NCOMP_S][sym_part_len(c)] at sym_part_offset(c)
    const double *vol;  // [npad] volumes
    int *flag;      // set to 1 when a rate or state is non-finite
    int n, npad, nchunk, chunk_len;
    int prec, uniform;
    int mode;       // stage this job runs when the combine is launched with MODE_PER_JOB
    int sym;        // binary32 symmetric pass: part holds NCOMP_S target-side
                    // components, jpart the source-side ones
    int tw;         // targets per block of the symmetric pass
    int sw;         // targets per strip (a whole number of blocks): the work of a
                    // strip's blocks against one source chunk is one work item
    double *jpart;  // source-side slots: strip s holds [NCOMP_S][npad - s sw]
                    // sums for particles s sw .. npad-1, at jpart_offset(s)
    double inv_s2;  // non-uniform core: unused; uniform: 1 (positions are prescaled)
    double rec_scale;  // records hold positions x rec_scale (1/sigma for a uniform core)
    double su, sq, sd;  // epilogue scales (see combine_kernel)
    double dt, nu;
    double uinf[3];
};

// Start of strip s's source-side slot row (rows shrink by sw per strip: the
// strip's blocks only reach particles from s sw on).
__host__ __device__ inline size_t jpart_offset(size_t s, size_t npad, size_t sw) {
    return (size_t)NCOMP_S * (s * npad - sw * (s * (s - 1) / 2));
}

// Target-side rows of a symmetric job.  Row c (source chunk c, L sources)
// holds the sums of the targets of the blocks that start before the chunk
// ends, [0, sym_part_len(c)) -- when tw divides L; otherwise every row holds
// all npad targets.
__host__ __device__ inline size_t sym_part_len(size_t c, size_t npad, size_t L, size_t tw) {
    if (L % tw != 0) return npad;
    const size_t e = (c + 1) * L;
    return e < npad ? e : npad;
}
__host__ __device__ inline size_t sym_part_offset(size_t c, size_t npad, size_t L, size_t tw) {
    if (L % tw != 0) return (size_t)NCOMP_S * c * npad;
    const size_t cm = npad / L;  // rows c' < cm hold (c' + 1) L targets, later ones npad
    if (c <= cm) return (size_t)NCOMP_S * L * (c * (c + 1) / 2);
    return (size_t)NCOMP_S * (L * (cm * (cm + 1) / 2) + (c - cm) * npad);
}

struct Item {
    int job, tile, chunk, pad;
};
// Combine item: particles [256 tile, 256 tile + 256) of a one-sided job,
// [64 tile, 64 tile + 64) of a symmetric one.
struct CItem {
    int job, tile;
};

}  // namespace pw
