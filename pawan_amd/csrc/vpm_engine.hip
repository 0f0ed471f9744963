// vpm_engine.hip -- host engine of the gfx950 interaction path.
#include "vpm_engine.hpp"
#include "vpm_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/pawan_hip.h"

namespace pw {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
            throw Error(PW_ENOMEM, std::string(what) + ": " + hipGetErrorString(e));
        throw Error(PW_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    }
}

DevBuf::DevBuf(size_t b) { reset(b); }
DevBuf::~DevBuf() {
    if (p) (void)hipFree(p);
}
void DevBuf::reset(size_t b) {
    if (p) {
        (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    if (b == 0) return;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        throw Error(PW_ENOMEM, "device allocation of " + std::to_string(b) + " bytes failed: " +
                                   hipGetErrorString(e));
    }
    bytes = b;
}

HostBuf::~HostBuf() {
    if (p) (void)hipHostFree(p);
}
void HostBuf::ensure(size_t b) {
    if (b <= bytes) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipHostMalloc(&p, b, hipHostMallocDefault);
    if (e != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        throw Error(PW_ENOMEM, "pinned host allocation failed");
    }
    bytes = b;
}

static size_t round_up(size_t a, size_t b) { return (a + b - 1) / b * b; }
// jobs one engine holds at once (one non-finite flag each)
constexpr int kMaxJobSlots = 65536;
static int grid_for(size_t n) { return (int)std::min<size_t>(4096, (n + 255) / 256 + 1); }

size_t chunk_for(size_t n, int TB) {
    if (n == 0) return (size_t)TB;
    size_t ntile = (n + TB - 1) / TB;
    // 16 chunks, or enough for ~4,096 items (tiles x chunks) when there are
    // few tiles, so that a small job still spreads over the chip
    size_t nchunk = std::min(ntile, std::max<size_t>(16, (4096 + ntile - 1) / ntile));
    return (size_t)TB * ((ntile + nchunk - 1) / nchunk);
}

size_t chunk_for_sym(size_t n, int TB, int TPT) {
    if (n == 0) return (size_t)TB;
    const size_t tw = (size_t)TB * TPT, nblk = (n + tw - 1) / tw;
    const size_t nstrip = std::min(nblk, std::max<size_t>(128, nblk / 8));
    return tw * ((nblk + nstrip - 1) / nstrip);
}

size_t chunk_len(size_t n, int TB, int TPT, bool sym, int chunks) {
    if (chunks <= 0) return sym ? chunk_for_sym(n, TB, TPT) : chunk_for(n, TB);
    const size_t unit = sym ? (size_t)TB * TPT : (size_t)TB;
    const size_t nunit = std::max<size_t>(1, (n + unit - 1) / unit);
    return unit * ((nunit + (size_t)chunks - 1) / (size_t)chunks);
}

size_t strip_for_sym(size_t L, int TB, int TPT) {
    const size_t tw = (size_t)TB * TPT;
    return tw * std::max<size_t>(1, L / tw);
}

static size_t rec_bytes(int precision) {
    return precision == PREC_FP64 ? sizeof(Rec64) : sizeof(Rec32);
}

static void fill_records(void *dst, int precision, size_t m, const double *pos, const double *vort,
                         const double *radius, const double *volume, double scale) {
    for (size_t i = 0; i < m; ++i) {
        if (precision == PREC_FP64) {
            Rec64 &r = ((Rec64 *)dst)[i];
            r.x = pos[3 * i] * scale;
            r.y = pos[3 * i + 1] * scale;
            r.z = pos[3 * i + 2] * scale;
            r.h = 0.5 * radius[i] * radius[i];
            r.ax = vort[3 * i];
            r.ay = vort[3 * i + 1];
            r.az = vort[3 * i + 2];
            r.v = volume[i];
        } else {
            Rec32 &r = ((Rec32 *)dst)[i];
            r.x = pos[3 * i] * scale;
            r.y = pos[3 * i + 1] * scale;
            r.z = pos[3 * i + 2] * scale;
            r.h = 0.5 * radius[i] * radius[i];
            r.ax = vort[3 * i];
            r.ay = vort[3 * i + 1];
            r.az = vort[3 * i + 2];
            r.fx = (float)vort[3 * i];
            r.fy = (float)vort[3 * i + 1];
            r.fz = (float)vort[3 * i + 2];
            r.v = (float)volume[i];
            r.vd = volume[i];
        }
    }
}

size_t jpart_count(size_t n, size_t npad, size_t sw) {
    const size_t nstrip = (n + sw - 1) / sw;
    return jpart_offset(nstrip, npad, sw);
}

size_t part_count(size_t n, size_t npad, size_t L, size_t tw, bool sym) {
    const size_t nchunk = (n + L - 1) / L;
    return sym ? sym_part_offset(nchunk, npad, L, tw) : (size_t)NCOMP * nchunk * npad;
}

int block_for(int block, int precision) {
    return precision == PREC_FP64 && block > 256 ? 256 : block;
}

size_t job_footprint(size_t n, int precision, int TB, int TPT, bool sym, int chunks) {
    TB = block_for(TB, precision);
    size_t npad = round_up(std::max<size_t>(n, 1), 256);
    size_t L = chunk_len(std::max<size_t>(n, 1), TB, TPT, sym, chunks);
    size_t b = npad * (rec_bytes(precision) + 5 * 6 * sizeof(double) + 2 * sizeof(double)) +
               part_count(n, npad, L, (size_t)TB * TPT, sym) * sizeof(double) + 256;
    if (sym) b += jpart_count(npad, npad, strip_for_sym(L, TB, TPT)) * sizeof(double);
    return b;
}

JobDev Job::dev() const {
    JobDev d;
    std::memset(&d, 0, sizeof d);
    double *s = soa.as<double>();
    const size_t blk = 6 * npad;
    d.rec = rec.p;
    d.x0 = s;
    d.xs = s + blk;
    d.k1 = s + 2 * blk;
    d.k23 = s + 3 * blk;
    d.rate = s + 4 * blk;
    d.part = part.as<double>();
    d.vol = vol.as<double>();
    d.flag = flag;
    d.mode = MODE_SOLVE;
    d.n = (int)n;
    d.npad = (int)npad;
    d.nchunk = nchunk;
    d.chunk_len = chunk_len;
    d.prec = par.precision;
    d.uniform = uniform ? 1 : 0;
    d.sym = sym ? 1 : 0;
    d.tw = tw;
    d.sw = sw;
    d.jpart = jpart.as<double>();
    d.dt = par.dt;
    d.nu = par.nu;
    for (int c = 0; c < 3; ++c) d.uinf[c] = par.uinf[c];
    d.rec_scale = rec_scale();
    // the pair kernels' sums carry kq (Q, F') and kz (Zs)
    const PairScale ps = pair_scale(kind, par.precision);
    const long double kq = ps.q, kz = ps.z;
    if (uniform) {
        // sigma^2 = (s^2 + s^2)/2 = s^2; records hold x / sigma, so
        // d = sigma d^, q = sigma^-3 Q, F = sigma^-5 F', Z = sigma^-3 Zs
        long double s2 = (long double)sigma * sigma;
        long double sg = sqrtl(s2);
        long double s3 = s2 * sg, s5 = s3 * s2;
        d.inv_s2 = 1.0;
        d.su = (double)(1.0L / (s2 * kq));
        d.sq = (double)(1.0L / (s3 * kq));
        d.sd = (double)(2.0L * par.nu / (s5 * kz));
    } else {
        d.inv_s2 = 0.0;
        d.su = d.sq = (double)(1.0L / kq);
        d.sd = (double)(2.0L * par.nu / kz);
    }
    return d;
}

// ------------------------------------------------------------ dispatch
typedef void (*PairFn)(const JobDev *, const Item *, const int *, int *);

// Variants compiled: every (block, targets per thread, unroll) for the
// HIGHORDER kernel; the default variant for GAUSSIAN.
#define PW_VARIANTS(X, KIND)                                                                    \
    X(KIND, 64, 1, 1) X(KIND, 64, 1, 2) X(KIND, 64, 1, 4) X(KIND, 64, 2, 1) X(KIND, 64, 2, 2)   \
    X(KIND, 64, 2, 4) X(KIND, 128, 1, 1) X(KIND, 128, 1, 2) X(KIND, 128, 1, 4)                \
    X(KIND, 128, 2, 1) X(KIND, 128, 2, 2) X(KIND, 128, 2, 4) X(KIND, 256, 1, 1)                \
    X(KIND, 256, 1, 2) X(KIND, 256, 1, 4) X(KIND, 256, 2, 1) X(KIND, 256, 2, 2)                \
    X(KIND, 256, 2, 4)

// Symmetric binary64 pass: every (block, targets per thread) for both kernels.
static PairFn pair_fn_sym64(int kind, int tb, int tpt) {
#define PW_CASES(K, B, T) \
    if (tb == B && tpt == T) return pair64s_kernel<K, B, T>;
    if (kind == KIND_HIGHORDER) {
        PW_CASES(KIND_HIGHORDER, 64, 1) PW_CASES(KIND_HIGHORDER, 64, 2)
        PW_CASES(KIND_HIGHORDER, 128, 1) PW_CASES(KIND_HIGHORDER, 128, 2)
        PW_CASES(KIND_HIGHORDER, 256, 1) PW_CASES(KIND_HIGHORDER, 256, 2)
    } else {
        PW_CASES(KIND_GAUSSIAN, 64, 1) PW_CASES(KIND_GAUSSIAN, 128, 1)
        PW_CASES(KIND_GAUSSIAN, 256, 1)
    }
#undef PW_CASES
    return nullptr;
}

// Symmetric binary32 pass: every (block, targets per thread) for both kernels.
static PairFn pair_fn_sym(int kind, int tb, int tpt) {
#define PW_CASES(K, B, T) \
    if (tb == B && tpt == T) return pair32s_kernel<K, B, T>;
    if (kind == KIND_HIGHORDER) {
        PW_CASES(KIND_HIGHORDER, 64, 1) PW_CASES(KIND_HIGHORDER, 64, 2)
        PW_CASES(KIND_HIGHORDER, 128, 1) PW_CASES(KIND_HIGHORDER, 128, 2)
        PW_CASES(KIND_HIGHORDER, 256, 1) PW_CASES(KIND_HIGHORDER, 256, 2)
        PW_CASES(KIND_HIGHORDER, 512, 1)
    } else {
        PW_CASES(KIND_GAUSSIAN, 64, 1) PW_CASES(KIND_GAUSSIAN, 128, 1)
        PW_CASES(KIND_GAUSSIAN, 256, 1)
    }
#undef PW_CASES
    return nullptr;
}

static PairFn pair_fn(int prec, int kind, int tb, int tpt, int unr) {
#define PW_CASE64(K, B, T, U) \
    if (tb == B && tpt == T && unr == U) return pair64_kernel<K, B, T, U>;
#define PW_CASE32(K, B, T, U) \
    if (tb == B && tpt == T && unr == U) return pair32_kernel<K, B, T, U>;
    if (kind == KIND_HIGHORDER) {
        if (prec == PREC_FP64) {
            PW_VARIANTS(PW_CASE64, KIND_HIGHORDER)
        } else {
            PW_VARIANTS(PW_CASE32, KIND_HIGHORDER)
        }
    } else {
        if (prec == PREC_FP64) {
            PW_CASE64(KIND_GAUSSIAN, 64, 1, 2)
            PW_CASE64(KIND_GAUSSIAN, 128, 1, 2)
            PW_CASE64(KIND_GAUSSIAN, 256, 1, 2)
        } else {
            PW_CASE32(KIND_GAUSSIAN, 64, 1, 4)
            PW_CASE32(KIND_GAUSSIAN, 128, 1, 4)
            PW_CASE32(KIND_GAUSSIAN, 256, 1, 4)
        }
    }
#undef PW_CASE64
#undef PW_CASE32
    return nullptr;
}

bool variant_exists(int kind, int tb, int tpt, int unr64, int unr32, int sym32, int sym64,
                    int tpt32) {
    const int t32 = tpt32 ? tpt32 : tpt, tb64 = block_for(tb, PREC_FP64);
    return (sym64 ? pair_fn_sym64(kind, tb64, tpt) != nullptr
                  : pair_fn(PREC_FP64, kind, tb64, tpt, unr64) != nullptr) &&
           (sym32 ? pair_fn_sym(kind, tb, t32) != nullptr
                  : pair_fn(PREC_FP32C, kind, tb, t32, unr32) != nullptr);
}

Engine::Engine(int device, int kind, const Variant &v)
    : device_(device), kind_(kind), tb_(block_for(v.block, PREC_FP64)), tpt_(v.tpt),
      unr64_(v.unroll64), unr32_(v.unroll32), sym32_(v.sym32 ? 1 : 0), sym64_(v.sym64 ? 1 : 0),
      tb32_(block_for(v.block, PREC_FP32C)), tpt32_(v.tpt32 ? v.tpt32 : v.tpt) {
    if (tb32_ != 64 && tb32_ != 128 && tb32_ != 256 && tb32_ != 512)
        throw Error(PW_EINVAL, "block must be 64, 128, 256 or 512");
    if (kind_ != KIND_HIGHORDER && kind_ != KIND_GAUSSIAN) throw Error(PW_EINVAL, "unknown kernel");
    fn64_ = sym64_ ? pair_fn_sym64(kind_, tb_, tpt_) : pair_fn(PREC_FP64, kind_, tb_, tpt_, unr64_);
    fn32_ = sym32_ ? pair_fn_sym(kind_, tb32_, tpt32_) : pair_fn(PREC_FP32C, kind_, tb32_, tpt32_, unr32_);
    if (!fn64_ || !fn32_) throw Error(PW_EINVAL, "no compiled kernel variant for these parameters");
    int ndev = 0;
    PW_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw Error(PW_EINVAL, "no such HIP device");
    PW_HIP(hipSetDevice(device_));
    PW_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    PW_HIP(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, device_));
    int b64 = 0, b32 = 0;
    PW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b64, (const void *)fn64_, tb_, 0));
    PW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b32, (const void *)fn32_, tb32_, 0));
    grid64_ = std::max(1, b64) * cus_;
    grid32_ = std::max(1, b32) * cus_;
    gridc_ = 4 * cus_;
    counters_.reset(2 * sizeof(int));
    PW_HIP(hipMemsetAsync(counters_.p, 0, 2 * sizeof(int), stream_));
    flags_.reset(kMaxJobSlots * sizeof(int));
    PW_HIP(hipMemsetAsync(flags_.p, 0, flags_.bytes, stream_));
    flags_host_.ensure(flags_.bytes);
    for (int k = kMaxJobSlots - 1; k >= 0; --k) free_slots_.push_back(k);
    PW_HIP(hipEventCreate(&ev_begin_));
    PW_HIP(hipEventCreate(&ev_end_));
    PW_HIP(hipStreamSynchronize(stream_));
}

Engine::~Engine() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (int r = 0; r < 4; ++r) {
        if (pgexec_[r]) (void)hipGraphExecDestroy(pgexec_[r]);
        if (pgraph_[r]) (void)hipGraphDestroy(pgraph_[r]);
    }
    if (upload_) {
        (void)hipStreamSynchronize(upload_);
        (void)hipStreamDestroy(upload_);
    }
    for (auto e : ev_pool_) (void)hipEventDestroy(e);
    if (ev_begin_) (void)hipEventDestroy(ev_begin_);
    if (ev_end_) (void)hipEventDestroy(ev_end_);
    jobs_.clear();
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::sync() { PW_HIP(hipStreamSynchronize(stream_)); }

Job *Engine::find(int64_t id) {
    for (auto &j : jobs_)
        if (j->id == id) return j.get();
    return nullptr;
}

hipStream_t Engine::upload_stream() {
    if (!upload_) PW_HIP(hipStreamCreateWithFlags(&upload_, hipStreamNonBlocking));
    return upload_;
}

Job *Engine::add_job(int64_t id, size_t n, const double *pos, const double *vort,
                     const double *radius, const double *volume, const JobParams &par,
                     size_t chunk_override, bool async_upload) {
    if (n == 0 || !pos || !vort || !radius || !volume) throw Error(PW_EINVAL, "empty job or null array");
    if (n > (size_t)INT32_MAX / 2) throw Error(PW_EINVAL, "too many particles");
    if (par.precision != PREC_FP64 && par.precision != PREC_FP32C)
        throw Error(PW_EINVAL, "unknown precision");
    if (free_slots_.empty()) throw Error(PW_ENOMEM, "too many jobs in one engine");
    PW_HIP(hipSetDevice(device_));
    // the upload runs on the compute stream, or beside it
    hipStream_t s = async_upload ? upload_stream() : stream_;
    DevBuf &dst = async_upload ? ustage_ : dstage_;
    auto j = std::make_unique<Job>();
    j->id = id;
    j->n = n;
    j->npad = round_up(n, 256);
    j->par = par;
    j->kind = kind_;
    j->sym = par.precision == PREC_FP32C ? sym32_ != 0 : sym64_ != 0;
    const int tpt = tpt_of(par.precision), tb = tb_of(par.precision);
    size_t L = chunk_override ? chunk_override : chunk_len(n, tb, tpt, j->sym, par.chunks);
    if (L % (size_t)tb != 0) throw Error(PW_EINVAL, "chunk must be a multiple of the block size");
    j->chunk_len = (int)L;
    j->nchunk = (int)((n + L - 1) / L);
    j->ntile = (int)((n + tb * tpt - 1) / (tb * tpt));
    j->chunk_override = chunk_override;
    j->tw = tb * tpt;
    j->sw = (int)strip_for_sym(L, tb, tpt);
    j->uniform = true;
    for (size_t i = 1; i < n; ++i)
        if (std::memcmp(&radius[i], &radius[0], sizeof(double)) != 0) {
            j->uniform = false;
            break;
        }
    j->sigma = radius[0];
    const size_t np = j->npad;
    j->rec.reset(np * rec_bytes(par.precision));
    j->soa.reset(5 * 6 * np * sizeof(double));
    j->vol.reset(2 * np * sizeof(double));  // volume | radius
    alloc_part(j.get());
    j->bytes = j->rec.bytes + j->soa.bytes + j->part.bytes + j->vol.bytes + sizeof(int) +
               j->jpart.bytes;
    // host-side staging of records, state and constants
    size_t need = std::max(np * rec_bytes(par.precision), 6 * np * sizeof(double));
    stage_.ensure(std::max(need, 2 * np * sizeof(double)));
    std::memset(stage_.p, 0, np * rec_bytes(par.precision));
    fill_records(stage_.p, par.precision, n, pos, vort, radius, volume, j->rec_scale());
    PW_HIP(hipMemcpyAsync(j->rec.p, stage_.p, np * rec_bytes(par.precision), hipMemcpyHostToDevice,
                          s));
    PW_HIP(hipStreamSynchronize(s));
    double *v = stage_.as<double>();
    std::memset(v, 0, 2 * np * sizeof(double));
    for (size_t i = 0; i < n; ++i) {
        v[i] = volume[i];
        v[np + i] = radius[i];
    }
    PW_HIP(hipMemcpyAsync(j->vol.p, v, 2 * np * sizeof(double), hipMemcpyHostToDevice, s));
    PW_HIP(hipStreamSynchronize(s));
    PW_HIP(hipMemsetAsync(j->soa.p, 0, j->soa.bytes, s));
    j->slot = free_slots_.back();
    j->flag = flags_.as<int>() + j->slot;
    PW_HIP(hipMemsetAsync(j->flag, 0, sizeof(int), s));
    upload_soa(j.get(), pos, vort, s, dst);
    free_slots_.pop_back();
    Job *raw = j.get();
    jobs_.push_back(std::move(j));
    plan_dirty_ = true;
    return raw;
}

void Engine::append(Job *j, size_t m, const double *pos, const double *vort, const double *radius,
                    const double *volume) {
    if (m == 0) return;
    if (!pos || !vort || !radius || !volume) throw Error(PW_EINVAL, "null array");
    const size_t n = j->n, n2 = n + m;
    if (n2 > (size_t)INT32_MAX / 2) throw Error(PW_EINVAL, "too many particles");
    PW_HIP(hipSetDevice(device_));
    PW_HIP(hipStreamSynchronize(stream_));
    const size_t rb = rec_bytes(j->par.precision);
    if (n2 > j->npad) {  // grow: new stride, keep the device contents
        const size_t np = j->npad, np2 = round_up(std::max(2 * np, n2), 256);
        DevBuf rec(np2 * rb), soa(5 * 6 * np2 * sizeof(double)), vol(2 * np2 * sizeof(double));
        PW_HIP(hipMemsetAsync(rec.p, 0, rec.bytes, stream_));
        PW_HIP(hipMemsetAsync(soa.p, 0, soa.bytes, stream_));
        PW_HIP(hipMemsetAsync(vol.p, 0, vol.bytes, stream_));
        PW_HIP(hipMemcpyAsync(rec.p, j->rec.p, n * rb, hipMemcpyDeviceToDevice, stream_));
        for (size_t b = 0; b < 30; ++b)
            PW_HIP(hipMemcpyAsync(soa.as<double>() + b * np2, j->soa.as<double>() + b * np,
                                  n * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        for (size_t b = 0; b < 2; ++b)
            PW_HIP(hipMemcpyAsync(vol.as<double>() + b * np2, j->vol.as<double>() + b * np,
                                  n * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        PW_HIP(hipStreamSynchronize(stream_));
        std::swap(j->rec.p, rec.p);
        std::swap(j->rec.bytes, rec.bytes);
        std::swap(j->soa.p, soa.p);
        std::swap(j->soa.bytes, soa.bytes);
        std::swap(j->vol.p, vol.p);
        std::swap(j->vol.bytes, vol.bytes);
        j->npad = np2;
    }
    const size_t np = j->npad;
    // new particles' records, state (x0 = xs between steps), volume, radius
    stage_.ensure(std::max(m * rb, 8 * m * sizeof(double)));
    bool uni = j->uniform;
    for (size_t i = 0; i < m && uni; ++i)
        if (std::memcmp(&radius[i], &j->sigma, sizeof(double)) != 0) uni = false;
    if (uni != j->uniform) {  // the job loses its uniform core: records back to plain positions
        j->uniform = false;
        hipLaunchKernelGGL(refresh_rec_kernel, dim3((n + 255) / 256), dim3(256), 0, stream_,
                           j->dev());
        PW_HIP(hipGetLastError());
        PW_HIP(hipStreamSynchronize(stream_));
    }
    fill_records(stage_.p, j->par.precision, m, pos, vort, radius, volume, j->rec_scale());
    PW_HIP(hipMemcpyAsync(j->rec.as<char>() + n * rb, stage_.p, m * rb, hipMemcpyHostToDevice,
                          stream_));
    PW_HIP(hipStreamSynchronize(stream_));
    double *h = stage_.as<double>();
    for (size_t i = 0; i < m; ++i)
        for (int c = 0; c < 3; ++c) {
            h[c * m + i] = pos[3 * i + c];
            h[(3 + c) * m + i] = vort[3 * i + c];
        }
    for (size_t i = 0; i < m; ++i) {
        h[6 * m + i] = volume[i];
        h[7 * m + i] = radius[i];
    }
    double *soa = j->soa.as<double>();
    for (int c = 0; c < 6; ++c) {
        PW_HIP(hipMemcpyAsync(soa + c * np + n, h + c * m, m * sizeof(double),
                              hipMemcpyHostToDevice, stream_));  // x0
        PW_HIP(hipMemcpyAsync(soa + (6 + c) * np + n, h + c * m, m * sizeof(double),
                              hipMemcpyHostToDevice, stream_));  // xs
    }
    PW_HIP(hipMemcpyAsync(j->vol.as<double>() + n, h + 6 * m, m * sizeof(double),
                          hipMemcpyHostToDevice, stream_));
    PW_HIP(hipMemcpyAsync(j->vol.as<double>() + np + n, h + 7 * m, m * sizeof(double),
                          hipMemcpyHostToDevice, stream_));
    PW_HIP(hipStreamSynchronize(stream_));
    j->n = n2;
    const int tpt = tpt_of(j->par.precision), tb = tb_of(j->par.precision);
    const size_t L = j->chunk_override ? j->chunk_override
                                       : chunk_len(n2, tb, tpt, j->sym, j->par.chunks);
    j->chunk_len = (int)L;
    j->nchunk = (int)((n2 + L - 1) / L);
    j->ntile = (int)((n2 + tb * tpt - 1) / (tb * tpt));
    j->sw = (int)strip_for_sym(L, tb, tpt);
    alloc_part(j);
    j->bytes = j->rec.bytes + j->soa.bytes + j->part.bytes + j->vol.bytes + sizeof(int) +
               j->jpart.bytes;
    plan_dirty_ = true;
}

// Partial-sum buffers of a job: the target-side rows (per source chunk) and,
// on the symmetric pass, the source-side slot rows (per strip).  Scratch,
// rewritten by every pass; grown only (particle emission reallocates them when
// the chunk plan needs more than they hold, not on every append).
void Engine::alloc_part(Job *j) {
    const size_t np = j->npad;
    const size_t pneed =
        part_count(j->n, np, (size_t)j->chunk_len, (size_t)j->tw, j->sym) * sizeof(double);
    if (pneed > j->part.bytes) j->part.reset(pneed);
    if (!j->sym) return;
    const size_t need = jpart_count(np, np, (size_t)j->sw) * sizeof(double);
    if (need > j->jpart.bytes) j->jpart.reset(need);
}

void Engine::remove_job(int64_t id) {
    PW_HIP(hipSetDevice(device_));
    PW_HIP(hipStreamSynchronize(stream_));
    for (size_t k = 0; k < jobs_.size(); ++k)
        if (jobs_[k]->id == id) {
            free_slots_.push_back(jobs_[k]->slot);
            jobs_.erase(jobs_.begin() + k);
            plan_dirty_ = true;
            return;
        }
    throw Error(PW_ENOTFOUND, "unknown job");
}

void Engine::set_params(Job *j, const JobParams &par) {
    if (par.precision != j->par.precision) throw Error(PW_EINVAL, "precision of a job is fixed");
    if (par.chunks != j->par.chunks) throw Error(PW_EINVAL, "chunk count of a job is fixed");
    j->par = par;
    plan_dirty_ = true;
}

// Work items of job j (table index k): pair items into dst, combine items
// into ci.
static void job_items(const Job &j, int k, std::vector<Item> &dst, std::vector<CItem> &ci) {
    if (!j.sym) {
        for (int c = 0; c < j.nchunk; ++c)
            for (int t = 0; t < j.ntile; ++t) dst.push_back(Item{k, t, c, j.chunk_len});
    } else {
        // strip s (blocks of tw targets from s sw on) pairs with the
        // sources from its start on: chunks from the one holding that start;
        // cost = sources its blocks visit
        const size_t nstrip = (j.n + j.sw - 1) / j.sw;
        for (size_t t = 0; t < nstrip; ++t) {
            const size_t s0 = t * j.sw, s1 = std::min(j.n, s0 + j.sw);
            for (int c = (int)(s0 / j.chunk_len); c < j.nchunk; ++c) {
                const size_t clo = (size_t)c * j.chunk_len;
                const size_t chi = std::min(j.n, clo + j.chunk_len);
                size_t cost = 0;
                for (size_t b0 = s0; b0 < s1; b0 += j.tw)
                    if (std::max(b0, clo) < chi) cost += chi - std::max(b0, clo);
                dst.push_back(Item{k, (int)t, c, (int)std::min<size_t>(cost, INT32_MAX)});
            }
        }
    }
    if (j.sym) {  // 64 particles per item
        for (int t = 0; t < (int)((j.n + 63) / 64); ++t) ci.push_back(CItem{k, t});
    } else {
        for (int t = 0; t < (int)((j.n + 255) / 256); ++t) ci.push_back(CItem{k, t});
    }
}

// larger items first (their cost is tile x chunk length): shorter tail
static void order_pair_items(std::vector<Item> &v) {
    std::stable_sort(v.begin(), v.end(), [](const Item &a, const Item &b) { return a.pad > b.pad; });
}

// Combine items: the combine grid takes item k, k + G, k + 2G, ... per block
// (G blocks, all resident); sorted by the values each reads, largest first,
// and every second round of G reversed, so that each block's items add up to
// about the same work.  tab[c.job] is the item's job.
static void order_combine_items(std::vector<CItem> &ci, const std::vector<const Job *> &tab,
                                size_t G) {
    auto cost = [&](const CItem &c) -> size_t {
        const Job &j = *tab[(size_t)c.job];
        if (!j.sym) return (size_t)NCOMP * 256 * (size_t)j.nchunk;
        const size_t b0 = (size_t)c.tile * 64 / (size_t)j.tw;
        const size_t s0 = (size_t)c.tile * 64 / (size_t)j.sw;
        const size_t cf = b0 * (size_t)j.tw / (size_t)j.chunk_len;
        return (size_t)NCOMP_S * 64 * (s0 + 1 + (size_t)j.nchunk - cf);
    };
    std::vector<std::pair<size_t, CItem>> kc;
    kc.reserve(ci.size());
    for (const CItem &c : ci) kc.push_back({cost(c), c});
    std::stable_sort(kc.begin(), kc.end(),
                     [](const std::pair<size_t, CItem> &a, const std::pair<size_t, CItem> &b) {
                         return a.first > b.first;
                     });
    G = std::max<size_t>(1, G);
    for (size_t r = 0; r * G < kc.size(); ++r)
        if (r & 1) std::reverse(kc.begin() + r * G, kc.begin() + std::min(kc.size(), (r + 1) * G));
    for (size_t k = 0; k < ci.size(); ++k) ci[k] = kc[k].second;
}

void Engine::build_plan() {
    PW_HIP(hipSetDevice(device_));
    // queued kernels may still read the current tables and their host copy
    PW_HIP(hipStreamSynchronize(stream_));
    std::vector<JobDev> jd;
    std::vector<Item> it64, it32;
    std::vector<CItem> ci;
    std::vector<const Job *> tab;
    for (size_t k = 0; k < jobs_.size(); ++k) {
        const Job &j = *jobs_[k];
        jd.push_back(j.dev());
        tab.push_back(&j);
        job_items(j, (int)k, j.par.precision == PREC_FP64 ? it64 : it32, ci);
    }
    order_pair_items(it64);
    order_pair_items(it32);
    order_combine_items(ci, tab, (size_t)gridc_);
    n64_ = (int)it64.size();
    n32_ = (int)it32.size();
    nci_ = (int)ci.size();
    const int g64 = std::max(1, std::min(grid64_, n64_)), g32 = std::max(1, std::min(grid32_, n32_)),
              gc = std::max(1, std::min(gridc_, nci_));
    // Region capacities only grow.  The region addresses are kernel arguments
    // baked into the captured step graph, so they must not move while only
    // the table contents change; a growth starts a new layout version, for
    // which the graph is captured again.
    const bool grow = !plan_.p || jd.size() > cap_jobs_ || it64.size() > cap_i64_ ||
                      it32.size() > cap_i32_ || ci.size() > cap_ci_;
    if (grow) {
        cap_jobs_ = std::max(cap_jobs_, round_up(std::max<size_t>(jd.size(), 1) * 2, 16));
        cap_i64_ = std::max(cap_i64_, round_up(std::max<size_t>(it64.size(), 1) * 2, 1024));
        cap_i32_ = std::max(cap_i32_, round_up(std::max<size_t>(it32.size(), 1) * 2, 1024));
        cap_ci_ = std::max(cap_ci_, round_up(std::max<size_t>(ci.size(), 1) * 2, 1024));
        o_i64_ = round_up(cap_jobs_ * sizeof(JobDev), 256);
        o_i32_ = round_up(o_i64_ + cap_i64_ * sizeof(Item), 256);
        o_ci_ = round_up(o_i32_ + cap_i32_ * sizeof(Item), 256);
        o_cnt_ = round_up(o_ci_ + cap_ci_ * sizeof(CItem), 256);
        plan_.reset(o_cnt_ + 256);
        plan_host_.ensure(plan_.bytes);
        ++layout_version_;
    }
    if (g64 != lgrid64_ || g32 != lgrid32_ || gc != lgridc_) {
        // the grid sizes are baked into a captured step too
        lgrid64_ = g64;
        lgrid32_ = g32;
        lgridc_ = gc;
        ++layout_version_;
    }
    char *h = plan_host_.as<char>();
    std::memcpy(h, jd.data(), jd.size() * sizeof(JobDev));
    std::memcpy(h + o_i64_, it64.data(), it64.size() * sizeof(Item));
    std::memcpy(h + o_i32_, it32.data(), it32.size() * sizeof(Item));
    std::memcpy(h + o_ci_, ci.data(), ci.size() * sizeof(CItem));
    int cnt[4] = {n64_, n32_, nci_, 0};
    std::memcpy(h + o_cnt_, cnt, sizeof cnt);
    PW_HIP(hipMemcpyAsync(plan_.p, h, o_cnt_ + 256, hipMemcpyHostToDevice, stream_));
    char *d = plan_.as<char>();
    d_jobs_ = (JobDev *)d;
    d_items64_ = (Item *)(d + o_i64_);
    d_items32_ = (Item *)(d + o_i32_);
    d_citems_ = (CItem *)(d + o_ci_);
    d_counts_ = (int *)(d + o_cnt_);
    plan_dirty_ = false;
}

double Engine::pairs_per_rhs() const {
    double p = 0.0;
    for (auto &j : jobs_) p += (double)j->n * (double)(j->n - 1);
    return p;
}

void Engine::launch_stage(int mode, bool timed) {
    int *cnt = counters_.as<int>();
    const bool graph = !timed;
    auto launch_pair = [&](int prec, int grid, const Item *items, const int *nitems, int *queue,
                           int count) {
        if (!graph && count == 0) return;
        PairFn fn = prec == PREC_FP64 ? fn64_ : fn32_;
        hipEvent_t a = nullptr, b = nullptr;
        if (timed) {
            size_t k = (size_t)pair_launches * 2;
            while (ev_pool_.size() < k + 2) {
                hipEvent_t e;
                PW_HIP(hipEventCreate(&e));
                ev_pool_.push_back(e);
            }
            a = ev_pool_[k];
            b = ev_pool_[k + 1];
            PW_HIP(hipEventRecord(a, stream_));
        }
        hipLaunchKernelGGL(fn, dim3(grid), dim3(tb_of(prec)), 0, stream_, d_jobs_, items, nitems,
                           queue);
        PW_HIP(hipGetLastError());
        if (timed) {
            PW_HIP(hipEventRecord(b, stream_));
            ++pair_launches;
        }
    };
    launch_pair(PREC_FP64, lgrid64_, d_items64_, d_counts_ + 0, cnt + 0, n64_);
    launch_pair(PREC_FP32C, lgrid32_, d_items32_, d_counts_ + 1, cnt + 1, n32_);
    hipLaunchKernelGGL(combine_kernel, dim3(lgridc_), dim3(256), 0, stream_, (const JobDev *)d_jobs_,
                       (const CItem *)d_citems_, (const int *)(d_counts_ + 2), cnt, mode);
    PW_HIP(hipGetLastError());
}

static void begin_timing(hipEvent_t e, hipStream_t s) { PW_HIP(hipEventRecord(e, s)); }

void Engine::solve() {
    PW_HIP(hipSetDevice(device_));
    if (plan_dirty_) build_plan();
    pair_launches = 0;
    begin_timing(ev_begin_, stream_);
    launch_stage(MODE_SOLVE, true);
    PW_HIP(hipEventRecord(ev_end_, stream_));
    PW_HIP(hipEventSynchronize(ev_end_));
    float ms = 0.f;
    PW_HIP(hipEventElapsedTime(&ms, ev_begin_, ev_end_));
    last_total_ms = ms;
    last_pair_ms = 0.0;
    for (int64_t k = 0; k < pair_launches; ++k) {
        PW_HIP(hipEventElapsedTime(&ms, ev_pool_[2 * k], ev_pool_[2 * k + 1]));
        last_pair_ms += ms;
    }
    last_pairs = pairs_per_rhs();
}

void Engine::capture_graph() {
    if (graph_exec_) {
        (void)hipGraphExecDestroy(graph_exec_);
        graph_exec_ = nullptr;
    }
    if (graph_) {
        (void)hipGraphDestroy(graph_);
        graph_ = nullptr;
    }
    PW_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
        launch_stage(MODE_RK4_1, false);
        launch_stage(MODE_RK4_2, false);
        launch_stage(MODE_RK4_3, false);
        launch_stage(MODE_RK4_4, false);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    PW_HIP(hipStreamEndCapture(stream_, &graph_));
    PW_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
    graph_layout_ = layout_version_;
}

void Engine::rk4(size_t nsteps, bool use_graph) {
    PW_HIP(hipSetDevice(device_));
    if (plan_dirty_) build_plan();
    if (jobs_.empty() || nsteps == 0) {
        last_total_ms = last_pair_ms = last_pairs = 0.0;
        return;
    }
    pair_launches = 0;
    if (use_graph && (!graph_exec_ || graph_layout_ != layout_version_)) capture_graph();
    begin_timing(ev_begin_, stream_);
    for (size_t s = 0; s < nsteps; ++s) {
        if (use_graph) {
            PW_HIP(hipGraphLaunch(graph_exec_, stream_));
        } else {
            launch_stage(MODE_RK4_1, true);
            launch_stage(MODE_RK4_2, true);
            launch_stage(MODE_RK4_3, true);
            launch_stage(MODE_RK4_4, true);
        }
    }
    PW_HIP(hipEventRecord(ev_end_, stream_));
    PW_HIP(hipEventSynchronize(ev_end_));
    float ms = 0.f;
    PW_HIP(hipEventElapsedTime(&ms, ev_begin_, ev_end_));
    last_total_ms = ms;
    last_pair_ms = 0.0;
    for (int64_t k = 0; k < pair_launches; ++k) {
        PW_HIP(hipEventElapsedTime(&ms, ev_pool_[2 * k], ev_pool_[2 * k + 1]));
        last_pair_ms += ms;
    }
    last_pairs = 4.0 * (double)nsteps * pairs_per_rhs();
}

void Engine::euler(size_t nsteps) {
    PW_HIP(hipSetDevice(device_));
    if (plan_dirty_) build_plan();
    pair_launches = 0;
    begin_timing(ev_begin_, stream_);
    for (size_t s = 0; s < nsteps; ++s) launch_stage(MODE_EULER, true);
    PW_HIP(hipEventRecord(ev_end_, stream_));
    PW_HIP(hipEventSynchronize(ev_end_));
    float ms = 0.f;
    PW_HIP(hipEventElapsedTime(&ms, ev_begin_, ev_end_));
    last_total_ms = ms;
    last_pair_ms = 0.0;
    for (int64_t k = 0; k < pair_launches; ++k) {
        PW_HIP(hipEventElapsedTime(&ms, ev_pool_[2 * k], ev_pool_[2 * k + 1]));
        last_pair_ms += ms;
    }
    last_pairs = (double)nsteps * pairs_per_rhs();
}


// ------------------------------------------------------------ server pass
void Engine::launch_round(int r, bool timed) {
    char *d = pplan_.as<char>() + (size_t)r * p_round_;
    const JobDev *jd = (const JobDev *)d;
    const int *cnt = (const int *)(d + po_cnt_);
    int *q = counters_.as<int>();
    const bool graph = !timed;
    if (graph || pn64_[r] > 0) {
        hipLaunchKernelGGL(fn64_, dim3(grid64_), dim3(tb_), 0, stream_, jd,
                           (const Item *)(d + po_i64_), cnt + 0, q + 0);
        PW_HIP(hipGetLastError());
    }
    if (graph || pn32_[r] > 0) {
        hipLaunchKernelGGL(fn32_, dim3(grid32_), dim3(tb32_), 0, stream_, jd,
                           (const Item *)(d + po_i32_), cnt + 1, q + 1);
        PW_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(combine_kernel, dim3(gridc_), dim3(256), 0, stream_, jd,
                       (const CItem *)(d + po_ci_), cnt + 2, q, (int)MODE_PER_JOB);
    PW_HIP(hipGetLastError());
}

void Engine::capture_pass(int rounds) {
    const int g = rounds - 1;
    if (pgexec_[g]) {
        (void)hipGraphExecDestroy(pgexec_[g]);
        pgexec_[g] = nullptr;
    }
    if (pgraph_[g]) {
        (void)hipGraphDestroy(pgraph_[g]);
        pgraph_[g] = nullptr;
    }
    PW_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
        for (int r = 0; r < rounds; ++r) launch_round(r, false);
    } catch (...) {
        hipGraph_t gr = nullptr;
        (void)hipStreamEndCapture(stream_, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        throw;
    }
    PW_HIP(hipStreamEndCapture(stream_, &pgraph_[g]));
    PW_HIP(hipGraphInstantiate(&pgexec_[g], pgraph_[g], nullptr, nullptr, 0));
    pgraph_layout_[g] = playout_;
}

void Engine::pass_begin(const std::vector<PassJob> &pj, bool use_graph) {
    PW_HIP(hipSetDevice(device_));
    cur_pass_.clear();
    if (pj.empty()) return;
    int R = 0;
    for (const PassJob &p : pj) {
        if (!p.job || p.stages < 1 || p.first_stage < 0 || p.first_stage + p.stages > 4)
            throw Error(PW_EINVAL, "a pass runs 1 to 4 stages of a job within one step");
        if (p.parts < 1 || p.part < 0 || p.part >= p.parts || (p.parts > 1 && p.stages != 1))
            throw Error(PW_EINVAL, "a stage slice is one of `parts` of a single stage");
        R = std::max(R, p.stages);
    }
    // the rounds' work: job k runs in rounds 0 .. stages - 1
    std::vector<const Job *> tab;
    for (const PassJob &p : pj) tab.push_back(p.job);
    std::vector<std::vector<Item>> r64(R), r32(R);
    std::vector<std::vector<CItem>> rci(R);
    size_t m64 = 0, m32 = 0, mci = 0;
    double pairs = 0.0;
    for (int r = 0; r < R; ++r) {
        for (size_t k = 0; k < pj.size(); ++k) {
            if (pj[k].stages <= r) continue;
            const Job &j = *pj[k].job;
            std::vector<Item> &dst = j.par.precision == PREC_FP64 ? r64[r] : r32[r];
            const double pj_pairs = (double)j.n * (double)(j.n - 1);
            if (pj[k].parts == 1) {
                job_items(j, (int)k, dst, rci[r]);
                pairs += pj_pairs;
                continue;
            }
            // a slice: the job's items in their fixed order, cut where the
            // running cost (Item::pad) passes q / parts of the total, each cut
            // moved up to whole rounds of the persistent grid while every
            // later part keeps at least a round
            std::vector<Item> all;
            std::vector<CItem> ci;
            job_items(j, (int)k, all, ci);
            const size_t G = (size_t)(j.par.precision == PREC_FP64 ? grid64_ : grid32_);
            const int parts = pj[k].parts;
            double w_all = 0.0;
            for (const Item &it : all) w_all += (double)it.pad;
            std::vector<size_t> cut((size_t)parts + 1, 0);
            cut[(size_t)parts] = all.size();
            {
                double acc = 0.0;
                size_t i = 0;
                for (int q = 1; q < parts; ++q) {
                    const double goal = w_all * q / parts;
                    while (i < all.size() && acc + (double)all[i].pad <= goal) acc += (double)all[i++].pad;
                    size_t b = i;
                    const size_t br = round_up(std::max<size_t>(b, 1), G);
                    if (br + (size_t)(parts - q) * G <= all.size()) b = br;
                    cut[(size_t)q] = std::max(cut[(size_t)q - 1], std::min(b, all.size()));
                }
            }
            const size_t lo = cut[(size_t)pj[k].part], hi = cut[(size_t)pj[k].part + 1];
            double w = 0.0;
            for (size_t i = lo; i < hi; ++i) w += (double)all[i].pad;
            dst.insert(dst.end(), all.begin() + lo, all.begin() + hi);
            if (pj[k].part + 1 == pj[k].parts) rci[r].insert(rci[r].end(), ci.begin(), ci.end());
            pairs += w_all > 0.0 ? pj_pairs * w / w_all : 0.0;
        }
        order_pair_items(r64[r]);
        order_pair_items(r32[r]);
        order_combine_items(rci[r], tab, (size_t)gridc_);
        m64 = std::max(m64, r64[r].size());
        m32 = std::max(m32, r32[r].size());
        mci = std::max(mci, rci[r].size());
    }
    // regions only grow: a growth is a new layout, for which the pass graphs
    // are captured again
    if (!pplan_.p || pj.size() > pcap_jobs_ || m64 > pcap_i64_ || m32 > pcap_i32_ || mci > pcap_ci_) {
        pcap_jobs_ = std::max(pcap_jobs_, round_up(pj.size() * 2, 16));
        pcap_i64_ = std::max(pcap_i64_, round_up(std::max<size_t>(m64, 1) * 2, 1024));
        pcap_i32_ = std::max(pcap_i32_, round_up(std::max<size_t>(m32, 1) * 2, 1024));
        pcap_ci_ = std::max(pcap_ci_, round_up(std::max<size_t>(mci, 1) * 2, 1024));
        po_i64_ = round_up(pcap_jobs_ * sizeof(JobDev), 256);
        po_i32_ = round_up(po_i64_ + pcap_i64_ * sizeof(Item), 256);
        po_ci_ = round_up(po_i32_ + pcap_i32_ * sizeof(Item), 256);
        po_cnt_ = round_up(po_ci_ + pcap_ci_ * sizeof(CItem), 256);
        p_round_ = round_up(po_cnt_ + 256, 256);
        PW_HIP(hipStreamSynchronize(stream_));
        pplan_.reset(4 * p_round_);
        pplan_host_.ensure(pplan_.bytes);
        ++playout_;
    }
    char *h = pplan_host_.as<char>();
    int hi = 0;
    for (int r = 0; r < R; ++r) {
        char *b = h + (size_t)r * p_round_;
        JobDev *jd = (JobDev *)b;
        for (size_t k = 0; k < pj.size(); ++k) {
            jd[k] = pj[k].job->dev();
            jd[k].mode = pj[k].stages > r ? MODE_RK4_1 + pj[k].first_stage + r : MODE_SOLVE;
        }
        std::memcpy(b + po_i64_, r64[r].data(), r64[r].size() * sizeof(Item));
        std::memcpy(b + po_i32_, r32[r].data(), r32[r].size() * sizeof(Item));
        std::memcpy(b + po_ci_, rci[r].data(), rci[r].size() * sizeof(CItem));
        const int cnt[4] = {(int)r64[r].size(), (int)r32[r].size(), (int)rci[r].size(), 0};
        std::memcpy(b + po_cnt_, cnt, sizeof cnt);
        pn64_[r] = cnt[0];
        pn32_[r] = cnt[1];
    }
    for (const PassJob &p : pj) hi = std::max(hi, p.job->slot + 1);
    PW_HIP(hipMemcpyAsync(pplan_.p, h, (size_t)R * p_round_, hipMemcpyHostToDevice, stream_));
    PW_HIP(hipEventRecord(ev_begin_, stream_));
    if (use_graph) {
        if (!pgexec_[R - 1] || pgraph_layout_[R - 1] != playout_) capture_pass(R);
        PW_HIP(hipGraphLaunch(pgexec_[R - 1], stream_));
    } else {
        for (int r = 0; r < R; ++r) launch_round(r, true);
    }
    // every listed job's flag, in one copy
    PW_HIP(hipMemcpyAsync(flags_host_.p, flags_.p, (size_t)hi * sizeof(int), hipMemcpyDeviceToHost,
                          stream_));
    PW_HIP(hipEventRecord(ev_end_, stream_));
    cur_pass_ = pj;
    flag_hi_ = hi;
    pass_pairs_ = pairs;
}

void Engine::pass_end(std::vector<char> *nonfinite) {
    PW_HIP(hipSetDevice(device_));
    nonfinite->assign(cur_pass_.size(), 0);
    if (cur_pass_.empty()) {
        last_total_ms = last_pair_ms = last_pairs = 0.0;
        return;
    }
    PW_HIP(hipEventSynchronize(ev_end_));
    float ms = 0.f;
    PW_HIP(hipEventElapsedTime(&ms, ev_begin_, ev_end_));
    last_total_ms = ms;
    last_pair_ms = 0.0;
    last_pairs = pass_pairs_;
    const int *h = flags_host_.as<int>();
    bool any = false;
    for (size_t k = 0; k < cur_pass_.size(); ++k) {
        (*nonfinite)[k] = h[cur_pass_[k].job->slot] != 0;
        any = any || (*nonfinite)[k];
    }
    if (any) PW_HIP(hipMemsetAsync(flags_.p, 0, (size_t)flag_hi_ * sizeof(int), stream_));
    cur_pass_.clear();
}

void Engine::upload_state_soa(Job *j, const double *pos, const double *vort) {
    PW_HIP(hipSetDevice(device_));
    upload_soa(j, pos, vort, stream_, dstage_);
}

void Engine::upload_soa(Job *j, const double *pos, const double *vort, hipStream_t s,
                        DevBuf &stage) {
    const size_t np = j->npad, n = j->n;
    if (stage.bytes < 3 * n * sizeof(double)) {
        PW_HIP(hipStreamSynchronize(s));
        stage.reset(round_up(6 * np * sizeof(double), 4096));
    }
    JobDev d = j->dev();
    const double *src[2] = {pos, vort};
    for (int h = 0; h < 2; ++h) {
        PW_HIP(hipMemcpyAsync(stage.p, src[h], 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(aos_soa_kernel, dim3(grid_for(n)), dim3(256), 0, s, stage.as<double>(),
                           d.x0 + (size_t)3 * h * np, (int)n, (int)np, 0);
        PW_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(copy_soa_kernel, dim3(grid_for(6 * np)), dim3(256), 0, s,
                       (const double *)d.x0, d.xs, 6 * np);
    hipLaunchKernelGGL(refresh_rec_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d);
    PW_HIP(hipGetLastError());
    PW_HIP(hipStreamSynchronize(s));
}

void Engine::download_aos(Job *j, int which, double *out) { download_aos(j, which, out, 0, j->n); }

void Engine::download_aos(Job *j, int which, double *out, size_t first, size_t count) {
    PW_HIP(hipSetDevice(device_));
    const size_t np = j->npad;
    if (first > j->n || count > j->n - first) throw Error(PW_EINVAL, "particle range out of bounds");
    if (count == 0) return;
    if (dstage_.bytes < 3 * count * sizeof(double)) {
        PW_HIP(hipStreamSynchronize(stream_));
        dstage_.reset(round_up(6 * np * sizeof(double), 4096));
    }
    JobDev d = j->dev();
    double *src = which == 0 ? d.x0 : which == 1 ? d.x0 + 3 * np : which == 2 ? d.rate : d.rate + 3 * np;
    hipLaunchKernelGGL(aos_soa_kernel, dim3(grid_for(count)), dim3(256), 0, stream_,
                       dstage_.as<double>(), src + first, (int)count, (int)np, 1);
    PW_HIP(hipGetLastError());
    PW_HIP(hipMemcpyAsync(out, dstage_.p, 3 * count * sizeof(double), hipMemcpyDeviceToHost,
                          stream_));
    PW_HIP(hipStreamSynchronize(stream_));
}

void Engine::snapshot_aos(Job *j, double *dst) {
    PW_HIP(hipSetDevice(device_));
    const size_t np = j->npad, n = j->n;
    JobDev d = j->dev();
    for (int h = 0; h < 2; ++h)
        hipLaunchKernelGGL(aos_soa_kernel, dim3(grid_for(n)), dim3(256), 0, stream_,
                           dst + (size_t)3 * h * n, d.x0 + (size_t)3 * h * np, (int)n, (int)np, 1);
    PW_HIP(hipGetLastError());
}

void Engine::upload_state_layout(Job *j, const double *state, const std::vector<int64_t> &woff) {
    PW_HIP(hipSetDevice(device_));
    const size_t np = j->npad, n = j->n;
    const size_t wbytes = woff.size() * sizeof(int64_t);
    const size_t need = 6 * n * sizeof(double) + 256 + wbytes;
    if (dstage_.bytes < need) {
        PW_HIP(hipStreamSynchronize(stream_));
        dstage_.reset(round_up(need + 6 * np * sizeof(double), 4096));
    }
    char *base = dstage_.as<char>();
    int64_t *dwoff = (int64_t *)(base + round_up(6 * n * sizeof(double), 256));
    PW_HIP(hipMemcpyAsync(base, state, 6 * n * sizeof(double), hipMemcpyHostToDevice, stream_));
    PW_HIP(hipMemcpyAsync(dwoff, woff.data(), wbytes, hipMemcpyHostToDevice, stream_));
    JobDev d = j->dev();
    hipLaunchKernelGGL(state_layout_kernel, dim3(grid_for(n)), dim3(256), 0, stream_,
                       (double *)base, d.x0, (int)np, (const int64_t *)dwoff, (int)woff.size() - 1, 0);
    hipLaunchKernelGGL(copy_soa_kernel, dim3(grid_for(6 * np)), dim3(256), 0, stream_,
                       (const double *)d.x0, d.xs, 6 * np);
    hipLaunchKernelGGL(refresh_rec_kernel, dim3((n + 255) / 256), dim3(256), 0, stream_, d);
    PW_HIP(hipGetLastError());
    PW_HIP(hipStreamSynchronize(stream_));
}

void Engine::download_state_layout(Job *j, int which, double *state,
                                   const std::vector<int64_t> &woff) {
    PW_HIP(hipSetDevice(device_));
    const size_t np = j->npad, n = j->n;
    const size_t wbytes = woff.size() * sizeof(int64_t);
    const size_t need = 6 * n * sizeof(double) + 256 + wbytes;
    if (dstage_.bytes < need) {
        PW_HIP(hipStreamSynchronize(stream_));
        dstage_.reset(round_up(need + 6 * np * sizeof(double), 4096));
    }
    char *base = dstage_.as<char>();
    int64_t *dwoff = (int64_t *)(base + round_up(6 * n * sizeof(double), 256));
    PW_HIP(hipMemcpyAsync(dwoff, woff.data(), wbytes, hipMemcpyHostToDevice, stream_));
    JobDev d = j->dev();
    hipLaunchKernelGGL(state_layout_kernel, dim3(grid_for(n)), dim3(256), 0, stream_,
                       (double *)base, which == 0 ? d.x0 : d.rate, (int)np,
                       (const int64_t *)dwoff, (int)woff.size() - 1, 1);
    PW_HIP(hipGetLastError());
    PW_HIP(hipMemcpyAsync(state, base, 6 * n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    PW_HIP(hipStreamSynchronize(stream_));
}

void Engine::diagnostics(Job *j, Diagnostics *out, bool quadratic) {
    PW_HIP(hipSetDevice(device_));
    const size_t np = j->npad, n = j->n;
    JobDev d = j->dev();
    const int nt = (int)((n + 255) / 256);
    {
        DevBuf buf(nt * 9 * sizeof(double));
        hipLaunchKernelGGL(linear_diag_kernel, dim3(nt), dim3(256), 0, stream_, (const double *)d.x0,
                           (int)n, (int)np, buf.as<double>());
        PW_HIP(hipGetLastError());
        std::vector<double> h(nt * 9);
        PW_HIP(hipMemcpyAsync(h.data(), buf.p, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                              stream_));
        PW_HIP(hipStreamSynchronize(stream_));
        double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int t = 0; t < nt; ++t)
            for (int m = 0; m < 9; ++m) s[m] += h[t * 9 + m];
        for (int c = 0; c < 3; ++c) {
            out->omega[c] = s[c];
            out->impulse[c] = s[3 + c] * 0.5;
            out->angular[c] = s[6 + c] * (1.0 / 3.0);
        }
    }
    out->ke = out->enstrophy = out->helicity = 0.0;
    if (!quadratic) return;
    const int L = (int)chunk_for(n, 256);
    const int nc = (int)((n + L - 1) / L);
    DevBuf part((size_t)nc * 3 * np * sizeof(double)), sum(3 * np * sizeof(double));
    if (kind_ == KIND_GAUSSIAN)
        hipLaunchKernelGGL(diag_pair_kernel<KIND_GAUSSIAN>, dim3(nt, nc), dim3(256), 0, stream_,
                           (const double *)d.x0, (const double *)(d.vol + np), (int)n, (int)np, L,
                           part.as<double>());
    else
        hipLaunchKernelGGL(diag_pair_kernel<KIND_HIGHORDER>, dim3(nt, nc), dim3(256), 0, stream_,
                           (const double *)d.x0, (const double *)(d.vol + np), (int)n, (int)np, L,
                           part.as<double>());
    PW_HIP(hipGetLastError());
    hipLaunchKernelGGL(chunk_sum_kernel, dim3(nt), dim3(256), 0, stream_, (const double *)part.p,
                       sum.as<double>(), nc, 3, (int)n, (int)np);
    PW_HIP(hipGetLastError());
    std::vector<double> h(3 * np);
    PW_HIP(hipMemcpyAsync(h.data(), sum.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    PW_HIP(hipStreamSynchronize(stream_));
    double e = 0.0, en = 0.0, he = 0.0;
    for (size_t i = 0; i < n; ++i) {
        e += h[i];
        en += h[np + i];
        he += h[2 * np + i];
    }
    out->ke = 2.0 * e;
    out->enstrophy = en;
    out->helicity = he;
}

void Engine::influence(Job *j, double *field_aos) {
    PW_HIP(hipSetDevice(device_));
    const size_t np = j->npad, n = j->n;
    JobDev d = j->dev();
    const int nt = (int)((n + 255) / 256);
    const int L = (int)chunk_for(n, 256);
    const int nc = (int)((n + L - 1) / L);
    DevBuf part((size_t)nc * 3 * np * sizeof(double)), sum(3 * np * sizeof(double));
    if (kind_ == KIND_GAUSSIAN)
        hipLaunchKernelGGL(influence_kernel<KIND_GAUSSIAN>, dim3(nt, nc), dim3(256), 0, stream_,
                           (const double *)d.x0, (const double *)(d.vol + np), (int)n, (int)np, L,
                           part.as<double>());
    else
        hipLaunchKernelGGL(influence_kernel<KIND_HIGHORDER>, dim3(nt, nc), dim3(256), 0, stream_,
                           (const double *)d.x0, (const double *)(d.vol + np), (int)n, (int)np, L,
                           part.as<double>());
    PW_HIP(hipGetLastError());
    hipLaunchKernelGGL(chunk_sum_kernel, dim3(nt), dim3(256), 0, stream_, (const double *)part.p,
                       sum.as<double>(), nc, 3, (int)n, (int)np);
    PW_HIP(hipGetLastError());
    if (dstage_.bytes < 3 * n * sizeof(double)) {
        PW_HIP(hipStreamSynchronize(stream_));
        dstage_.reset(round_up(6 * np * sizeof(double), 4096));
    }
    hipLaunchKernelGGL(aos_soa_kernel, dim3(grid_for(n)), dim3(256), 0, stream_,
                       dstage_.as<double>(), sum.as<double>(), (int)n, (int)np, 1);
    PW_HIP(hipGetLastError());
    PW_HIP(hipMemcpyAsync(field_aos, dstage_.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost,
                          stream_));
    PW_HIP(hipStreamSynchronize(stream_));
}

bool Engine::take_flag(Job *j) {
    PW_HIP(hipSetDevice(device_));
    int f = 0;
    PW_HIP(hipMemcpyAsync(&f, j->flag, sizeof(int), hipMemcpyDeviceToHost, stream_));
    PW_HIP(hipMemsetAsync(j->flag, 0, sizeof(int), stream_));
    PW_HIP(hipStreamSynchronize(stream_));
    return f != 0;
}

std::vector<char> Engine::take_flags() {
    PW_HIP(hipSetDevice(device_));
    std::vector<char> out(jobs_.size(), 0);
    int hi = 0;
    for (auto &j : jobs_) hi = std::max(hi, j->slot + 1);
    if (hi == 0) return out;
    int *h = flags_host_.as<int>();
    PW_HIP(hipMemcpyAsync(h, flags_.p, (size_t)hi * sizeof(int), hipMemcpyDeviceToHost, stream_));
    PW_HIP(hipStreamSynchronize(stream_));
    bool any = false;
    for (size_t k = 0; k < jobs_.size(); ++k) {
        out[k] = h[jobs_[k]->slot] != 0;
        any = any || out[k];
    }
    if (any) {
        PW_HIP(hipMemsetAsync(flags_.p, 0, (size_t)hi * sizeof(int), stream_));
        PW_HIP(hipStreamSynchronize(stream_));
    }
    return out;
}


// ------------------------------------------------ rates without allocation
// One wake's rates (a binary64 solve) for callers that cannot allocate device
// memory: every intermediate lives in `scratch`, a fixed array of this
// library in each device's memory, laid out for at most NOALLOC_MAX_N
// particles.  Calls on one device share it, so calls that overlap on a
// device race; each call rewrites every region it reads before reading it.
namespace {

constexpr size_t NOALLOC_MAX_N = 16384;
constexpr int NOALLOC_TB = 256;  // one-sided binary64 pass, 256 x 1, unroll 2
constexpr size_t NOALLOC_MAX_CHUNKS = 16;
constexpr size_t NOALLOC_MAX_ITEMS = (NOALLOC_MAX_N / NOALLOC_TB) * NOALLOC_MAX_CHUNKS;

constexpr size_t al256(size_t b) { return (b + 255) / 256 * 256; }
// byte offsets of the regions
constexpr size_t NA_JOB = 0;
constexpr size_t NA_ITEMS = al256(NA_JOB + sizeof(JobDev));
constexpr size_t NA_CITEMS = al256(NA_ITEMS + NOALLOC_MAX_ITEMS * sizeof(Item));
constexpr size_t NA_INTS = al256(NA_CITEMS + (NOALLOC_MAX_N / 256) * sizeof(CItem));
// ints: [0] pair items [1] (unused) [2] combine items [3] pad | [4] [5] queue
// counters | [8] non-finite flag
constexpr size_t NA_REC = al256(NA_INTS + 16 * sizeof(int));
constexpr size_t NA_XS = al256(NA_REC + NOALLOC_MAX_N * sizeof(Rec64));
constexpr size_t NA_VOL = al256(NA_XS + 6 * NOALLOC_MAX_N * sizeof(double));
constexpr size_t NA_RATE = al256(NA_VOL + NOALLOC_MAX_N * sizeof(double));
constexpr size_t NA_PART = al256(NA_RATE + 6 * NOALLOC_MAX_N * sizeof(double));
constexpr size_t NOALLOC_SCRATCH_BYTES =
    al256(NA_PART + NOALLOC_MAX_CHUNKS * NCOMP * NOALLOC_MAX_N * sizeof(double));
static_assert(NOALLOC_SCRATCH_BYTES < (32u << 20), "scratch stays under 32 MiB");

}  // namespace

size_t noalloc_max_n() { return NOALLOC_MAX_N; }
size_t noalloc_scratch_bytes() { return NOALLOC_SCRATCH_BYTES; }

// The scratch array.  In host code its name is a stand-in, not a device
// address: host code uses only the per-device address from
// hipGetSymbolAddress below.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Work tables, job descriptor, zeroed counters and flag, and the records of
// the caller's particles (N x 3 positions / vorticities, N radii / volumes,
// device arrays) written into the scratch regions: the plan build and record
// fill of the context path, on the device, so that the call needs no copy
// from host memory.
__global__ __launch_bounds__(256) void noalloc_prepare_kernel(
    JobDev J, int ntile, int nchunk, Item *__restrict__ items, CItem *__restrict__ citems,
    int *__restrict__ ints, JobDev *__restrict__ jd, const double *__restrict__ pos,
    const double *__restrict__ vort, const double *__restrict__ radius,
    const double *__restrict__ volume) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    const int nitems = ntile * nchunk, nci = (J.n + 255) / 256;
    if (k == 0) {
        *jd = J;
        ints[0] = nitems;
        ints[1] = 0;
        ints[2] = nci;
        ints[3] = 0;
        ints[4] = 0;  // the pair kernel's queue
        ints[5] = 0;
        ints[8] = 0;  // non-finite flag
    }
    if (k < nitems) items[k] = Item{0, k % ntile, k / ntile, J.chunk_len};
    if (k < nci) citems[k] = CItem{0, k};
    if (k < J.n) {
        const size_t np = (size_t)J.npad;
        Rec64 r;
        r.x = pos[3 * (size_t)k];
        r.y = pos[3 * (size_t)k + 1];
        r.z = pos[3 * (size_t)k + 2];
        r.h = 0.5 * radius[k] * radius[k];
        r.ax = vort[3 * (size_t)k];
        r.ay = vort[3 * (size_t)k + 1];
        r.az = vort[3 * (size_t)k + 2];
        r.v = volume[k];
        ((Rec64 *)J.rec)[k] = r;
        J.xs[3 * np + k] = r.ax;
        J.xs[4 * np + k] = r.ay;
        J.xs[5 * np + k] = r.az;
        ((double *)J.vol)[k] = r.v;
    }
}

namespace {
struct NoallocDevice {
    bool ready = false;
    unsigned char *ptr = nullptr;  // device address of scratch on this device
    int grid = 0;                  // persistent grid of the pair kernel
};
std::mutex g_noalloc_mu;
NoallocDevice g_noalloc[64];
}  // namespace

void rates_noalloc(int kind, double nu, const double uinf[3], size_t n, const double *pos,
                   const double *vort, const double *radius, const double *volume, double *vel,
                   double *dvort, hipStream_t stream) {
    if (kind != KIND_HIGHORDER && kind != KIND_GAUSSIAN) throw Error(PW_EINVAL, "unknown kernel");
    if (n == 0 || !pos || !vort || !radius || !volume || !vel || !dvort)
        throw Error(PW_EINVAL, "empty system or null array");
    if (n > NOALLOC_MAX_N)
        throw Error(PW_EINVAL, "pw_rates_noalloc takes at most " + std::to_string(NOALLOC_MAX_N) +
                                   " particles (its fixed device scratch is laid out for that)");
    const double u[3] = {uinf ? uinf[0] : 0.0, uinf ? uinf[1] : 0.0, uinf ? uinf[2] : 0.0};
    if (!std::isfinite(nu) || !std::isfinite(u[0]) || !std::isfinite(u[1]) || !std::isfinite(u[2]))
        throw Error(PW_EINVAL, "non-finite parameter");
    PairFn fn = kind == KIND_HIGHORDER ? pair64_kernel<KIND_HIGHORDER, NOALLOC_TB, 1, 2>
                                       : pair64_kernel<KIND_GAUSSIAN, NOALLOC_TB, 1, 2>;
    int dev = 0;
    PW_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw Error(PW_EINVAL, "device ordinal beyond 63");
    NoallocDevice *D = &g_noalloc[dev];
    {
        // set-up once per device, with that device current: the array's
        // address on it, and the pair kernel's resident grid
        std::lock_guard<std::mutex> g(g_noalloc_mu);
        if (!D->ready) {
            void *p = nullptr;
            PW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            int cus = 0, b64 = 0, b64g = 0;
            PW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            PW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &b64, (const void *)pair64_kernel<KIND_HIGHORDER, NOALLOC_TB, 1, 2>, NOALLOC_TB, 0));
            PW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &b64g, (const void *)pair64_kernel<KIND_GAUSSIAN, NOALLOC_TB, 1, 2>, NOALLOC_TB, 0));
            D->ptr = (unsigned char *)p;
            D->grid = std::max(1, std::min(b64, b64g)) * cus;
            D->ready = true;
        }
    }
    unsigned char *base = D->ptr;
    // the job as the context path would describe it, with every array in the
    // scratch regions; cores are taken as non-uniform (the radii are on the
    // device), so sigma is formed per pair
    JobDev J;
    std::memset(&J, 0, sizeof J);
    // at most NOALLOC_MAX_CHUNKS chunks: the scratch's partial sums and tables
    // are laid out for them
    const size_t npad = round_up(n, 256),
                 L = chunk_len(n, NOALLOC_TB, 1, false, (int)NOALLOC_MAX_CHUNKS);
    J.rec = base + NA_REC;
    J.xs = (double *)(base + NA_XS);
    J.x0 = J.xs;
    J.rate = (double *)(base + NA_RATE);
    J.part = (double *)(base + NA_PART);
    J.vol = (const double *)(base + NA_VOL);
    J.flag = (int *)(base + NA_INTS) + 8;
    J.n = (int)n;
    J.npad = (int)npad;
    J.chunk_len = (int)L;
    J.nchunk = (int)((n + L - 1) / L);
    J.prec = PREC_FP64;
    J.uniform = 0;
    J.inv_s2 = 0.0;
    J.rec_scale = 1.0;
    J.nu = nu;
    for (int c = 0; c < 3; ++c) J.uinf[c] = u[c];
    const PairScale ps = pair_scale(kind, PREC_FP64);
    J.su = J.sq = (double)(1.0L / (long double)ps.q);
    J.sd = (double)(2.0L * nu / (long double)ps.z);
    const int ntile = (int)((n + NOALLOC_TB - 1) / NOALLOC_TB);
    if ((size_t)J.nchunk > NOALLOC_MAX_CHUNKS || (size_t)ntile * J.nchunk > NOALLOC_MAX_ITEMS)
        throw Error(PW_EINVAL, "work plan exceeds the scratch tables");
    JobDev *jd = (JobDev *)(base + NA_JOB);
    Item *items = (Item *)(base + NA_ITEMS);
    CItem *citems = (CItem *)(base + NA_CITEMS);
    int *ints = (int *)(base + NA_INTS);
    const int nprep = (int)std::max<size_t>(n, (size_t)ntile * J.nchunk);
    hipLaunchKernelGGL(noalloc_prepare_kernel, dim3((nprep + 255) / 256), dim3(256), 0, stream, J,
                       ntile, J.nchunk, items, citems, ints, jd, pos, vort, radius, volume);
    PW_HIP(hipGetLastError());
    hipLaunchKernelGGL(fn, dim3(D->grid), dim3(NOALLOC_TB), 0, stream, (const JobDev *)jd,
                       (const Item *)items, (const int *)ints, ints + 4);
    PW_HIP(hipGetLastError());
    hipLaunchKernelGGL(combine_kernel, dim3(4 * std::max(1, ntile)), dim3(256), 0, stream,
                       (const JobDev *)jd, (const CItem *)citems, (const int *)(ints + 2), ints + 4,
                       (int)MODE_SOLVE);
    PW_HIP(hipGetLastError());
    hipLaunchKernelGGL(aos_soa_kernel, dim3(grid_for(n)), dim3(256), 0, stream, vel, J.rate, (int)n,
                       (int)npad, 1);
    hipLaunchKernelGGL(aos_soa_kernel, dim3(grid_for(n)), dim3(256), 0, stream, dvort,
                       J.rate + 3 * npad, (int)n, (int)npad, 1);
    PW_HIP(hipGetLastError());
}

}  // namespace pw
