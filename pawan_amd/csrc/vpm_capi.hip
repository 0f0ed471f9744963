// vpm_capi.hip -- the extern "C" surface of include/pawan_hip.h over the engine.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <exception>
#include <map>
#include <string>
#include <vector>

#include <cstdlib>
#include <fstream>
#include <sstream>

#include "../../include/pawan_hip.h"
#include "vpm_engine.hpp"
#include "vpm_writer.hpp"

using pw::Engine;
using pw::Error;
using pw::Job;
using pw::JobParams;

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(std::string *err, F &&f) {
    try {
        f();
        return PW_OK;
    } catch (const Error &e) {
        if (err) *err = e.what();
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        if (err) *err = "host allocation failed";
        g_last_error = "host allocation failed";
        return PW_ENOMEM;
    } catch (const std::exception &e) {
        if (err) *err = e.what();
        g_last_error = e.what();
        return PW_EINVAL;
    }
}

// One target per thread is the faster binary32 symmetric pass (59.9 against
// 68.2 ms at N = 262,144; 244 against 270 ms at 524,288), but its slots take
// 80 bytes per particle per strip, and past 128 strips the strip count is
// N / (8 x block x targets per thread): past N = 524,288 (43 GB at 256 x 1 and
// N = 1,048,576) a binary32 symmetric HIGHORDER job takes 512 threads x 1
// target, which halves them as 256 x 2 does with the faster one-target pass
// (973 against 1,052 ms per RHS at N = 1,048,576).  From 131,072 to 524,288
// 512 x 1 and 256 x 1 are within 1% of each other (the autotuner picks).
// An explicit two targets per thread, or a chunk length that 512 does not
// divide, keeps 256 threads there (the one 512-thread variant takes one
// target per thread; chunks are multiples of the block).
int default_block(size_t n, int precision, int kernel, int scheme, int tpt = 0, int chunk = 0) {
    if (precision == PW_FP32C && kernel == PW_KERNEL_HIGHORDER && scheme != PW_SCHEME_ONESIDED &&
        n > 524288 && tpt != 2 && chunk % 512 == 0)
        return 512;
    if (n >= 32768) return 256;
    if (n >= 8192) return 128;
    return 64;
}

// Two targets per thread for the binary32 one-sided HIGHORDER pass past
// N = 524,288 (the symmetric one takes 512 x 1 there).  (The GAUSSIAN kernel is
// compiled with one target per thread only.)
int default_tpt(size_t n, int precision, int kernel) {
    return kernel == PW_KERNEL_HIGHORDER && precision == PW_FP32C && n > 524288 ? 2 : 1;
}

bool finite_params(double nu, double dt, const double *u) {
    return std::isfinite(nu) && std::isfinite(dt) && std::isfinite(u[0]) && std::isfinite(u[1]) &&
           std::isfinite(u[2]);
}

}  // namespace

// ---------------------------------------------------------------- context
struct pw_ctx {
    pw_config cfg;
    std::string err;
    std::unique_ptr<Engine> eng;
    Job *job = nullptr;
    bool dirty = false;
    // wakes as added (host copies; refreshed from the device before a rebuild)
    std::vector<size_t> wake_n;
    std::vector<int64_t> woff{0};
    std::vector<double> pos, vort, rad, vol;
    double dt = 0.0;
    // asynchronous frames (pw_write_async); declared after eng so that it is
    // destroyed first, once its frames are written
    std::unique_ptr<pw::FrameWriter> writer;

    size_t total() const { return (size_t)woff.back(); }

    // the variant for a context of n particles (default: the current total);
    // *chunks receives the source chunk count its job takes (0: the policy)
    pw::Variant choose_variant(int *chunks = nullptr) const { return choose_variant(total(), chunks); }
    pw::Variant choose_variant(size_t n, int *chunks = nullptr) const {
        if (chunks) *chunks = cfg.chunks;
        pw::Variant v;
        v.block = cfg.block ? cfg.block
                            : default_block(n, cfg.precision, cfg.kernel, cfg.scheme,
                                            cfg.targets_per_thread, cfg.chunk);
        v.tpt = cfg.targets_per_thread ? cfg.targets_per_thread : default_tpt(n, cfg.precision, cfg.kernel);
        if (v.block == 512) v.tpt = 1;  // the one 512-thread variant
        if (cfg.unroll) v.unroll64 = v.unroll32 = cfg.unroll;
        v.sym32 = cfg.scheme != PW_SCHEME_ONESIDED;
        v.sym64 = cfg.scheme == PW_SCHEME_SYMMETRIC;
        const char *cache = std::getenv("PW_TUNE_CACHE");
        if (cache && !cfg.block && !cfg.targets_per_thread && !cfg.unroll && !cfg.scheme) {
            pw_variant t;
            if (pw_tune_lookup(cache, cfg.kernel, cfg.precision, n, &t) == PW_OK) {
                v.block = t.block;
                v.tpt = t.targets_per_thread;
                if (chunks && !cfg.chunks) *chunks = t.chunks;
                if (cfg.precision == PW_FP64) {
                    if (t.unroll) v.unroll64 = t.unroll;
                    v.sym64 = t.scheme == PW_SCHEME_SYMMETRIC;
                } else {
                    v.sym32 = t.scheme != PW_SCHEME_ONESIDED;
                    if (t.unroll) v.unroll32 = t.unroll;
                }
            }
        }
        return v;
    }

    // the engine runs variant v already
    bool engine_runs(const pw::Variant &v) const {
        if (!eng) return false;
        const pw::Variant cur = eng->variant();
        const int cur32 = cur.tpt32 ? cur.tpt32 : cur.tpt, v32 = v.tpt32 ? v.tpt32 : v.tpt;
        return cur.block == v.block && cur.tpt == v.tpt && cur32 == v32 &&
               cur.unroll64 == v.unroll64 && cur.unroll32 == v.unroll32 && cur.sym32 == v.sym32 &&
               cur.sym64 == v.sym64;
    }

    void rebuild() {
        if (total() == 0) throw Error(PW_EINVAL, "no particles: add a wake first");
        int chunks = 0;
        pw::Variant v = choose_variant(&chunks);
        if (writer) writer->quiesce();  // queued frames read the current device state
        if (!engine_runs(v)) {
            job = nullptr;
            eng.reset();
            eng.reset(new Engine(cfg.device, cfg.kernel, v));
        }
        if (job) {
            eng->remove_job(job->id);
            job = nullptr;
        }
        JobParams par;
        par.nu = cfg.nu;
        par.dt = dt;
        std::memcpy(par.uinf, cfg.uinf, sizeof par.uinf);
        par.precision = cfg.precision;
        par.chunks = chunks;
        job = eng->add_job(1, total(), pos.data(), vort.data(), rad.data(), vol.data(), par,
                           (size_t)std::max(0, cfg.chunk));
        dirty = false;
    }
    void ready() {
        if (dirty || !job) rebuild();
    }
    void set_dt(double d) {
        if (d != dt) {
            dt = d;
            JobParams par = job->par;
            par.dt = d;
            eng->set_params(job, par);
        }
    }
    int flag_status() { return eng->take_flag(job) ? PW_ENONFINITE : PW_OK; }
};

extern "C" {

const char *pw_version(void) { return "pawan_hip 1.0 gfx950"; }

int pw_config_default(pw_config *cfg) {
    if (!cfg) return PW_EINVAL;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->device = 0;
    cfg->kernel = PW_KERNEL_HIGHORDER;
    cfg->precision = PW_FP64;
    cfg->nu = 2.5e-3;
    return PW_OK;
}

int pw_create(pw_ctx **out, const pw_config *cfg) {
    if (!out) return PW_EINVAL;
    *out = nullptr;
    return guarded(nullptr, [&] {
        pw_config c;
        if (cfg)
            c = *cfg;
        else
            pw_config_default(&c);
        if (c.kernel != PW_KERNEL_HIGHORDER && c.kernel != PW_KERNEL_GAUSSIAN)
            throw Error(PW_EINVAL, "unknown kernel");
        if (c.precision != PW_FP64 && c.precision != PW_FP32C)
            throw Error(PW_EINVAL, "unknown precision");
        if (c.block != 0 && c.block != 64 && c.block != 128 && c.block != 256 && c.block != 512)
            throw Error(PW_EINVAL, "block must be 0, 64, 128, 256 or 512");
        if (c.block == 512 && (c.precision != PW_FP32C || c.kernel != PW_KERNEL_HIGHORDER ||
                               c.scheme == PW_SCHEME_ONESIDED || c.targets_per_thread == 2))
            throw Error(PW_EINVAL,
                        "block 512 is the binary32 symmetric HIGHORDER pass, one target per thread");
        if (c.chunk < 0 || (c.chunk && c.block && c.chunk % c.block))
            throw Error(PW_EINVAL, "chunk must be a non-negative multiple of block");
        if (c.chunks < 0) throw Error(PW_EINVAL, "chunks must be non-negative");
        if (c.targets_per_thread < 0 || c.targets_per_thread > 2)
            throw Error(PW_EINVAL, "targets_per_thread must be 0, 1 or 2");
        if (c.unroll != 0 && c.unroll != 1 && c.unroll != 2 && c.unroll != 4)
            throw Error(PW_EINVAL, "unroll must be 0, 1, 2 or 4");
        if (c.scheme < PW_SCHEME_DEFAULT || c.scheme > PW_SCHEME_SYMMETRIC)
            throw Error(PW_EINVAL, "unknown pair scheme");
        if (!finite_params(c.nu, 0.0, c.uinf)) throw Error(PW_EINVAL, "non-finite parameter");
        int ndev = 0;
        pw::hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        if (c.device < 0 || c.device >= ndev) throw Error(PW_EINVAL, "no such HIP device");
        pw_ctx *x = new pw_ctx();
        x->cfg = c;
        *out = x;
    });
}

int pw_destroy(pw_ctx *ctx) {
    if (!ctx) return PW_EINVAL;
    int rc = guarded(nullptr, [&] {
        ctx->writer.reset();
        ctx->eng.reset();
    });
    delete ctx;
    return rc;
}

const char *pw_last_error(const pw_ctx *ctx) {
    return ctx ? ctx->err.c_str() : g_last_error.c_str();
}

int pw_set_nu(pw_ctx *ctx, double nu) {
    if (!ctx) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        if (!std::isfinite(nu)) throw Error(PW_EINVAL, "non-finite viscosity");
        ctx->cfg.nu = nu;
        if (ctx->job) {
            JobParams par = ctx->job->par;
            par.nu = nu;
            ctx->eng->set_params(ctx->job, par);
        }
    });
}

int pw_add_wake(pw_ctx *ctx, size_t n, const double *pos, const double *vort,
                const double *radius, const double *volume) {
    if (!ctx) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        if (n == 0 || !pos || !vort || !radius || !volume)
            throw Error(PW_EINVAL, "empty wake or null array");
        // keep the device state of the wakes already present; once the host
        // arrays are ahead of the device (a wake added, particles inserted),
        // they hold that state already and the device arrays no longer match
        // their layout
        if (ctx->job && !ctx->dirty) {
            ctx->eng->download_aos(ctx->job, 0, ctx->pos.data());
            ctx->eng->download_aos(ctx->job, 1, ctx->vort.data());
        }
        ctx->pos.insert(ctx->pos.end(), pos, pos + 3 * n);
        ctx->vort.insert(ctx->vort.end(), vort, vort + 3 * n);
        ctx->rad.insert(ctx->rad.end(), radius, radius + n);
        ctx->vol.insert(ctx->vol.end(), volume, volume + n);
        ctx->wake_n.push_back(n);
        ctx->woff.push_back(ctx->woff.back() + (int64_t)n);
        ctx->dirty = true;
    });
}

size_t pw_size(const pw_ctx *ctx) { return ctx ? 6 * ctx->total() : 0; }
size_t pw_num_wakes(const pw_ctx *ctx) { return ctx ? ctx->wake_n.size() : 0; }

int pw_set_state(pw_ctx *ctx, const double *state) {
    if (!ctx || !state) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        ctx->ready();
        ctx->eng->upload_state_layout(ctx->job, state, ctx->woff);
    });
}

int pw_get_state(pw_ctx *ctx, double *state) {
    if (!ctx || !state) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        ctx->ready();
        ctx->eng->download_state_layout(ctx->job, 0, state, ctx->woff);
    });
}

int pw_solve(pw_ctx *ctx) {
    if (!ctx) return PW_EINVAL;
    int st = PW_OK;
    int rc = guarded(&ctx->err, [&] {
        ctx->ready();
        ctx->eng->solve();
        st = ctx->flag_status();
        if (st) ctx->err = "non-finite rates (coincident particles?)";
    });
    return rc ? rc : st;
}

int pw_get_rates(pw_ctx *ctx, double *rates) {
    if (!ctx || !rates) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        ctx->ready();
        ctx->eng->download_state_layout(ctx->job, 1, rates, ctx->woff);
    });
}

int pw_get_wake(pw_ctx *ctx, size_t wake, double *pos, double *vort, double *vel,
                double *dvort) {
    if (!ctx) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        if (wake >= ctx->wake_n.size()) throw Error(PW_EINVAL, "no such wake");
        ctx->ready();
        double *outs[4] = {pos, vort, vel, dvort};
        // only this wake's particles cross to the host
        for (int k = 0; k < 4; ++k)
            if (outs[k])
                ctx->eng->download_aos(ctx->job, k, outs[k], (size_t)ctx->woff[wake],
                                       ctx->wake_n[wake]);
    });
}

// Contexts up to this many pair interactions per RHS (about 32,768
// particles) replay each RK4 step as one captured graph: their kernels take
// microseconds, and per-launch costs set the step time (a 2,048-particle
// ring).  Larger ones launch eagerly, with the pair kernels timed one by one.
constexpr double kGraphStepPairs = 1073741824.0;

int pw_rk4_steps(pw_ctx *ctx, double dt, size_t nsteps) {
    if (!ctx) return PW_EINVAL;
    int st = PW_OK;
    int rc = guarded(&ctx->err, [&] {
        if (!std::isfinite(dt)) throw Error(PW_EINVAL, "non-finite dt");
        ctx->ready();
        ctx->set_dt(dt);
        const double n = (double)ctx->total();
        ctx->eng->rk4(nsteps, n * (n - 1.0) <= kGraphStepPairs);
        st = ctx->flag_status();
        if (st) ctx->err = "non-finite rates or state";
    });
    return rc ? rc : st;
}

int pw_euler_steps(pw_ctx *ctx, double dt, size_t nsteps) {
    if (!ctx) return PW_EINVAL;
    int st = PW_OK;
    int rc = guarded(&ctx->err, [&] {
        if (!std::isfinite(dt)) throw Error(PW_EINVAL, "non-finite dt");
        ctx->ready();
        ctx->set_dt(dt);
        ctx->eng->euler(nsteps);
        st = ctx->flag_status();
        if (st) ctx->err = "non-finite rates or state";
    });
    return rc ? rc : st;
}

int pw_diagnostics(pw_ctx *ctx, double omega[3], double impulse[3], double *ke) {
    if (!ctx) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        ctx->ready();
        pw::Diagnostics d;
        ctx->eng->diagnostics(ctx->job, &d, ke != nullptr);
        if (omega) std::memcpy(omega, d.omega, sizeof d.omega);
        if (impulse) std::memcpy(impulse, d.impulse, sizeof d.impulse);
        if (ke) *ke = d.ke;
    });
}

int pw_diagnose(pw_ctx *ctx, pw_diag *out) {
    if (!ctx || !out) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        ctx->ready();
        pw::Diagnostics d;
        ctx->eng->diagnostics(ctx->job, &d, true);
        std::memcpy(out->total_vorticity, d.omega, sizeof d.omega);
        std::memcpy(out->linear_impulse, d.impulse, sizeof d.impulse);
        std::memcpy(out->angular_impulse, d.angular, sizeof d.angular);
        out->helicity = d.helicity;
        out->enstrophy = d.enstrophy;
        out->kinetic_energy = d.ke;
    });
}

int pw_resolve(pw_ctx *ctx, double *field) {
    if (!ctx || !field) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        ctx->ready();
        ctx->eng->influence(ctx->job, field);
    });
}

int pw_write(pw_ctx *ctx, FILE *f) {
    if (!ctx || !f) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        ctx->ready();
        const size_t n = ctx->total();
        std::vector<double> pos(3 * n), vort(3 * n);
        ctx->eng->download_aos(ctx->job, 0, pos.data());
        ctx->eng->download_aos(ctx->job, 1, vort.data());
        size_t nw = ctx->wake_n.size();
        bool ok = std::fwrite(&nw, sizeof(size_t), 1, f) == 1;
        for (size_t w = 0; w < nw && ok; ++w) {
            const size_t m = ctx->wake_n[w], o = (size_t)ctx->woff[w];
            ok = std::fwrite(&m, sizeof(size_t), 1, f) == 1 &&
                 std::fwrite(pos.data() + 3 * o, sizeof(double), 3 * m, f) == 3 * m &&
                 std::fwrite(vort.data() + 3 * o, sizeof(double), 3 * m, f) == 3 * m &&
                 std::fwrite(ctx->rad.data() + o, sizeof(double), m, f) == m;
        }
        if (!ok) throw Error(PW_EINVAL, "write to the file failed");
    });
}

int pw_write_async(pw_ctx *ctx, FILE *f, double t, int with_time) {
    if (!ctx || !f) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        ctx->ready();
        if (!ctx->writer) ctx->writer.reset(new pw::FrameWriter(ctx->cfg.device));
        std::string m;
        if (ctx->writer->status(&m) != PW_OK) {  // an earlier frame failed: report it
            const int code = ctx->writer->drain(&m);
            throw Error(code, m);
        }
        ctx->writer->submit(ctx->eng.get(), ctx->job, f, t, with_time != 0, ctx->wake_n,
                            ctx->rad);
    });
}

int pw_write_wait(pw_ctx *ctx) {
    if (!ctx) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        if (!ctx->writer) return;
        std::string m;
        const int code = ctx->writer->drain(&m);
        if (code != PW_OK) throw Error(code, m);
    });
}

int pw_append_particles(pw_ctx *ctx, size_t wake, size_t n, const double *pos, const double *vort,
                        const double *radius, const double *volume) {
    if (!ctx) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        if (wake >= ctx->wake_n.size()) throw Error(PW_EINVAL, "no such wake");
        if (n == 0) return;
        if (!pos || !vort || !radius || !volume) throw Error(PW_EINVAL, "null array");
        const size_t last = ctx->wake_n.size() - 1;
        // in place only while the grown context keeps its kernel variant and
        // chunk count (a binary32 context past 524,288 particles moves to 512
        // threads, which halves its slots); otherwise the next call rebuilds it
        int chunks = 0;
        const pw::Variant grown = ctx->choose_variant(ctx->total() + n, &chunks);
        if (ctx->job && !ctx->dirty && wake == last && ctx->engine_runs(grown) &&
            chunks == ctx->job->par.chunks) {
            // the wake's particles end the device arrays: append in place
            ctx->eng->append(ctx->job, n, pos, vort, radius, volume);
            ctx->pos.insert(ctx->pos.end(), pos, pos + 3 * n);
            ctx->vort.insert(ctx->vort.end(), vort, vort + 3 * n);
        } else {
            if (ctx->job && !ctx->dirty) {  // keep the device state, then insert
                ctx->eng->download_aos(ctx->job, 0, ctx->pos.data());
                ctx->eng->download_aos(ctx->job, 1, ctx->vort.data());
            }
            const size_t at = (size_t)ctx->woff[wake + 1];
            ctx->pos.insert(ctx->pos.begin() + 3 * at, pos, pos + 3 * n);
            ctx->vort.insert(ctx->vort.begin() + 3 * at, vort, vort + 3 * n);
            ctx->dirty = true;
        }
        const size_t at = (size_t)ctx->woff[wake + 1];
        ctx->rad.insert(ctx->rad.begin() + at, radius, radius + n);
        ctx->vol.insert(ctx->vol.begin() + at, volume, volume + n);
        ctx->wake_n[wake] += n;
        for (size_t w = wake + 1; w < ctx->woff.size(); ++w) ctx->woff[w] += (int64_t)n;
    });
}

int pw_get_variant(pw_ctx *ctx, pw_variant *out) {
    if (!ctx || !out) return PW_EINVAL;
    return guarded(&ctx->err, [&] {
        ctx->ready();
        const pw::Variant v = ctx->eng->variant();
        const bool fp32 = ctx->cfg.precision == PW_FP32C;
        out->block = pw::block_for(v.block, ctx->cfg.precision);
        out->targets_per_thread = fp32 && v.tpt32 ? v.tpt32 : v.tpt;
        const bool sym = fp32 ? v.sym32 != 0 : v.sym64 != 0;
        out->unroll = sym ? 0 : (fp32 ? v.unroll32 : v.unroll64);
        out->scheme = sym ? PW_SCHEME_SYMMETRIC : PW_SCHEME_ONESIDED;
        out->chunks = ctx->job->nchunk;
    });
}

int pw_rates_noalloc(int kernel, double nu, const double *uinf, size_t n, const double *pos,
                     const double *vort, const double *radius, const double *volume, double *vel,
                     double *dvort, void *stream) {
    static_assert(PW_NOALLOC_MAX_N == 16384, "header limit matches the scratch layout");
    return guarded(nullptr, [&] {
        if (pw::noalloc_max_n() != PW_NOALLOC_MAX_N) throw Error(PW_EINVAL, "scratch layout mismatch");
        pw::rates_noalloc(kernel, nu, uinf, n, pos, vort, radius, volume, vel, dvort,
                          (hipStream_t)stream);
    });
}

size_t pw_noalloc_scratch_bytes(void) { return pw::noalloc_scratch_bytes(); }

int pw_last_timing(const pw_ctx *ctx, double *total_ms, double *pair_ms, double *pairs) {
    if (!ctx || !ctx->eng) return PW_EINVAL;
    if (total_ms) *total_ms = ctx->eng->last_total_ms;
    if (pair_ms) *pair_ms = ctx->eng->last_pair_ms;
    if (pairs) *pairs = ctx->eng->last_pairs;
    return PW_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batch
struct pw_batch {
    std::string err;
    std::unique_ptr<Engine> eng;
    int chunks = 0;  // source chunks per job (pw_variant::chunks)
    int64_t next_id = 1;
    std::map<int64_t, int> status;
};

static JobParams params_of(const pw_job_spec *s) {
    if (!finite_params(s->nu, s->dt, s->uinf)) throw Error(PW_EINVAL, "non-finite job parameter");
    if (s->precision != PW_FP64 && s->precision != PW_FP32C)
        throw Error(PW_EINVAL, "unknown precision");
    JobParams p;
    p.nu = s->nu;
    p.dt = s->dt;
    std::memcpy(p.uinf, s->uinf, sizeof p.uinf);
    p.precision = s->precision;
    return p;
}

extern "C" {

// Batches and servers: 256 threads; binary64 jobs two targets per thread
// (the fastest one-sided HIGHORDER variant), binary32 jobs one (the fastest
// symmetric pass); GAUSSIAN is compiled with one target per thread only.  An
// explicit targets_per_thread applies to both.
static pw::Variant variant_of(const pw_variant *v, int precision_hint, int kernel) {
    pw::Variant r;
    r.block = 256;
    r.tpt = kernel == PW_KERNEL_HIGHORDER ? 2 : 1;
    r.tpt32 = 1;
    if (!v) return r;
    if (v->scheme < PW_SCHEME_DEFAULT || v->scheme > PW_SCHEME_SYMMETRIC)
        throw Error(PW_EINVAL, "unknown pair scheme");
    if (v->chunks < 0) throw Error(PW_EINVAL, "chunks must be non-negative");
    r.block = v->block ? v->block : 256;
    if (v->targets_per_thread) r.tpt = r.tpt32 = v->targets_per_thread;
    r.sym32 = v->scheme != PW_SCHEME_ONESIDED;
    r.sym64 = v->scheme == PW_SCHEME_SYMMETRIC;
    if (v->unroll) {
        if (precision_hint == PW_FP32C)
            r.unroll32 = v->unroll;
        else if (precision_hint == PW_FP64)
            r.unroll64 = v->unroll;
        else
            r.unroll64 = r.unroll32 = v->unroll;
    }
    return r;
}

int pw_batch_create(pw_batch **out, int device, int kernel, const pw_variant *variant) {
    if (!out) return PW_EINVAL;
    *out = nullptr;
    return guarded(nullptr, [&] {
        auto b = std::make_unique<pw_batch>();
        b->eng.reset(new Engine(device, kernel, variant_of(variant, -1, kernel)));
        b->chunks = variant ? variant->chunks : 0;
        *out = b.release();
    });
}

int pw_batch_destroy(pw_batch *b) {
    if (!b) return PW_EINVAL;
    int rc = guarded(nullptr, [&] { b->eng.reset(); });
    delete b;
    return rc;
}

const char *pw_batch_last_error(const pw_batch *b) { return b ? b->err.c_str() : g_last_error.c_str(); }

int pw_batch_add(pw_batch *b, const pw_job_spec *spec, int64_t *id) {
    if (!b || !spec || !id) return PW_EINVAL;
    return guarded(&b->err, [&] {
        JobParams p = params_of(spec);
        p.chunks = b->chunks;
        int64_t jid = b->next_id++;
        b->eng->add_job(jid, spec->n, spec->pos, spec->vort, spec->radius, spec->volume, p);
        b->status[jid] = PW_OK;
        *id = jid;
    });
}

int pw_batch_remove(pw_batch *b, int64_t id) {
    if (!b) return PW_EINVAL;
    return guarded(&b->err, [&] {
        b->eng->remove_job(id);
        b->status.erase(id);
    });
}

static void batch_flags(pw_batch *b) {
    const std::vector<char> f = b->eng->take_flags();  // every job's, one copy
    const auto &jobs = b->eng->jobs();
    for (size_t k = 0; k < jobs.size(); ++k)
        if (f[k]) b->status[jobs[k]->id] = PW_ENONFINITE;
}

int pw_batch_step(pw_batch *b, size_t nsteps, int use_graph) {
    if (!b) return PW_EINVAL;
    return guarded(&b->err, [&] {
        b->eng->rk4(nsteps, use_graph != 0);
        batch_flags(b);
    });
}

int pw_batch_solve(pw_batch *b) {
    if (!b) return PW_EINVAL;
    return guarded(&b->err, [&] {
        b->eng->solve();
        batch_flags(b);
    });
}

int pw_batch_get(pw_batch *b, int64_t id, double *pos, double *vort, double *vel, double *dvort) {
    if (!b) return PW_EINVAL;
    return guarded(&b->err, [&] {
        Job *j = b->eng->find(id);
        if (!j) throw Error(PW_ENOTFOUND, "unknown job");
        double *outs[4] = {pos, vort, vel, dvort};
        for (int k = 0; k < 4; ++k)
            if (outs[k]) b->eng->download_aos(j, k, outs[k]);
    });
}

int pw_batch_status(pw_batch *b, int64_t id) {
    if (!b) return PW_EINVAL;
    auto it = b->status.find(id);
    return it == b->status.end() ? PW_ENOTFOUND : it->second;
}

int pw_batch_last_timing(const pw_batch *b, double *total_ms, double *pair_ms, double *pairs) {
    if (!b) return PW_EINVAL;
    if (total_ms) *total_ms = b->eng->last_total_ms;
    if (pair_ms) *pair_ms = b->eng->last_pair_ms;
    if (pairs) *pairs = b->eng->last_pairs;
    return PW_OK;
}

size_t pw_chunk_for(size_t n, int block) { return pw::chunk_for(n, block ? block : 128); }

}  // extern "C"

// ---------------------------------------------------------------- serving
namespace {
enum JobState { JS_QUEUED = 0, JS_RUNNING = 1, JS_DONE = 2, JS_FAILED = 3, JS_CANCELLED = 4 };

struct SJob {
    int64_t id = 0;
    JobState state = JS_QUEUED;
    size_t n = 0, nsteps = 0, done = 0, bytes = 0;
    JobParams par;
    double arrival = 0.0, admitted = 0.0, finished = 0.0;
    std::vector<double> pos, vort, rad, vol;  // inputs until admitted, results when done
    // while resident: the device job, the next RK4 stage (0-3) and the stages left;
    // a stage larger than a pass's work runs in `parts` slices, `part` the next
    Job *job = nullptr;
    int stage = 0;
    size_t stages_left = 0;
    int part = 0, parts = 1;
};

// Stages a resident job runs in one pass: 4 if its stage costs at most a
// quarter of the pass's reference cost (the largest resident stage, clamped
// to [kPassFloor, kPassCap] pair interactions), 2 if at most half, else 1;
// never past the end of its current step or of its run.  A stage that costs
// more than the reference runs in slices of about the reference's work, one
// per pass.  A binary32 stage is priced at half a binary64 one (its pair pass
// runs at about twice the rate).  kPassCap, ~28 ms of binary64 pairs, bounds
// a pass: a 262,144-particle binary64 stage (6.9e10) runs in 5 slices.
constexpr double kPassFloor = 2e9;
constexpr double kPassCap = 1.6e10;
double stage_cost(const SJob &j) {
    return (double)j.n * (double)(j.n - 1) * (j.par.precision == PW_FP32C ? 0.5 : 1.0);
}
int stages_this_pass(const SJob &j, double ref) {
    const double c = stage_cost(j);
    const int k = 4 * c <= ref ? 4 : (2 * c <= ref ? 2 : 1);
    return (int)std::min<size_t>({(size_t)k, (size_t)(4 - j.stage), j.stages_left});
}

double pct(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    double r = q * (double)(v.size() - 1);
    size_t lo = (size_t)std::floor(r), hi = (size_t)std::ceil(r);
    return v[lo] + (v[hi] - v[lo]) * (r - (double)lo);
}
}  // namespace

struct pw_server {
    std::string err;
    std::unique_ptr<Engine> eng;
    int chunks = 0;  // source chunks per job (pw_variant::chunks)
    size_t budget = 0, in_use = 0;
    int64_t next_id = 1;
    std::map<int64_t, SJob> jobs;
    std::deque<int64_t> queue;     // arrival order
    std::vector<int64_t> resident;  // admission order
    int64_t cursor = 0;             // large jobs: the next pass starts at this id
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::vector<double> step_ms, job_ms, job_step_ms;
    pw_server_stats st{};

    // queued jobs in arrival order that fit the budget, uploaded (beside a
    // running pass when async) and appended to `into`
    void admit(bool async, std::vector<int64_t> *into) {
        for (auto it = queue.begin(); it != queue.end();) {
            SJob &j = jobs[*it];
            if (in_use + j.bytes > budget) {
                ++it;
                continue;
            }
            j.job = eng->add_job(j.id, j.n, j.pos.data(), j.vort.data(), j.rad.data(),
                                 j.vol.data(), j.par, 0, async);
            in_use += j.bytes;
            j.state = JS_RUNNING;
            j.admitted = now();
            j.stage = 0;
            j.stages_left = 4 * j.nsteps;
            j.pos.clear();
            j.vort.clear();
            j.pos.shrink_to_fit();
            j.vort.shrink_to_fit();
            st.admitted++;
            into->push_back(j.id);
            it = queue.erase(it);
        }
    }

    double now() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

extern "C" {

size_t pw_job_footprint(size_t n, int precision) {
    const int b = default_block(n, precision, PW_KERNEL_HIGHORDER, PW_SCHEME_DEFAULT);
    return pw::job_footprint(n, precision, b,
                             b == 512 ? 1 : default_tpt(n, precision, PW_KERNEL_HIGHORDER),
                             precision == PW_FP32C);
}

int pw_server_create(pw_server **out, int device, int kernel, size_t hbm_budget_bytes,
                     const pw_variant *variant) {
    if (!out) return PW_EINVAL;
    *out = nullptr;
    return guarded(nullptr, [&] {
        auto s = std::make_unique<pw_server>();
        s->eng.reset(new Engine(device, kernel, variant_of(variant, -1, kernel)));
        s->chunks = variant ? variant->chunks : 0;
        s->budget = hbm_budget_bytes;
        *out = s.release();
    });
}

int pw_server_destroy(pw_server *s) {
    if (!s) return PW_EINVAL;
    int rc = guarded(nullptr, [&] { s->eng.reset(); });
    delete s;
    return rc;
}

const char *pw_server_last_error(const pw_server *s) { return s ? s->err.c_str() : g_last_error.c_str(); }

double pw_server_now_ms(const pw_server *s) { return s ? s->now() : 0.0; }

int pw_server_submit(pw_server *s, const pw_job_spec *spec, size_t nsteps, double arrival_ms,
                     int64_t *id) {
    if (!s || !spec || !id) return PW_EINVAL;
    return guarded(&s->err, [&] {
        if (spec->n == 0 || !spec->pos || !spec->vort || !spec->radius || !spec->volume)
            throw Error(PW_EINVAL, "empty job or null array");
        if (nsteps == 0) throw Error(PW_EINVAL, "a job needs at least one step");
        SJob j;
        j.par = params_of(spec);
        j.par.chunks = s->chunks;
        j.id = s->next_id++;
        j.n = spec->n;
        j.nsteps = nsteps;
        const pw::Variant ev = s->eng->variant();
        const int tpt = spec->precision == PW_FP32C && ev.tpt32 ? ev.tpt32 : ev.tpt;
        j.bytes = pw::job_footprint(spec->n, spec->precision, ev.block, tpt,
                                    spec->precision == PW_FP32C ? ev.sym32 != 0 : ev.sym64 != 0,
                                    s->chunks);
        if (j.bytes > s->budget) throw Error(PW_ENOMEM, "job larger than the HBM budget");
        j.arrival = arrival_ms < 0 ? s->now() : arrival_ms;
        j.pos.assign(spec->pos, spec->pos + 3 * spec->n);
        j.vort.assign(spec->vort, spec->vort + 3 * spec->n);
        j.rad.assign(spec->radius, spec->radius + spec->n);
        j.vol.assign(spec->volume, spec->volume + spec->n);
        *id = j.id;
        s->queue.push_back(j.id);
        s->jobs.emplace(j.id, std::move(j));
        s->st.submitted++;
    });
}

int pw_server_cancel(pw_server *s, int64_t id) {
    if (!s) return PW_EINVAL;
    return guarded(&s->err, [&] {
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) throw Error(PW_ENOTFOUND, "unknown job");
        SJob &j = it->second;
        if (j.state == JS_RUNNING) {
            s->eng->remove_job(id);
            s->in_use -= j.bytes;
            j.job = nullptr;
            s->resident.erase(std::find(s->resident.begin(), s->resident.end(), id));
        } else if (j.state == JS_QUEUED) {
            s->queue.erase(std::find(s->queue.begin(), s->queue.end(), id));
        } else {
            return;  // already finished
        }
        j.state = JS_CANCELLED;
        j.pos.clear();
        j.vort.clear();
        s->st.cancelled++;
    });
}

int pw_server_step(pw_server *s, int64_t *running) {
    if (!s) return PW_EINVAL;
    return guarded(&s->err, [&] {
        const double t_begin = s->now();
        // an idle server admits first, so that arrivals start in this pass
        if (s->resident.empty()) s->admit(false, &s->resident);
        // the pass: every small job's stages (a whole step at most), and the
        // large jobs' (1 or 2 stages each) taken in turn from a rotating
        // cursor while their work stays within one largest stage, at least one
        std::vector<pw::PassJob> pj;
        std::vector<int64_t> pid;
        double ref = kPassFloor;
        for (int64_t id : s->resident) ref = std::max(ref, stage_cost(s->jobs[id]));
        ref = std::min(ref, std::max(kPassFloor, kPassCap));
        std::vector<int64_t> large;
        for (int64_t id : s->resident) {
            SJob &j = s->jobs[id];
            const int k = stages_this_pass(j, ref);
            if (4 * stage_cost(j) <= ref) {
                pj.push_back(pw::PassJob{j.job, j.stage, k});
                pid.push_back(id);
            } else {
                large.push_back(id);
            }
        }
        if (!large.empty()) {
            // in id order, from the first large job at or after the cursor
            std::sort(large.begin(), large.end());
            size_t first = 0;
            while (first < large.size() && large[first] < s->cursor) ++first;
            if (first == large.size()) first = 0;
            double work = 0.0;
            for (size_t t = 0; t < large.size(); ++t) {
                const int64_t id = large[(first + t) % large.size()];
                SJob &j = s->jobs[id];
                const double c = stage_cost(j);
                // a stage starting now is sliced by the current reference
                if (j.part == 0) j.parts = c > ref ? (int)std::ceil(c / ref) : 1;
                const int k = j.parts > 1 ? 1 : stages_this_pass(j, ref);
                const double wk = j.parts > 1 ? c / j.parts : k * c;
                if (t > 0 && work + wk > ref) break;
                work += wk;
                pw::PassJob p{j.job, j.stage, k};
                p.part = j.part;
                p.parts = j.parts;
                pj.push_back(p);
                pid.push_back(id);
                s->cursor = id + 1;
            }
        }
        std::vector<int64_t> joining;
        std::exception_ptr admit_err;
        if (!pj.empty()) {
            s->eng->pass_begin(pj, true);
            // admission beside the running pass; the new jobs join the next
            // one.  The pass has run whatever admission does: its stages are
            // accounted for before an admission error is passed on.
            try {
                s->admit(true, &joining);
            } catch (...) {
                admit_err = std::current_exception();
            }
            std::vector<char> bad;
            s->eng->pass_end(&bad);
            s->st.steps++;
            s->st.pairs += s->eng->last_pairs;
            s->st.busy_ms += s->eng->last_total_ms;
            // advance the jobs of the pass, then retire finished and failed
            // ones; the others keep their place (admission order)
            std::map<int64_t, size_t> in_pass;
            for (size_t k = 0; k < pid.size(); ++k) in_pass[pid[k]] = k;
            std::vector<int64_t> keep;
            for (const int64_t id : s->resident) {
                auto ip = in_pass.find(id);
                if (ip == in_pass.end()) {
                    keep.push_back(id);
                    continue;
                }
                const size_t k = ip->second;
                SJob &j = s->jobs[id];
                if (pj[k].part + 1 < pj[k].parts) {  // the stage continues next time
                    j.part++;
                    keep.push_back(id);
                    continue;
                }
                j.part = 0;
                j.parts = 1;
                j.stage = (j.stage + pj[k].stages) % 4;
                j.stages_left -= (size_t)pj[k].stages;
                j.done = (4 * j.nsteps - j.stages_left) / 4;
                if (!bad[k] && j.stages_left > 0) {
                    keep.push_back(id);
                    continue;
                }
                if (!bad[k]) {
                    j.pos.resize(3 * j.n);
                    j.vort.resize(3 * j.n);
                    s->eng->download_aos(j.job, 0, j.pos.data());
                    s->eng->download_aos(j.job, 1, j.vort.data());
                }
                s->eng->remove_job(id);
                j.job = nullptr;
                s->in_use -= j.bytes;
                j.state = bad[k] ? JS_FAILED : JS_DONE;
                j.finished = s->now();
                if (bad[k]) {
                    s->st.failed++;
                } else {
                    s->st.finished++;
                    s->job_ms.push_back(j.finished - j.arrival);
                    s->job_step_ms.push_back((j.finished - j.admitted) / (double)j.nsteps);
                }
            }
            s->resident.swap(keep);
            s->step_ms.push_back(s->now() - t_begin);
        }
        s->resident.insert(s->resident.end(), joining.begin(), joining.end());
        if (running) *running = (int64_t)s->resident.size();
        if (admit_err) std::rethrow_exception(admit_err);
    });
}

int pw_server_job_state(pw_server *s, int64_t id, int64_t *steps_done) {
    if (!s) return -1;
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return -1;
    if (steps_done) *steps_done = (int64_t)it->second.done;
    return (int)it->second.state;
}

int pw_server_fetch(pw_server *s, int64_t id, double *pos, double *vort) {
    if (!s) return PW_EINVAL;
    return guarded(&s->err, [&] {
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) throw Error(PW_ENOTFOUND, "unknown job");
        SJob &j = it->second;
        if (j.state == JS_FAILED) throw Error(PW_ENONFINITE, "job failed with non-finite values");
        if (j.state != JS_DONE) throw Error(PW_EINVAL, "job not finished");
        if (pos) std::memcpy(pos, j.pos.data(), j.pos.size() * sizeof(double));
        if (vort) std::memcpy(vort, j.vort.data(), j.vort.size() * sizeof(double));
        s->jobs.erase(it);
    });
}

int pw_server_get_stats(pw_server *s, pw_server_stats *st) {
    if (!s || !st) return PW_EINVAL;
    *st = s->st;
    st->queued = (int64_t)s->queue.size();
    st->running = (int64_t)s->resident.size();
    st->step_p50_ms = pct(s->step_ms, 0.50);
    st->step_p99_ms = pct(s->step_ms, 0.99);
    st->job_p50_ms = pct(s->job_ms, 0.50);
    st->job_p99_ms = pct(s->job_ms, 0.99);
    st->job_step_p50_ms = pct(s->job_step_ms, 0.50);
    st->job_step_p99_ms = pct(s->job_step_ms, 0.99);
    st->hbm_in_use = s->in_use;
    st->hbm_budget = s->budget;
    return PW_OK;
}

}  // extern "C"

// -------------------------------------------------------------- autotuner
extern "C" {

int pw_autotune(int device, int kernel, int precision, size_t n, const double *pos,
                const double *vort, const double *radius, const double *volume, int reps,
                pw_tune_entry *all, size_t max_all, size_t *n_all, pw_tune_entry *best) {
    if (!best || !pos || !vort || !radius || !volume || n == 0) return PW_EINVAL;
    return guarded(nullptr, [&] {
        if (precision != PW_FP64 && precision != PW_FP32C) throw Error(PW_EINVAL, "precision");
        // mean device time of a whole solve (pair pass + reduction)
        auto time_solve = [&](const pw::Variant &v, int chunks) {
            Engine eng(device, kernel, v);
            JobParams par;
            par.precision = precision;
            par.chunks = chunks;
            eng.add_job(1, n, pos, vort, radius, volume, par);
            eng.solve();  // warm-up (plan, code object load)
            double ms = 0.0;
            const int r = std::max(1, reps);
            for (int k = 0; k < r; ++k) {
                eng.solve();
                ms += eng.last_total_ms;
            }
            return ms / r;
        };
        auto entry = [&](int b, int t, int u, int sc, int chunks, double ms) {
            pw_tune_entry e;
            e.variant.block = b;
            e.variant.targets_per_thread = t;
            e.variant.unroll = sc == PW_SCHEME_SYMMETRIC ? 0 : u;
            e.variant.scheme = sc;
            e.variant.chunks = chunks;
            e.pair_ms = ms;
            e.pairs_per_s = (double)n * (double)(n - 1) / (ms * 1e-3);
            return e;
        };
        // 1. every compiled variant at the chunk policy.  512: binary32
        // symmetric pass only (binary64 jobs would run at 256)
        const int blocks[4] = {64, 128, 256, 512}, tpts[2] = {1, 2}, unrs[3] = {1, 2, 4};
        const int schemes[2] = {PW_SCHEME_SYMMETRIC, PW_SCHEME_ONESIDED};
        std::vector<pw_tune_entry> res;
        std::vector<pw::Variant> vars;
        for (int sc : schemes)
            for (int b : blocks)
                for (int t : tpts)
                    for (int u : unrs) {
                        const bool sym = sc == PW_SCHEME_SYMMETRIC;
                        if (sym && u != unrs[0]) continue;
                        if (b == 512 && (precision != PW_FP32C || !sym)) continue;
                        pw::Variant v;
                        v.block = b;
                        v.tpt = t;
                        v.sym32 = sym ? 1 : 0;
                        v.sym64 = sym ? 1 : 0;
                        if (!sym) {
                            if (precision == PW_FP64)
                                v.unroll64 = u;
                            else
                                v.unroll32 = u;
                        }
                        if (!pw::variant_exists(kernel, v.block, v.tpt, v.unroll64, v.unroll32,
                                                v.sym32, v.sym64))
                            continue;
                        res.push_back(entry(b, t, u, sc, 0, time_solve(v, 0)));
                        vars.push_back(v);
                    }
        if (res.empty()) throw Error(PW_EINVAL, "no variant to tune");
        // 2. the fastest variant of each scheme at other chunk counts (on the
        // symmetric pass the count is also the strip count: fewer strips mean
        // fewer slot rows to reduce, longer items to balance)
        const size_t nbase = res.size();
        for (int sc : schemes) {
            size_t k = nbase;
            for (size_t i = 0; i < nbase; ++i)
                if (res[i].variant.scheme == sc && (k == nbase || res[i].pair_ms < res[k].pair_ms))
                    k = i;
            if (k == nbase) continue;
            const pw::Variant &v = vars[k];
            const pw_variant &pv = res[k].variant;
            const bool sym = sc == PW_SCHEME_SYMMETRIC;
            const int tb = pw::block_for(pv.block, precision), tpt = pv.targets_per_thread;
            const size_t L0 = pw::chunk_len(n, tb, tpt, sym, 0);
            const int c0 = (int)((n + L0 - 1) / L0);
            const size_t unit = sym ? (size_t)tb * tpt : (size_t)tb;
            const int cmax = (int)std::min<size_t>((n + unit - 1) / unit, 4096);
            std::vector<size_t> seen{L0};
            for (int c : {c0 / 2, (3 * c0) / 4, (3 * c0) / 2, 2 * c0}) {
                if (c < 1 || c > cmax) continue;
                const size_t L = pw::chunk_len(n, tb, tpt, sym, c);
                if (std::find(seen.begin(), seen.end(), L) != seen.end()) continue;
                seen.push_back(L);
                const int got = (int)((n + L - 1) / L);
                res.push_back(entry(pv.block, tpt, pv.unroll, sc, got, time_solve(v, got)));
            }
        }
        size_t k = 0;
        for (size_t i = 1; i < res.size(); ++i)
            if (res[i].pair_ms < res[k].pair_ms) k = i;
        *best = res[k];
        if (all)
            for (size_t i = 0; i < res.size() && i < max_all; ++i) all[i] = res[i];
        if (n_all) *n_all = res.size();
    });
}

static int n_bucket(size_t n) {
    int b = 0;
    while ((size_t)1 << (b + 1) <= n) ++b;
    return b;
}

int pw_tune_save(const char *path, int kernel, int precision, size_t n, const pw_tune_entry *e) {
    if (!path || !e) return PW_EINVAL;
    return guarded(nullptr, [&] {
        std::vector<std::string> keep;
        const int bk = n_bucket(n);
        {
            std::ifstream in(path);
            std::string line;
            while (std::getline(in, line)) {
                std::istringstream ss(line);
                int k, p, b;
                if (ss >> k >> p >> b && k == kernel && p == precision && b == bk) continue;
                if (!line.empty()) keep.push_back(line);
            }
        }
        std::ofstream out(path, std::ios::trunc);
        if (!out) throw Error(PW_EINVAL, "cannot write the tuning cache");
        for (auto &l : keep) out << l << "\n";
        out << kernel << " " << precision << " " << bk << " " << e->variant.block << " "
            << e->variant.targets_per_thread << " " << e->variant.unroll << " "
            << e->variant.scheme << " " << e->pair_ms << " " << e->variant.chunks << "\n";
    });
}

int pw_tune_lookup(const char *path, int kernel, int precision, size_t n, pw_variant *out) {
    if (!path || !out) return PW_EINVAL;
    std::ifstream in(path);
    if (!in) return PW_ENOTFOUND;
    const int bk = n_bucket(n);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        int k, p, b, blk, tpt, unr, sc;
        if (ss >> k >> p >> b >> blk >> tpt >> unr >> sc && k == kernel && p == precision &&
            b == bk) {
            const int sym32 = precision == PW_FP32C && sc != PW_SCHEME_ONESIDED;
            const int sym64 = precision == PW_FP64 && sc == PW_SCHEME_SYMMETRIC;
            if (sc < PW_SCHEME_DEFAULT || sc > PW_SCHEME_SYMMETRIC ||
                !pw::variant_exists(kernel, blk, tpt, precision == PW_FP64 && unr ? unr : 2,
                                    precision == PW_FP32C && unr ? unr : 4, sym32, sym64))
                return PW_EINVAL;
            double ms = 0.0;
            int chunks = 0;  // absent in lines written before chunk counts were tuned
            if (ss >> ms && !(ss >> chunks)) chunks = 0;
            if (chunks < 0) return PW_EINVAL;
            out->block = blk;
            out->targets_per_thread = tpt;
            out->unroll = unr;
            out->scheme = sc;
            out->chunks = chunks;
            return PW_OK;
        }
    }
    return PW_ENOTFOUND;
}

}  // extern "C"
