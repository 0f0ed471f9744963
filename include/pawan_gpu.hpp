// pawan_gpu.hpp -- drop-in C++ adapter: pawan's interaction and RK4 classes
// backed by the gfx950 C ABI (pawan_hip.h).
//
// The adapter is a template over pawan's own classes so that the same header
// compiles inside pawan's tree:
//
//   using __gpu_interaction = pawan_gpu::interaction<pawan::__interaction, pawan::__wake, gsl_vector>;
//   using __gpu_rk4 = pawan_gpu::rk4<pawan::__rk4, pawan::__system, gsl_vector>;
//
// Requirements on the template arguments (all met by pawan's classes):
//   Interaction  virtual addWake(Wake*), solve(), resolve(), write(FILE*), diagnose(),
//                setStates(const Vector*), getRates(Vector*), getStates(Vector*)
//                (interaction.h:181-234); protected double _nu (interaction.h:24)
//   Wake         _size, _numParticles; _position, _vorticity, _velocity,
//                _retvorcity and _vorticityfield with ->data holding N x 3
//                row-major doubles (tda == 3);
//                _radius, _volume with ->data holding N doubles (wake.h:33-42)
//   Vector       ->data holding the state/rate vector (gsl_vector, stride 1)
//   RK4          RK4(const double &t, const size_t &n); protected virtual
//                step(const double &dt, System *S, Vector *state) (rk4.h:14-41)
//
// Semantics follow the reference: solve() fills every wake's _velocity and
// _retvorcity; setStates() also updates the wakes' _position/_vorticity;
// write() and diagnose() see the device state.  A wake whose state is not
// [positions 3N | vorticities 3N] (_size != 6 _numParticles: __vring, whose
// state is its cross-section strengths, vring.cpp:55,110-128) keeps its own
// state and rates: setStates/getStates/getRates then go through the wakes'
// own virtuals (interaction.cpp:204-229) and the device is given the
// particles' positions and vorticities those produce, so __resolve::rebuild
// (resolve.cpp:10-57) runs on the device-backed interaction unchanged.  RK4 steps run entirely on
// the device (pw_rk4_steps); the integrator's state vector is refreshed after
// each step, as __rk4::step leaves it.
//
// Lifetime.  pawan's __system and __interaction destructors are not virtual
// (system.h:34, interaction.h:175), so `delete S` through an __interaction*
// never runs this adapter's destructor.  Delete the adapter through its own
// type (`auto *S = new __gpu_interaction(W, o); ... delete S;`), or call
// pawan_gpu::release(S) before deleting it through the base.  Every device
// context the adapter creates is also listed in a process-wide registry that
// destroys the ones still alive when the program exits.
#pragma once

#include <cstddef>
#include <cstdio>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "pawan_hip.h"

namespace pawan_gpu {

inline void check(int rc, const pw_ctx *ctx) {
    if (rc != PW_OK)
        throw std::runtime_error(std::string("pawan_hip: ") + pw_last_error(ctx) + " (status " +
                                 std::to_string(rc) + ")");
}

namespace detail {
// Device contexts created by adapters and not yet destroyed.  Constructed by
// the first adapter after its pw_create returns, so that it is destroyed at
// exit before the HIP runtime state that pw_create brought up; its destructor
// destroys the contexts still listed (adapters deleted through a base pointer
// or never deleted).
struct registry {
    std::mutex m;
    std::set<pw_ctx *> live;
    ~registry() {
        for (pw_ctx *c : live) pw_destroy(c);
    }
};
inline registry &contexts() {
    static registry r;
    return r;
}
inline void track(pw_ctx *c) {
    registry &r = contexts();
    std::lock_guard<std::mutex> g(r.m);
    r.live.insert(c);
}
inline void untrack_and_destroy(pw_ctx *c) {
    if (!c) return;
    {
        registry &r = contexts();
        std::lock_guard<std::mutex> g(r.m);
        r.live.erase(c);
    }
    pw_destroy(c);
}
}  // namespace detail

// device contexts created by adapters and still alive
inline size_t live_contexts() {
    detail::registry &r = detail::contexts();
    std::lock_guard<std::mutex> g(r.m);
    return r.live.size();
}

// Common base of device-backed systems, so an integrator can reach the context
// through the reference's __system pointer.
class device_system {
  public:
    virtual ~device_system() = default;
    virtual pw_ctx *context() = 0;
    // copy device positions, vorticities and last rates into the wakes
    virtual void pull() = 0;
    // destroy the device context now (its stream and device memory); later
    // calls on the system fail with PW_EINVAL.  Idempotent.
    virtual void release() = 0;
};

// Release the device context of a pawan_gpu system held through any of its
// bases (e.g. pawan::__interaction *), before a `delete` through that base,
// which cannot reach the adapter's destructor.  No effect on other systems.
template <class System>
void release(System *S) {
    if (device_system *d = dynamic_cast<device_system *>(S)) d->release();
}

struct options {
    double nu = 2.5e-3;  // reference default, interaction.cpp:13
    int precision = PW_FP64;
    int kernel = PW_KERNEL_HIGHORDER;
    int device = 0;
    double uinf[3] = {0.0, 0.0, 0.0};
};

template <class Interaction, class Wake, class Vector>
class interaction : public Interaction, public device_system {
  public:
    explicit interaction(const options &o = options()) : Interaction() {
        pw_config c;
        pw_config_default(&c);
        c.nu = o.nu;
        c.precision = o.precision;
        c.kernel = o.kernel;
        c.device = o.device;
        for (int k = 0; k < 3; ++k) c.uinf[k] = o.uinf[k];
        check(pw_create(&ctx_, &c), nullptr);
        detail::track(ctx_);
        this->_nu = o.nu;
    }
    // __interaction(W) and __interaction(W1, W2), interaction.cpp:20-27
    explicit interaction(Wake *W, const options &o = options()) : interaction(o) { addWake(W); }
    interaction(Wake *W1, Wake *W2, const options &o = options()) : interaction(o) {
        addWake(W1);
        addWake(W2);
    }
    ~interaction() override { release(); }
    interaction(const interaction &) = delete;
    interaction &operator=(const interaction &) = delete;

    pw_ctx *context() override { return ctx_; }
    void release() override {
        detail::untrack_and_destroy(ctx_);
        ctx_ = nullptr;
    }

    // interaction.cpp:35-39, plus a copy of the particles to the device
    void addWake(Wake *W) override {
        Interaction::addWake(W);
        wakes_.push_back(W);
        check(pw_add_wake(ctx_, W->_numParticles, W->_position->data, W->_vorticity->data,
                          W->_radius->data, W->_volume->data),
              ctx_);
    }

    // interaction.cpp:41-43 -> interact(): rates of every particle
    void solve() override {
        check(pw_solve(ctx_), ctx_);
        pull_rates();
    }

    // interaction.cpp:204-211
    void setStates(const Vector *state) override {
        Interaction::setStates(state);
        if (plain()) {
            check(pw_set_state(ctx_, state->data), ctx_);
            return;
        }
        // the wakes' particles as the wakes' own setStates left them
        std::vector<double> st;
        for (Wake *W : wakes_) {
            const size_t m = 3 * W->_numParticles;
            st.insert(st.end(), W->_position->data, W->_position->data + m);
            st.insert(st.end(), W->_vorticity->data, W->_vorticity->data + m);
        }
        check(pw_set_state(ctx_, st.data()), ctx_);
    }
    // interaction.cpp:213-229
    void getRates(Vector *rate) override {
        if (plain())
            check(pw_get_rates(ctx_, rate->data), ctx_);
        else
            Interaction::getRates(rate);  // from the rates solve()/resolve() left in the wakes
    }
    void getStates(Vector *state) override {
        if (plain()) {
            check(pw_get_state(ctx_, state->data), ctx_);
        } else {
            pull();
            Interaction::getStates(state);
        }
    }

    // interaction.cpp:45-47 -> influence(): every wake's _vorticityfield
    void resolve() override {
        size_t total = 0;
        for (Wake *W : wakes_) total += W->_numParticles;
        std::vector<double> field(3 * total);
        check(pw_resolve(ctx_, field.data()), ctx_);
        size_t o = 0;
        for (Wake *W : wakes_) {
            for (size_t k = 0; k < 3 * W->_numParticles; ++k) W->_vorticityfield->data[k] = field[o + k];
            o += 3 * W->_numParticles;
        }
    }

    // interaction.cpp:197-202: frames carry the device state
    void write(FILE *f) override {
        pull();
        Interaction::write(f);
    }
    // interaction.cpp:49-93 on the device state
    void diagnose() override {
        pull();
        Interaction::diagnose();
    }

    void pull() override {
        for (size_t k = 0; k < wakes_.size(); ++k) {
            Wake *W = wakes_[k];
            check(pw_get_wake(ctx_, k, W->_position->data, W->_vorticity->data,
                              W->_velocity->data, W->_retvorcity->data),
                  ctx_);
        }
    }

    // kinetic energy (interaction.cpp:351-383) with total vorticity and
    // linear impulse, computed on the device
    void device_diagnostics(double omega[3], double impulse[3], double *ke) {
        check(pw_diagnostics(ctx_, omega, impulse, ke), ctx_);
    }

  private:
    // every wake's state is its particles' [positions | vorticities]
    bool plain() const {
        for (Wake *W : wakes_)
            if (W->_size != 6 * W->_numParticles) return false;
        return true;
    }
    void pull_rates() {
        for (size_t k = 0; k < wakes_.size(); ++k) {
            Wake *W = wakes_[k];
            check(pw_get_wake(ctx_, k, nullptr, nullptr, W->_velocity->data,
                              W->_retvorcity->data),
                  ctx_);
        }
    }
    pw_ctx *ctx_ = nullptr;
    std::vector<Wake *> wakes_;
};

// __integration::integrate (integration.cpp:16-35) with its frames written
// asynchronously (pw_write_async): f receives the same bytes -- the time, then
// S->write(f)'s frame, at t = 0 and after every step -- but a step does not
// wait for the previous frame's copy or file write.  f stays open.
inline void integrate_frames(device_system *d, FILE *f, double dt, size_t n, bool rk) {
    pw_ctx *c = d->context();
    check(pw_write_async(c, f, 0.0, 1), c);
    for (size_t i = 1; i <= n; ++i) {
        check(rk ? pw_rk4_steps(c, dt, 1) : pw_euler_steps(c, dt, 1), c);
        check(pw_write_async(c, f, i * dt, 1), c);
    }
    check(pw_write_wait(c), c);
    d->pull();
}

template <class System>
device_system *device_of(System *S) {
    device_system *d = dynamic_cast<device_system *>(S);
    if (!d) throw std::invalid_argument("pawan_gpu integrators step a pawan_gpu::interaction");
    return d;
}

// __rk4 whose step runs on the device (rk4.cpp:12-77 fused into pw_rk4_steps)
template <class RK4, class System, class Vector>
class rk4 : public RK4 {
  public:
    rk4(const double &t, const size_t &n) : RK4(t, n) {}
    // integrate(S, IO) with f = IO->create_binary_file(".wake") and the
    // frames written asynchronously; the caller closes f
    void integrate_async(System *S, FILE *f) {
        integrate_frames(device_of(S), f, this->_dt, this->_n, true);
    }

  protected:
    void step(const double &dt, System *S, Vector *states) override {
        device_system *d = dynamic_cast<device_system *>(S);
        if (!d) throw std::invalid_argument("pawan_gpu::rk4 steps a pawan_gpu::interaction");
        check(pw_rk4_steps(d->context(), dt, 1), d->context());
        S->getStates(states);
    }
};

// forward Euler, integration.cpp:37-45
template <class Integration, class System, class Vector>
class euler : public Integration {
  public:
    euler(const double &t, const size_t &n) : Integration(t, n) {}
    // integrate(S, IO) with asynchronous frames, as rk4::integrate_async
    void integrate_async(System *S, FILE *f) {
        integrate_frames(device_of(S), f, this->_dt, this->_n, false);
    }

  protected:
    void step(const double &dt, System *S, Vector *states) override {
        device_system *d = dynamic_cast<device_system *>(S);
        if (!d) throw std::invalid_argument("pawan_gpu::euler steps a pawan_gpu::interaction");
        check(pw_euler_steps(d->context(), dt, 1), d->context());
        S->getStates(states);
    }
};

}  // namespace pawan_gpu
