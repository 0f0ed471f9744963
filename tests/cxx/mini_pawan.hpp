// mini_pawan.hpp -- test double of pawan's class API (namespace pawan: __system,
// __wake, __interaction, __integration, __rk4) with plain vector/matrix structs
// in place of gsl_vector/gsl_matrix, so that the drop-in adapter
// include/pawan_gpu.hpp can be compiled and run here.  It reproduces the
// reference's observable behaviour that the adapter relies on: the state
// layout (wake.cpp:87-120), addWake bookkeeping (interaction.cpp:35-39), the
// .wake frame format (interaction.cpp:197-202, wake.cpp:68-73) and the
// integrate() loop (integration.cpp:16-35).  Its own CPU solve() is unused.
// Destructors are virtual exactly where the reference's are: __wake's is
// (wake.h:54); __system's, __interaction's, __integration's, __rk4's and
// __io's are not (system.h:34, interaction.h:175, integration.h:49,
// rk4.h:40, io.h:35), so deleting an adapter through a base pointer skips its
// destructor here as it does in pawan.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace mini {

struct vec {
    size_t size = 0;
    double *data = nullptr;
    bool owner = true;
    explicit vec(size_t n) : size(n), data(new double[n]()) {}
    // a view of n doubles at d, as gsl_vector_subvector gives (not owned)
    vec(double *d, size_t n) : size(n), data(d), owner(false) {}
    ~vec() {
        if (owner) delete[] data;
    }
    vec(const vec &) = delete;
    vec &operator=(const vec &) = delete;
};
struct mat {
    size_t size1 = 0, size2 = 3, tda = 3;
    double *data = nullptr;
    explicit mat(size_t n) : size1(n), data(new double[3 * n]()) {}
    ~mat() { delete[] data; }
};

class system {
  public:
    size_t _size = 0;
    ~system() = default;
    virtual void solve() {}
    virtual void write(FILE *) {}
    virtual void setStates(const vec *) {}
    virtual void getRates(vec *) {}
    virtual void getStates(vec *) {}
};

class wake {
  public:
    size_t _size;
    size_t _numParticles;
    mat *_position;
    mat *_velocity;
    mat *_vorticity;
    mat *_retvorcity;
    vec *_radius;
    vec *_volume;
    mat *_vorticityfield;
    explicit wake(size_t n)
        : _size(6 * n), _numParticles(n), _position(new mat(n)), _velocity(new mat(n)),
          _vorticity(new mat(n)), _retvorcity(new mat(n)), _radius(new vec(n)),
          _volume(new vec(n)), _vorticityfield(new mat(n)) {}
    virtual ~wake() {
        delete _position; delete _velocity; delete _vorticity; delete _retvorcity;
        delete _vorticityfield; delete _radius; delete _volume;
    }
    void write(FILE *f) {  // wake.cpp:68-73
        fwrite(&_numParticles, sizeof(size_t), 1, f);
        fwrite(_position->data, sizeof(double), 3 * _numParticles, f);
        fwrite(_vorticity->data, sizeof(double), 3 * _numParticles, f);
        fwrite(_radius->data, sizeof(double), _numParticles, f);
    }
    virtual void setStates(const vec *state) {  // wake.cpp:87-96
        const double *s = state->data;
        std::memcpy(_position->data, s, 3 * _numParticles * sizeof(double));
        std::memcpy(_vorticity->data, s + 3 * _numParticles, 3 * _numParticles * sizeof(double));
    }
    virtual void getStates(vec *state) {  // wake.cpp:109-120
        double *s = state->data;
        std::memcpy(s, _position->data, 3 * _numParticles * sizeof(double));
        std::memcpy(s + 3 * _numParticles, _vorticity->data, 3 * _numParticles * sizeof(double));
    }
    virtual void getRates(vec *rate) {  // wake.cpp:98-107
        double *r = rate->data;
        std::memcpy(r, _velocity->data, 3 * _numParticles * sizeof(double));
        std::memcpy(r + 3 * _numParticles, _retvorcity->data, 3 * _numParticles * sizeof(double));
    }
};

// A wake whose state is not its particles' positions and vorticities, as
// __vring's (vring.cpp:55,110-128): one strength per particle along a fixed
// unit direction; its rates are the magnitudes of the vorticity field that
// resolve() leaves in _vorticityfield.  _size = N, not 6 N.
class strength_wake : public wake {
  public:
    std::vector<double> dir, strength;
    explicit strength_wake(size_t n) : wake(n), dir(3 * n), strength(n) { _size = n; }
    void setStates(const vec *state) override {
        const double *s = state->data;
        for (size_t i = 0; i < _numParticles; ++i) {
            strength[i] = s[i];
            for (int c = 0; c < 3; ++c) _vorticity->data[3 * i + c] = s[i] * dir[3 * i + c];
        }
    }
    void getStates(vec *state) override {
        for (size_t i = 0; i < _numParticles; ++i) state->data[i] = strength[i];
    }
    void getRates(vec *rate) override {
        for (size_t i = 0; i < _numParticles; ++i) {
            const double *f = _vorticityfield->data + 3 * i;
            rate->data[i] = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        }
    }
};

class interaction : public system {
  protected:
    double _nu = 2.5e-3;
    size_t _nWake = 0;

  private:
    std::vector<wake *> _W;

  public:
    interaction() = default;
    ~interaction() {}
    virtual void addWake(wake *W) {
        _W.push_back(W);
        _size += W->_size;
        _nWake++;
    }
    virtual void diagnose() {}
    virtual void solve() {}
    virtual void resolve() {}
    void write(FILE *f) override {
        fwrite(&_nWake, sizeof(size_t), 1, f);
        for (auto *w : _W) w->write(f);
    }
    // interaction.cpp:204-229: each wake reads or writes its own _size slice
    // (a subvector view)
    void setStates(const vec *state) override {
        size_t off = 0;
        for (auto *w : _W) {
            vec st(state->data + off, w->_size);
            w->setStates(&st);
            off += w->_size;
        }
    }
    void getStates(vec *state) override {
        size_t off = 0;
        for (auto *w : _W) {
            vec st(state->data + off, w->_size);
            w->getStates(&st);
            off += w->_size;
        }
    }
    void getRates(vec *rate) override {
        size_t off = 0;
        for (auto *w : _W) {
            vec rt(rate->data + off, w->_size);
            w->getRates(&rt);
            off += w->_size;
        }
    }
};

// __io (io.h:20-60): create_binary_file(suffix) opens root + name + suffix
class io {
  public:
    std::string _root, _name;
    io(const std::string &root, const std::string &name) : _root(root), _name(name) {}
    ~io() = default;
    virtual FILE *create_binary_file(std::string suffix) {
        return std::fopen((_root + _name + suffix).c_str(), "wb");
    }
};

class integration {
  protected:
    double _dt, _t;
    size_t _n;
    virtual void step(const double &, system *, vec *) {}

  public:
    integration(const double &t, const size_t &n) : _dt(t / n), _t(t), _n(n) {}
    ~integration() = default;
    // integration.cpp:16-35
    void integrate(system *S, io *IO) {
        vec states(S->_size);
        FILE *f = IO->create_binary_file(".wake");
        double t = 0.0;
        std::fwrite(&t, sizeof(double), 1, f);
        S->write(f);
        S->getStates(&states);
        for (size_t i = 1; i <= _n; ++i) {
            t = i * _dt;
            step(_dt, S, &states);
            std::fwrite(&t, sizeof(double), 1, f);
            S->write(f);
        }
        std::fclose(f);
    }
};

class rk4 : public integration {
  protected:
    void step(const double &, system *, vec *) override {}  // rk4.cpp:12-77 (unused here)

  public:
    rk4(const double &t, const size_t &n) : integration(t, n) {}
    ~rk4() = default;
};

}  // namespace mini
