// adapter_driver.cpp -- runs pawan's call sequence through the drop-in adapter:
//   W = ring; S = new __gpu_interaction(W); S->solve(); read W->_velocity ...
//   IN = new __gpu_rk4(t, n); IN->integrate(S, file)
// and checks every result against the CPU oracle (liboracle.so).
// Usage: adapter_driver <wake-file>   (exit status 0 = all checks passed)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/pawan_gpu.hpp"
#include "../../oracle/vpm_oracle.h"
#include "mini_pawan.hpp"

using gpu_interaction = pawan_gpu::interaction<mini::interaction, mini::wake, mini::vec>;
using gpu_rk4 = pawan_gpu::rk4<mini::rk4, mini::system, mini::vec>;

static mini::wake *ring(size_t n_azi, size_t n_cs, double R, double z0) {
    auto *W = new mini::wake(n_azi * n_cs);
    const double core = 0.1, h = 2 * core / 4.0, dpsi = 2 * M_PI / n_azi;
    size_t p = 0;
    for (size_t a = 0; a < n_azi; ++a)
        for (size_t c = 0; c < n_cs; ++c, ++p) {
            double xi = -core + h * (c % 4 + 0.5), eta = -core + h * (c / 4 + 0.5);
            double psi = a * dpsi + 1e-3 * std::sin(7.0 * p), rho = R + xi;
            W->_position->data[3 * p] = rho * std::cos(psi);
            W->_position->data[3 * p + 1] = rho * std::sin(psi);
            W->_position->data[3 * p + 2] = z0 + eta;
            double g = std::exp(-(xi * xi + eta * eta) / (core * core)) / n_cs * rho * dpsi;
            W->_vorticity->data[3 * p] = -g * std::sin(psi);
            W->_vorticity->data[3 * p + 1] = g * std::cos(psi);
            W->_vorticity->data[3 * p + 2] = 0.0;
            W->_radius->data[p] = 1.3 * h;
            W->_volume->data[p] = h * h * rho * dpsi;
        }
    return W;
}

static double relerr(const std::vector<double> &a, const std::vector<double> &b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num = std::fmax(num, std::fabs(a[i] - b[i]));
        den = std::fmax(den, std::fabs(b[i]));
    }
    return num / den;
}

int main(int argc, char **argv) {
    const char *path = argc > 1 ? argv[1] : "adapter.wake";
    mini::wake *W1 = ring(64, 16, 1.0, 0.0), *W2 = ring(48, 16, 0.8, 0.3);
    const size_t n1 = W1->_numParticles, n2 = W2->_numParticles, n = n1 + n2;
    pawan_gpu::options o;
    o.nu = 1e-3;
    int bad = 0;
    try {
        // pawan.cpp:30 with the adapter named by its own type, so that the
        // `delete S` of pawan.cpp:67 reaches its destructor (pawan's
        // __interaction destructor is not virtual, interaction.h:175)
        auto *S = new gpu_interaction(W1, W2, o);  // __interaction(W1, W2)
        std::printf("driver: N = %zu + %zu, _size = %zu\n", n1, n2, S->_size);
        // the oracle's view of the same particles
        std::vector<double> pos(3 * n), vort(3 * n), rad(n), vol(n);
        size_t wn[2] = {n1, n2};
        for (int k = 0; k < 2; ++k) {
            mini::wake *W = k ? W2 : W1;
            size_t o0 = k ? n1 : 0;
            for (size_t i = 0; i < W->_numParticles; ++i) {
                for (int c = 0; c < 3; ++c) {
                    pos[3 * (o0 + i) + c] = W->_position->data[3 * i + c];
                    vort[3 * (o0 + i) + c] = W->_vorticity->data[3 * i + c];
                }
                rad[o0 + i] = W->_radius->data[i];
                vol[o0 + i] = W->_volume->data[i];
            }
        }
        std::vector<double> vel(3 * n), dv(3 * n);
        orc_rates(ORC_MODE_OMP_CLEAN, ORC_KERNEL_HIGHORDER, 1e-3, 2, wn, pos.data(), vort.data(),
                  rad.data(), vol.data(), nullptr, vel.data(), dv.data());
        // solve(): rates land in each wake's _velocity / _retvorcity
        S->solve();
        std::vector<double> gv(W1->_velocity->data, W1->_velocity->data + 3 * n1),
            gd(W1->_retvorcity->data, W1->_retvorcity->data + 3 * n1);
        gv.insert(gv.end(), W2->_velocity->data, W2->_velocity->data + 3 * n2);
        gd.insert(gd.end(), W2->_retvorcity->data, W2->_retvorcity->data + 3 * n2);
        double ev = relerr(gv, vel), ed = relerr(gd, dv);
        std::printf("driver: solve() vs oracle: velocity %.3e vorticity rate %.3e\n", ev, ed);
        bad |= !(ev < 1e-10 && ed < 1e-10);
        {  // the compensated binary32 path (symmetric pass) through the same adapter
            pawan_gpu::options o32 = o;
            o32.precision = PW_FP32C;
            gpu_interaction S32(W1, W2, o32);
            S32.solve();
            std::vector<double> sv(W1->_velocity->data, W1->_velocity->data + 3 * n1),
                sd(W1->_retvorcity->data, W1->_retvorcity->data + 3 * n1);
            sv.insert(sv.end(), W2->_velocity->data, W2->_velocity->data + 3 * n2);
            sd.insert(sd.end(), W2->_retvorcity->data, W2->_retvorcity->data + 3 * n2);
            double e32v = relerr(sv, vel), e32d = relerr(sd, dv);
            std::printf("driver: binary32 solve() vs oracle: velocity %.3e vorticity rate %.3e\n",
                        e32v, e32d);
            bad |= !(e32v < 1e-5 && e32d < 1e-5);
        }
        // resolve(): each wake's _vorticityfield
        std::vector<double> fld(3 * n);
        orc_influence(ORC_KERNEL_HIGHORDER, 2, wn, pos.data(), vort.data(), rad.data(), fld.data());
        S->resolve();
        std::vector<double> gf(W1->_vorticityfield->data, W1->_vorticityfield->data + 3 * n1);
        gf.insert(gf.end(), W2->_vorticityfield->data, W2->_vorticityfield->data + 3 * n2);
        double ef = relerr(gf, fld);
        std::printf("driver: resolve() field vs oracle: %.3e\n", ef);
        bad |= !(ef < 1e-12);
        // getStates / getRates layout
        mini::vec st(S->_size), rt(S->_size), ref(S->_size);
        S->getStates(&st);
        orc_pack_state(2, wn, pos.data(), vort.data(), ref.data);
        bool same = true;
        for (size_t i = 0; i < S->_size; ++i) same = same && st.data[i] == ref.data[i];
        std::printf("driver: getStates layout %s\n", same ? "exact" : "WRONG");
        bad |= !same;
        // __rk4(t, n).integrate(S, IO): 3 steps of dt = 0.01, a frame per step
        std::string base(path);
        if (base.size() > 5 && base.compare(base.size() - 5, 5, ".wake") == 0)
            base.resize(base.size() - 5);
        mini::io IO("", base);  // IO->create_binary_file(".wake") opens `path`
        gpu_rk4 IN(0.03, 3);
        IN.integrate(S, &IO);
        std::vector<double> x0(ref.data, ref.data + S->_size), xg(S->_size);
        mini::vec fin(S->_size);
        S->getStates(&fin);
        xg.assign(fin.data, fin.data + S->_size);
        orc_rk4(ORC_MODE_OMP_CLEAN, ORC_KERNEL_HIGHORDER, 1e-3, 2, wn, rad.data(), vol.data(),
                nullptr, x0.data(), 0.01, 3);
        double ex = relerr(xg, x0);
        std::printf("driver: integrate() 3 RK4 steps vs oracle: %.3e\n", ex);
        bad |= !(ex < 1e-12);
        // the last frame of the .wake file equals the final state
        FILE *f = std::fopen(path, "rb");
        std::fseek(f, 0, SEEK_END);
        long sz = std::ftell(f);
        long frame = 8 + 8 + 2 * 8 + 8 * (7 * n);
        std::printf("driver: .wake file %ld bytes = 4 frames of %ld\n", sz, frame);
        bad |= sz != 4 * frame;
        std::fclose(f);
        {  // the same integration with asynchronous frames: the same bytes
            mini::wake *A1 = ring(64, 16, 1.0, 0.0), *A2 = ring(48, 16, 0.8, 0.3);
            std::string p2 = std::string(path) + ".async";
            {
                gpu_interaction SA(A1, A2, o);
                gpu_rk4 INA(0.03, 3);
                FILE *fa = std::fopen(p2.c_str(), "wb");
                INA.integrate_async(&SA, fa);
                std::fclose(fa);
            }
            FILE *f1 = std::fopen(path, "rb"), *f2 = std::fopen(p2.c_str(), "rb");
            std::vector<char> b1(sz), b2(sz + 1);
            size_t r1 = std::fread(b1.data(), 1, sz, f1), r2 = std::fread(b2.data(), 1, sz + 1, f2);
            std::fclose(f1);
            std::fclose(f2);
            bool eq = r1 == (size_t)sz && r2 == (size_t)sz &&
                      std::equal(b1.begin(), b1.end(), b2.begin());
            std::printf("driver: integrate_async() frames %s\n",
                        eq ? "byte-identical" : "DIFFER");
            bad |= !eq;
            delete A1;
            delete A2;
        }
        {  // __resolve::rebuild's sequence (resolve.cpp:18-31) on a wake whose state is one
           // strength per particle: column i of the influence matrix is getRates() after
           // setStates(e_i) and resolve()
            const size_t m = 48;
            auto *R = new mini::strength_wake(m);
            for (size_t i = 0; i < m; ++i) {
                const double psi = 2.0 * M_PI * (double)i / (double)m;
                R->_position->data[3 * i] = std::cos(psi);
                R->_position->data[3 * i + 1] = std::sin(psi);
                R->_position->data[3 * i + 2] = 0.02 * std::sin(3.0 * psi);
                R->dir[3 * i] = -std::sin(psi);
                R->dir[3 * i + 1] = std::cos(psi);
                R->dir[3 * i + 2] = 0.0;
                R->_radius->data[i] = 0.15;
                R->_volume->data[i] = 1e-3;
            }
            std::vector<double> A(m * m), Aref(m * m);
            {
                gpu_interaction SR(R, o);
                mini::vec st(SR._size), col(SR._size);
                for (size_t i = 0; i < m; ++i) {
                    for (size_t k = 0; k < m; ++k) st.data[k] = k == i ? 1.0 : 0.0;
                    SR.setStates(&st);
                    SR.resolve();
                    SR.getRates(&col);
                    for (size_t k = 0; k < m; ++k) A[k * m + i] = col.data[k];
                }
                SR.getStates(&st);
                bool sst = true;
                for (size_t k = 0; k < m; ++k) sst = sst && st.data[k] == (k == m - 1 ? 1.0 : 0.0);
                std::printf("driver: strength-wake getStates %s\n", sst ? "exact" : "WRONG");
                bad |= !sst;
            }
            std::vector<double> rp(R->_position->data, R->_position->data + 3 * m), rv(3 * m),
                rr(R->_radius->data, R->_radius->data + m), fl(3 * m);
            size_t wm[1] = {m};
            for (size_t i = 0; i < m; ++i) {
                std::fill(rv.begin(), rv.end(), 0.0);
                for (int c = 0; c < 3; ++c) rv[3 * i + c] = R->dir[3 * i + c];
                orc_influence(ORC_KERNEL_HIGHORDER, 1, wm, rp.data(), rv.data(), rr.data(), fl.data());
                for (size_t k = 0; k < m; ++k)
                    Aref[k * m + i] = std::sqrt(fl[3 * k] * fl[3 * k] + fl[3 * k + 1] * fl[3 * k + 1] +
                                                fl[3 * k + 2] * fl[3 * k + 2]);
            }
            const double ea = relerr(A, Aref);
            std::printf("driver: rebuild influence matrix (strength wake, %zu x %zu) vs oracle: %.3e\n",
                        m, m, ea);
            bad |= !(ea < 1e-12);
            delete R;
        }
        double om[3], im[3], ke;
        S->device_diagnostics(om, im, &ke);
        std::printf("driver: kinetic energy %.12e\n", ke);
        {  // a caller holding the adapter through pawan's base type releases it first
            mini::interaction *B = new gpu_interaction(W1, o);
            const size_t before = pawan_gpu::live_contexts();
            pawan_gpu::release(B);
            const size_t after = pawan_gpu::live_contexts();
            bool threw = false;
            try {
                B->solve();
            } catch (const std::exception &) {
                threw = true;  // the context is gone
            }
            delete static_cast<gpu_interaction *>(B);
            std::printf("driver: release() through the base: contexts %zu -> %zu, later solve() %s\n",
                        before, after, threw ? "refused" : "RAN");
            bad |= !(after + 1 == before && threw);
        }
        delete S;  // pawan.cpp:67
        std::printf("driver: contexts alive after delete S: %zu\n", pawan_gpu::live_contexts());
        bad |= pawan_gpu::live_contexts() != 0;
    } catch (const std::exception &e) {
        std::printf("driver: exception %s\n", e.what());
        bad = 1;
    }
    delete W1;
    delete W2;
    std::printf("driver: %s\n", bad ? "FAILED" : "passed");
    return bad;
}
