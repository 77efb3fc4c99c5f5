// csl_host.hpp — host-side helpers shared by the C-ABI translation units:
// error reporting, device buffers and the integrator time grids.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/csl_gpu.h"

namespace cslh {

// last error message of this thread (csl_last_error)
inline thread_local std::string g_err;

struct CslError : std::runtime_error {
    int code;
    CslError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw CslError(code, msg); }

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(CSL_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return CSL_OK;
    } catch (const CslError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return CSL_ERR_HIP;
    }
}

// Device buffer owned by a handle.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void ensure(size_t count) {
        if (count <= n && p) return;
        release();
        hip_check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
        n = count;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        ensure(count);
        if (count) hip_check(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s), "H2D");
    }
};

// Time grid of an integration (host side): start time of every step plus the
// time after the last step.
inline std::vector<double> make_grid(double T0, double T1, double dt, int policy) {
    if (!(dt > 0.0) || !std::isfinite(T0) || !std::isfinite(T1))
        fail(CSL_ERR_ARG, "time grid needs finite T0, T1 and dt > 0");
    std::vector<double> g;
    if (policy == CSL_TGRID_BOOST) {
        // boost::numeric::odeint integrate_const (stepper_tag): t_k = T0 + k*dt,
        // step while t_k + dt - T1 <= eps (less_eq_with_sign).
        size_t k = 0;
        double t = T0;
        while ((t + dt) - T1 <= std::numeric_limits<double>::epsilon()) {
            g.push_back(t);
            ++k;
            t = T0 + double(k) * dt;
        }
        g.push_back(t);
    } else if (policy == CSL_TGRID_GPU || policy == CSL_TGRID_BATCHED) {
        // gpu::ode::integrate (integrator.cuh:470-487) / batched loop
        // (lattice_gpu.cu:557-565): accumulated t += dt.
        const double lim = policy == CSL_TGRID_GPU ? T1 : T1 - 1e-12;
        double t = T0;
        while (t < lim) {
            g.push_back(t);
            t += dt;
        }
        g.push_back(t);
    } else {
        fail(CSL_ERR_ARG, "unknown time-grid policy");
    }
    return g;
}

inline size_t records_for(size_t n_steps, size_t si) { return (n_steps + si - 1) / si + 1; }

// Work split of B trajectories of n steps over G workgroups (B > G):
// workgroup g takes the range [g*W, (g+1)*W) of the trajectory-major step
// order, W = ceil(B*n/G) >= n, as pieces (trajectory, first step, end step,
// slot to load, slot to save): first the head of the trajectory it shares
// with g+1 (saved to slot g), then its whole trajectories, last the rest of
// the trajectory begun by g-1 (loaded from slot g-1).  Rows [G][max_p][5],
// unused pieces -1.
inline std::vector<int> split_schedule(size_t B, size_t G, size_t n, size_t& max_p) {
    const size_t W = B * n, per = (W + G - 1) / G;
    std::vector<std::vector<std::array<int, 5>>> pc(G);
    max_p = 1;
    for (size_t g = 0; g < G; ++g) {
        const size_t a = g * per, e = std::min(W, (g + 1) * per);
        std::vector<std::array<int, 5>> lay;
        for (size_t u = a; u < e;) {
            const size_t j = u / n, s0 = u - j * n, s1 = std::min(n, e - j * n);
            lay.push_back({int(j), int(s0), int(s1), -1, -1});
            u = j * n + s1;
        }
        auto& out = pc[g];
        if (!lay.empty() && lay.back()[1] == 0 && size_t(lay.back()[2]) < n) {
            lay.back()[4] = int(g);
            out.push_back(lay.back());
            lay.pop_back();
        }
        const bool tail = !lay.empty() && lay.front()[1] > 0;
        if (tail) lay.front()[3] = int(g) - 1;
        for (size_t i = tail ? 1 : 0; i < lay.size(); ++i) out.push_back(lay[i]);
        if (tail) out.push_back(lay.front());
        max_p = std::max(max_p, out.size());
    }
    std::vector<int> rows(G * max_p * 5, -1);
    for (size_t g = 0; g < G; ++g)
        for (size_t i = 0; i < pc[g].size(); ++i)
            for (int k = 0; k < 5; ++k) rows[(g * max_p + i) * 5 + k] = pc[g][i][k];
    return rows;
}

inline bool method_is_rk4(const char* m) { return m == nullptr || std::strcmp(m, "rk4") == 0; }

inline void check_finite(const double* v, size_t n, const char* what) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) fail(CSL_ERR_NONFINITE, std::string("non-finite value in ") + what);
}

inline void copy_back(double* host, const double* dev, size_t n, hipStream_t s) {
    if (!host || !n) return;
    hip_check(hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
}

}  // namespace cslh
