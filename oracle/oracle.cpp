/*
 * oracle.cpp — CPU FP64 restatement of the reference's Landau–Lifshitz path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the "port" CPU
 * baseline.  The product library never links, loads or calls this file.
 *
 * Built by oracle/Makefile with  g++ -O2 -fopenmp -ffp-contract=off  (no
 * fast-math, no FMA contraction), so its arithmetic is the plain IEEE order
 * written below.
 *
 * Pinning (each oracle function is checked by tests/ against something that
 * is not a retyping of its own formula):
 *   lattice expansion ....... neighbour counts/partners vs hand-derived
 *                             honeycomb geometry; reverse bond = J^T; CSR
 *                             order vs a brute-force enumeration
 *                             (tests/test_oracle_model.py)
 *   local field / RHS ....... H = dE/dS by central finite differences of
 *                             orc_energy; dense O(N^2) brute force on 1x1x1 ..
 *                             2x2x1 cells; S.dS/dt = 0; sum H.dS/dt = 0
 *                             (tests/test_oracle_physics.py)
 *   energy .................. closed-form FM / Neel energies of the Kitaev
 *                             model; E conservation of undamped RK4 runs
 *   SU(3) contraction ....... f_abc = -(i/4) Tr(l_a [l_b, l_c]) from the
 *                             Gell-Mann matrices (numpy)
 *   drive envelope .......... closed form exp(-(dt/2w)^2) cos(w dt) and a
 *                             single-spin driven precession vs scipy
 *   RK4 step ................ single spin in a static field: the RK4
 *                             amplification of a rotation is known exactly;
 *                             fitted global order 4
 *   integrate_const grid .... step counts vs hand-evaluated cases
 *   pulse segments .......... the reference's own test cases
 *                             (test_pulse_chunking_alignment.cpp)
 *   pump-probe driver ....... M1(tau) equals M0 shifted by tau when the ground
 *                             state is stationary (W1 identity); serial and
 *                             OpenMP tau loops agree
 *   tangent linear .......... central finite differences of the trajectory
 *   Dormand-Prince 5(4) ..... tableau order conditions (exact rationals);
 *                             5th-order global convergence and 4th-order
 *                             embedded error on y' = l y + cos(w t) (closed
 *                             form); tolerance-proportional error vs a
 *                             converged solution; the accept/reject
 *                             sequence of Boost.Odeint's controller is
 *                             parity unpinned (Boost is not in the tree)
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct SiteBond {
    size_t partner;
    std::vector<double> J;   // D*D row-major copy (reference keeps one copy per bond)
    int ucb;                 // unit-cell bond index
    bool transposed;         // reverse bond carries J^T
};

}  // namespace

struct OracleLattice {
    int D = 3;
    size_t n_atoms = 1, d1 = 1, d2 = 1, d3 = 1, N = 0;
    std::vector<double> field;    // [N*D]
    std::vector<double> onsite;   // [N*D*D]
    std::vector<std::vector<SiteBond>> bonds;  // nested tables (lattice.h:785-803)
    // flat mirror (lattice.h:868-903)
    std::vector<size_t> bi_offset, bi_partner;
    std::vector<double> bi_J;
    std::vector<double> frames;    // [n_atoms*D*D]
    std::vector<double> afm;       // [n_atoms]
    // drive state (lattice.h:553-560)
    std::vector<double> field_drive[2];  // [n_atoms*D]
    double t_pulse[2] = {0.0, 0.0};
    double amp = 0.0, width = 1.0, freq = 0.0;
    double alpha = 0.0;
};

namespace {

size_t flatten(const OracleLattice& L, size_t i, size_t j, size_t k, size_t atom) {
    return ((i * L.d2 + j) * L.d3 + k) * L.n_atoms + atom;  // lattice.h:1057-1059
}

int wrap(int c, int n) {  // periodic_boundary (lattice.h:1064-1071): one wrap only
    if (c < 0) return c + n;
    if (c >= n) return c - n;
    return c;
}

void build_flat(OracleLattice& L) {  // build_flat_bilinear_tables (lattice.h:868-903)
    const size_t D2 = size_t(L.D) * L.D;
    L.bi_offset.assign(L.N + 1, 0);
    for (size_t i = 0; i < L.N; ++i) L.bi_offset[i + 1] = L.bi_offset[i] + L.bonds[i].size();
    const size_t nb = L.bi_offset[L.N];
    L.bi_partner.assign(nb, 0);
    L.bi_J.assign(nb * D2, 0.0);
    for (size_t i = 0; i < L.N; ++i)
        for (size_t n = 0; n < L.bonds[i].size(); ++n) {
            const size_t k = L.bi_offset[i] + n;
            L.bi_partner[k] = L.bonds[i][n].partner;
            std::copy(L.bonds[i][n].J.begin(), L.bonds[i][n].J.end(), L.bi_J.begin() + k * D2);
        }
}

// Lattice::get_local_field_flat without twist/trilinear (lattice.h:1653-1711).
void local_field(const OracleLattice& L, const double* x, size_t site, double* H) {
    const int D = L.D;
    for (int d = 0; d < D; ++d) H[d] = -L.field[site * D + d];
    const double* S = x + site * D;
    const double* A = &L.onsite[site * D * D];
    for (int d = 0; d < D; ++d)
        for (int e = 0; e < D; ++e) H[d] += 2.0 * A[d * D + e] * S[e];
    const size_t D2 = size_t(D) * D;
    for (size_t k = L.bi_offset[site]; k < L.bi_offset[site + 1]; ++k) {
        const double* Sp = x + L.bi_partner[k] * D;
        const double* J = &L.bi_J[k * D2];
        for (int d = 0; d < D; ++d) {
            double row = 0.0;
            for (int e = 0; e < D; ++e) row += J[d * D + e] * Sp[e];
            H[d] += row;
        }
    }
}

void drive_factors(const OracleLattice& L, double t, double& f1, double& f2) {
    // lattice_md.cpp:45-56
    const double dt1 = t - L.t_pulse[0];
    const double dt2 = t - L.t_pulse[1];
    f1 = L.amp * std::exp(-std::pow(dt1 / (2.0 * L.width), 2)) * std::cos(L.freq * dt1);
    f2 = L.amp * std::exp(-std::pow(dt2 / (2.0 * L.width), 2)) * std::cos(L.freq * dt2);
}

// cross_prod_SU3_flat (simple_linear_alg.h:78-169), written out per triple.
void su3_cross(const double* H, const double* S, double* out) {
    static const int tri[9][3] = {{0, 1, 2}, {0, 3, 6}, {0, 4, 5}, {1, 3, 5}, {1, 4, 6},
                                  {2, 3, 4}, {2, 5, 6}, {3, 4, 7}, {5, 6, 7}};
    static const double val[9] = {1.0, 0.5, -0.5, 0.5, 0.5, 0.5, -0.5,
                                  0.8660254037844386, 0.8660254037844386};
    double d[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < 9; ++q) {
        const int a = tri[q][0], b = tri[q][1], c = tri[q][2];
        const double v = val[q];
        d[a] += v * (H[b] * S[c] - H[c] * S[b]);
        d[b] += v * (H[c] * S[a] - H[a] * S[c]);
        d[c] += v * (H[a] * S[b] - H[b] * S[a]);
    }
    for (int i = 0; i < 8; ++i) out[i] = d[i];
}

// One site of landau_lifshitz_flat (lattice_md.cpp:70-127).
void rhs_site(const OracleLattice& L, const double* x, double* dxdt, size_t i,
              bool drive, double f1, double f2) {
    const int D = L.D;
    double H[16];
    local_field(L, x, i, H);
    if (drive) {  // apply_drive_field_flat (lattice_md.cpp:178-191)
        const size_t atom = i % L.n_atoms;
        const double* fd0 = &L.field_drive[0][atom * D];
        const double* fd1 = &L.field_drive[1][atom * D];
        for (int d = 0; d < D; ++d) H[d] -= fd0[d] * f1 + fd1[d] * f2;
    }
    const double* S = x + i * D;
    double* o = dxdt + i * D;
    if (D == 3) {
        double dSx = H[1] * S[2] - H[2] * S[1];
        double dSy = H[2] * S[0] - H[0] * S[2];
        double dSz = H[0] * S[1] - H[1] * S[0];
        if (L.alpha > 0.0) {  // lattice_md.cpp:88-98
            const double a = S[1] * dSz - S[2] * dSy;
            const double b = S[2] * dSx - S[0] * dSz;
            const double c = S[0] * dSy - S[1] * dSx;
            const double inv_S = 1.0 / std::sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
            dSx -= L.alpha * inv_S * a;
            dSy -= L.alpha * inv_S * b;
            dSz -= L.alpha * inv_S * c;
        }
        o[0] = dSx; o[1] = dSy; o[2] = dSz;
    } else if (D == 8) {
        su3_cross(H, S, o);
    } else {
        for (int d = 0; d < D; ++d) o[d] = 0.0;
    }
}

void rhs(const OracleLattice& L, const double* x, double* dxdt, double t) {
    const bool drive = (L.amp != 0.0);
    double f1 = 0.0, f2 = 0.0;
    if (drive) drive_factors(L, t, f1, f2);
    const long n = long(L.N);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (n >= 64)
#endif
    for (long i = 0; i < n; ++i) rhs_site(L, x, dxdt, size_t(i), drive, f1, f2);
}

// Boost.Odeint runge_kutta4 (explicit_generic_rk<4,4>): a = {1/2; 0,1/2; 0,0,1},
// b = {1/6,1/3,1/3,1/6}, c = {0,1/2,1/2,1}; stage sums evaluated left to right.
struct RK4Work {
    std::vector<double> k1, k2, k3, k4, tmp;
    explicit RK4Work(size_t n) : k1(n), k2(n), k3(n), k4(n), tmp(n) {}
};

template <class F>
void rk4_step_generic(F&& f, double* x, size_t n, double t, double dt, RK4Work& w) {
    const double b1 = 1.0 / 6.0, b2 = 1.0 / 3.0, b3 = 1.0 / 3.0, b4 = 1.0 / 6.0;
    f(x, w.k1.data(), t);
    for (size_t i = 0; i < n; ++i) w.tmp[i] = x[i] + (0.5 * dt) * w.k1[i];
    f(w.tmp.data(), w.k2.data(), t + 0.5 * dt);
    for (size_t i = 0; i < n; ++i) w.tmp[i] = x[i] + (0.5 * dt) * w.k2[i];
    f(w.tmp.data(), w.k3.data(), t + 0.5 * dt);
    for (size_t i = 0; i < n; ++i) w.tmp[i] = x[i] + (1.0 * dt) * w.k3[i];
    f(w.tmp.data(), w.k4.data(), t + 1.0 * dt);
    for (size_t i = 0; i < n; ++i)
        x[i] = x[i] + (b1 * dt) * w.k1[i] + (b2 * dt) * w.k2[i] + (b3 * dt) * w.k3[i] +
               (b4 * dt) * w.k4[i];
}

// boost::numeric::odeint::detail::less_eq_with_sign for dt > 0.
bool less_eq_with_sign(double t1, double t2) {
    return t1 - t2 <= std::numeric_limits<double>::epsilon();
}

// integrate_const (stepper_tag): observe(t_k) then step while t_k + dt <= end.
template <class Step, class Obs>
size_t integrate_const(Step&& step, double t0, double t1, double dt, Obs&& obs) {
    double t = t0;
    size_t k = 0;
    while (less_eq_with_sign(t + dt, t1)) {
        obs(t);
        step(t, dt);
        ++k;
        t = t0 + double(k) * dt;
    }
    obs(t);
    return k;
}

void mags(const OracleLattice& L, const double* x, double* out) {
    // out = [M_antiferro | M_local | M_global], each D long, all divided by N.
    const int D = L.D;
    double Ml[16] = {0}, Ma[16] = {0}, Mg[16] = {0};
    for (size_t i = 0; i < L.N; ++i)  // compute_magnetizations_from_flat (lattice.h:4962-4975)
        for (int d = 0; d < D; ++d) Ml[d] += x[i * D + d];
    for (size_t i = 0; i < L.N; ++i) {  // compute_magnetization_global_from_flat (:4461-4485)
        const size_t atom = i % L.n_atoms;
        const double* R = &L.frames[atom * D * D];
        for (int mu = 0; mu < D; ++mu)
            for (int nu = 0; nu < D; ++nu) Mg[mu] += R[mu * D + nu] * x[i * D + nu];
    }
    for (size_t i = 0; i < L.N; ++i) {  // compute_magnetization_staggered_from_flat (:4492-4512)
        const size_t atom = i % L.n_atoms;
        const double* R = &L.frames[atom * D * D];
        const double s = L.afm[atom];
        for (int mu = 0; mu < D; ++mu)
            for (int nu = 0; nu < D; ++nu) Ma[mu] += s * R[mu * D + nu] * x[i * D + nu];
    }
    const double n = double(L.N);
    for (int d = 0; d < D; ++d) {
        out[d] = Ma[d] / n;
        out[D + d] = Ml[d] / n;
        out[2 * D + d] = Mg[d] / n;
    }
}

struct Recorder {
    size_t max_rec, n = 0;
    double* times;
    double* mag;
    int MD3;
    void push(double t, const double* m) {
        if (n < max_rec) {
            if (times) times[n] = t;
            if (mag) std::memcpy(mag + n * MD3, m, sizeof(double) * MD3);
        }
        ++n;
    }
};

// Segment list used by the pulse drivers (pulse_chunking.h:82-161).
struct Seg {
    double t0, t1, dt;
};

std::vector<Seg> pulse_segments(double T_start, double T_end, std::vector<double> centers,
                                double window, double T_step, double free_dt_factor) {
    std::vector<Seg> segs;
    if (T_end <= T_start || T_step <= 0.0) return segs;
    auto snap = [&](double t) { return T_start + std::round((t - T_start) / T_step) * T_step; };
    const double T_end_s = T_start + std::floor((T_end - T_start) / T_step + 1e-9) * T_step;
    if (T_end_s <= T_start) return segs;
    std::vector<std::pair<double, double>> win;
    for (double tc : centers) {
        const double w0 = std::max(T_start, snap(tc - window));
        const double w1 = std::min(T_end_s, snap(tc + window));
        if (w0 < w1) win.emplace_back(w0, w1);
    }
    std::sort(win.begin(), win.end());
    std::vector<std::pair<double, double>> merged;
    for (auto& w : win) {
        if (!merged.empty() && w.first <= merged.back().second)
            merged.back().second = std::max(merged.back().second, w.second);
        else
            merged.push_back(w);
    }
    const double free_dt = std::max(T_step, free_dt_factor * T_step);
    double cur = T_start;
    for (auto& w : merged) {
        if (cur < w.first) segs.push_back({cur, w.first, std::min(free_dt, w.first - cur)});
        segs.push_back({w.first, w.second, T_step});
        cur = w.second;
    }
    if (cur < T_end_s) segs.push_back({cur, T_end_s, std::min(free_dt, T_end_s - cur)});
    return segs;
}

// Shared body of single_pulse_drive / double_pulse_drive (CPU branch).
size_t pulse_drive(OracleLattice& L, const double* init, std::vector<double> centers,
                   double T0, double T1, double step, int chunking, size_t max_rec,
                   double* times, double* out_mags) {
    const size_t n = L.N * L.D;
    std::vector<double> x(init, init + n);
    RK4Work w(n);
    Recorder rec{max_rec, 0, times, out_mags, 3 * L.D};
    double last = T0;
    std::vector<double> m(3 * L.D);
    auto obs = [&](double t) {  // lattice_md.cpp:532-554
        if (t - last >= step - 1e-10 || t >= T1 - 1e-10) {
            mags(L, x.data(), m.data());
            rec.push(t, m.data());
            last = t;
        }
    };
    auto f = [&](const double* a, double* b, double t) { rhs(L, a, b, t); };
    auto stp = [&](double t, double dt) { rk4_step_generic(f, x.data(), n, t, dt, w); };
    if (chunking) {
        // kPulseWindowSigmas = 9, kFreeDtFactor = 20 (pulse_chunking.h:173,178)
        auto segs = pulse_segments(T0, T1, centers, 9.0 * L.width, step, 20.0);
        for (auto& s : segs) integrate_const(stp, s.t0, s.t1, step, obs);
    } else {
        integrate_const(stp, T0, T1, step, obs);
    }
    // Reset pulse (lattice_md.cpp:578-583)
    std::fill(L.field_drive[0].begin(), L.field_drive[0].end(), 0.0);
    std::fill(L.field_drive[1].begin(), L.field_drive[1].end(), 0.0);
    L.amp = 0.0;
    return rec.n;
}

void set_pulse(OracleLattice& L, const double* f1, double t1, const double* f2, double t2,
               double amp, double width, double freq) {
    const int D = L.D;
    for (size_t a = 0; a < L.n_atoms; ++a) {
        const double* R = &L.frames[a * D * D];
        for (int mu = 0; mu < D; ++mu) {
            double s1 = 0.0, s2 = 0.0;
            for (int nu = 0; nu < D; ++nu) {
                s1 += R[mu * D + nu] * (f1 ? f1[a * D + nu] : 0.0);
                s2 += R[mu * D + nu] * (f2 ? f2[a * D + nu] : 0.0);
            }
            L.field_drive[0][a * D + mu] = s1;
            L.field_drive[1][a * D + mu] = s2;
        }
    }
    L.t_pulse[0] = t1;
    L.t_pulse[1] = t2;
    L.amp = amp;
    L.width = width;
    L.freq = freq;
}

}  // namespace

// ---- Dormand-Prince 5(4) -------------------------------------------------
//
// Tableau of Dormand & Prince, J. Comput. Appl. Math. 6 (1980) 19-26, the one
// Boost.Odeint's runge_kutta_dopri5 implements (FSAL: the derivative at the
// accepted x_new starts the next step).  The controller restates Boost.Odeint's
// controlled_runge_kutta with default_error_checker (a_x = a_dxdt = 1) and
// default_step_adjuster; Boost is a third-party dependency absent from the
// reference tree, so the accept/reject sequence is parity unpinned against
// Boost and pinned instead by the tableau's order conditions, the 5th-order
// convergence of the propagated solution and the 4th-order embedded error on
// closed-form problems, and convergence to a tight-tolerance solution.
namespace {

struct DP5 {
    static constexpr double c2 = 1.0 / 5.0, c3 = 3.0 / 10.0, c4 = 4.0 / 5.0, c5 = 8.0 / 9.0;
    static constexpr double a21 = 1.0 / 5.0;
    static constexpr double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
    static constexpr double a41 = 44.0 / 45.0, a42 = -56.0 / 15.0, a43 = 32.0 / 9.0;
    static constexpr double a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0,
                            a54 = -212.0 / 729.0;
    static constexpr double a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0,
                            a64 = 49.0 / 176.0, a65 = -5103.0 / 18656.0;
    static constexpr double b1 = 35.0 / 384.0, b3 = 500.0 / 1113.0, b4 = 125.0 / 192.0,
                            b5 = -2187.0 / 6784.0, b6 = 11.0 / 84.0;
    // b - b* (4th-order weights 5179/57600, 7571/16695, 393/640, -92097/339200, 187/2100, 1/40)
    static constexpr double e1 = 35.0 / 384.0 - 5179.0 / 57600.0, e3 = 500.0 / 1113.0 - 7571.0 / 16695.0,
                            e4 = 125.0 / 192.0 - 393.0 / 640.0, e5 = -2187.0 / 6784.0 + 92097.0 / 339200.0,
                            e6 = 11.0 / 84.0 - 187.0 / 2100.0, e7 = -1.0 / 40.0;
};

struct DP5Work {
    std::vector<double> k2, k3, k4, k5, k6, k7, tmp, xnew, xerr;
    explicit DP5Work(size_t n) : k2(n), k3(n), k4(n), k5(n), k6(n), k7(n), tmp(n), xnew(n), xerr(n) {}
};

// One attempt from (x, k1 = f(x, t)): xnew, k7 = f(xnew, t + dt) and the
// embedded error xerr.  Stage sums are evaluated left to right.
template <class F>
void dp5_attempt(F&& f, const double* x, const double* k1, double t, double dt, size_t n, DP5Work& w) {
    using C = DP5;
    for (size_t i = 0; i < n; ++i) w.tmp[i] = x[i] + dt * C::a21 * k1[i];
    f(w.tmp.data(), w.k2.data(), t + C::c2 * dt);
    for (size_t i = 0; i < n; ++i) w.tmp[i] = x[i] + dt * (C::a31 * k1[i] + C::a32 * w.k2[i]);
    f(w.tmp.data(), w.k3.data(), t + C::c3 * dt);
    for (size_t i = 0; i < n; ++i)
        w.tmp[i] = x[i] + dt * (C::a41 * k1[i] + C::a42 * w.k2[i] + C::a43 * w.k3[i]);
    f(w.tmp.data(), w.k4.data(), t + C::c4 * dt);
    for (size_t i = 0; i < n; ++i)
        w.tmp[i] = x[i] + dt * (C::a51 * k1[i] + C::a52 * w.k2[i] + C::a53 * w.k3[i] + C::a54 * w.k4[i]);
    f(w.tmp.data(), w.k5.data(), t + C::c5 * dt);
    for (size_t i = 0; i < n; ++i)
        w.tmp[i] = x[i] + dt * (C::a61 * k1[i] + C::a62 * w.k2[i] + C::a63 * w.k3[i] + C::a64 * w.k4[i] +
                                C::a65 * w.k5[i]);
    f(w.tmp.data(), w.k6.data(), t + dt);
    for (size_t i = 0; i < n; ++i)
        w.xnew[i] = x[i] + dt * (C::b1 * k1[i] + C::b3 * w.k3[i] + C::b4 * w.k4[i] + C::b5 * w.k5[i] +
                                 C::b6 * w.k6[i]);
    f(w.xnew.data(), w.k7.data(), t + dt);
    for (size_t i = 0; i < n; ++i)
        w.xerr[i] = dt * (C::e1 * k1[i] + C::e3 * w.k3[i] + C::e4 * w.k4[i] + C::e5 * w.k5[i] +
                          C::e6 * w.k6[i] + C::e7 * w.k7[i]);
}

// default_error_checker::error: max_i |xerr_i| / (abs + rel (|x_i| + |dt| |dxdt_i|)).
double dp5_error(const double* x, const double* k1, const double* xerr, double dt, double abs_tol,
                 double rel_tol, size_t n) {
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double r = std::abs(xerr[i]) / (abs_tol + rel_tol * (std::abs(x[i]) + std::abs(dt) * std::abs(k1[i])));
        m = std::max(m, r);
    }
    return m;
}

// default_step_adjuster: shrink after a rejection (error order 4), grow after
// an accepted step with err < 1/2 (stepper order 5, at most x5).
double dp5_decrease(double dt, double err) {
    return dt * std::max(0.9 * std::pow(err, -1.0 / 3.0), 0.2);
}
double dp5_increase(double dt, double err) {
    if (err < 0.5) {
        err = std::max(1.0 / 3125.0, err);  // 5^-5
        dt = dt * (0.9 * std::pow(err, -1.0 / 5.0));
    }
    return dt;
}

// boost::numeric::odeint::detail::less_with_sign for dt > 0.
bool less_with_sign(double t1, double t2) { return t2 - t1 > std::numeric_limits<double>::epsilon(); }

constexpr int kMaxRejections = 500;  // failed_step_checker

// Controlled DP5 state: x and its FSAL derivative k1 at time t.
struct DP5Run {
    std::function<void(const double*, double*, double)> f;
    double* x;
    size_t n;
    std::vector<double> k1;
    DP5Work w;
    double abs_tol, rel_tol;
    size_t accepted = 0, rejected = 0;
    DP5Run(std::function<void(const double*, double*, double)> f_, double* x_, size_t n_, double t,
           double a, double r)
        : f(std::move(f_)), x(x_), n(n_), k1(n_), w(n_), abs_tol(a), rel_tol(r) {
        f(x, k1.data(), t);
    }
    // controlled try_step: true when the step was accepted (t, dt advanced)
    bool try_step(double& t, double& dt) {
        dp5_attempt(f, x, k1.data(), t, dt, n, w);
        const double err = dp5_error(x, k1.data(), w.xerr.data(), dt, abs_tol, rel_tol, n);
        if (!(err <= 1.0)) {  // NaN counts as a failure
            dt = dp5_decrease(dt, err);
            ++rejected;
            return false;
        }
        std::memcpy(x, w.xnew.data(), n * sizeof(double));
        std::memcpy(k1.data(), w.k7.data(), n * sizeof(double));
        t += dt;
        dt = dp5_increase(dt, err);
        ++accepted;
        return true;
    }
    // a new copy of the controlled stepper: the first try_step takes the
    // derivative at (x, t) again
    void restart(double t) { f(x, k1.data(), t); }
    // integrate_adaptive (controlled_stepper_tag) from t to t_end with the
    // step size dt (updated in place, as Boost's reference argument); obs(t)
    // at the start and after every accepted step when with_obs.  Returns
    // false after kMaxRejections rejections in a row.
    template <class Obs>
    bool adaptive(double t, double t_end, double& dt, Obs&& obs, bool with_obs) {
        while (less_with_sign(t, t_end)) {
            if (with_obs) obs(t);
            if (less_with_sign(t_end, t + dt)) dt = t_end - t;
            int fails = 0;
            while (!try_step(t, dt))
                if (++fails >= kMaxRejections) return false;
        }
        if (with_obs) obs(t);
        return true;
    }
};

// The three ways the reference drives the Dormand-Prince stepper
// (integrate_ode_system, lattice.h:4072-4085) over a state of n doubles with
// right-hand side f: control 0 fixed steps on the integrate_const grid,
// 1 integrate_adaptive, 2 integrate_const with the controlled stepper (one
// call per segment row (t0, t1) of segs when n_seg > 0).  obs(t) is called
// for every observer call; false after kMaxRejections rejections in a row.
template <class F, class Obs>
bool dp5_controls(F&& f, double* state, size_t n, double T0, double T1, double dt, int control, double abs_tol,
                  double rel_tol, const double* segs, size_t n_seg, Obs&& obs, size_t& acc, size_t& rej) {
    bool ok = true;
    if (control == 0) {
        // fixed steps on the integrate_const grid, no error control
        DP5Work w(n);
        std::vector<double> k1(n);
        auto stp = [&](double t, double h) {
            f(state, k1.data(), t);
            dp5_attempt(f, state, k1.data(), t, h, n, w);
            std::memcpy(state, w.xnew.data(), n * sizeof(double));
            ++acc;
        };
        integrate_const(stp, T0, T1, dt, obs);
    } else {
        DP5Run run(f, state, n, T0, abs_tol, rel_tol);
        if (control == 1) {
            double h = dt;
            ok = run.adaptive(T0, T1, h, obs, true);  // integrate_adaptive
        } else {
            // integrate_const with a controlled stepper (Boost's
            // controlled_stepper_tag loop): observe at time_k = s0 + k dt,
            // integrate_adaptive from time_k to time_k + dt with a copy of
            // the stepper made at the call (so the derivative is taken again
            // at time_k) and the step size carried over from the previous
            // interval; the step hint is dt at the start of each call.
            auto integrate_const_ctl = [&](double s0, double s1) {
                double time = s0, h = dt;
                size_t k = 0;
                while (ok && less_eq_with_sign(time + dt, s1)) {
                    obs(time);
                    run.restart(time);
                    ok = run.adaptive(time, time + dt, h, obs, false);
                    ++k;
                    time = s0 + double(k) * dt;
                }
                if (ok) obs(time);
            };
            if (n_seg == 0) {
                integrate_const_ctl(T0, T1);
            } else {
                for (size_t j = 0; ok && j < n_seg; ++j) integrate_const_ctl(segs[2 * j], segs[2 * j + 1]);
            }
        }
        acc = run.accepted;
        rej = run.rejected;
    }
    return ok;
}

// ---- the reference GPU stepper gpu::ode::step (integrator.cuh:91-458) ----
// Restated operation by operation on host arrays: every buffer update is one
// of the reference's element-wise kernels (gpu_common_helpers.cu:9-120), in
// the order the reference launches them.  f(in, out, t) is the RHS.

struct GpuStepWork {
    std::vector<double> k[11], tmp;
    explicit GpuStepWork(size_t n) : tmp(n) {
        for (auto& v : k) v.assign(n, 0.0);
    }
};

// update_arrays_kernel: out = a1*in1 + a2*in2
void upd2(double* out, double a1, const double* in1, double a2, const double* in2, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = a1 * in1[i] + a2 * in2[i];
}
// update_arrays_three_kernel: out = a1*in1 + a2*in2 + a3*in3
void upd3(double* out, double a1, const double* in1, double a2, const double* in2, double a3, const double* in3,
          size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = a1 * in1[i] + a2 * in2[i] + a3 * in3[i];
}
// rk_stage_update_m_kernel: out = x + dt * (sum a_j k_j)
void stage_upd(double* out, const double* x, double dt, std::initializer_list<std::pair<double, const double*>> t,
               size_t n) {
    for (size_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (const auto& p : t) acc += p.first * p.second[i];
        out[i] = x[i] + dt * acc;
    }
}

// returns false for an unknown method name (the caller falls back to
// ssprk53 as integrator.cuh:454-458 does)
template <class F>
bool gpu_step(F&& f, double* x, size_t n, double t, double dt, const char* method, GpuStepWork& w) {
    const std::string m = method ? method : "rk4";
    double* k1 = w.k[0].data();
    double* k2 = w.k[1].data();
    double* k3 = w.k[2].data();
    double* k4 = w.k[3].data();
    double* k5 = w.k[4].data();
    double* k6 = w.k[5].data();
    double* tmp = w.tmp.data();
    if (m == "euler") {  // integrator.cuh:91-99
        f(x, k1, t);
        upd2(x, 1.0, x, dt, k1, n);
    } else if (m == "rk2" || m == "midpoint") {  // :100-114
        f(x, k1, t);
        upd2(tmp, 1.0, x, 0.5 * dt, k1, n);
        f(tmp, k2, t + 0.5 * dt);
        upd2(x, 1.0, x, dt, k2, n);
    } else if (m == "rk4") {  // :115-139 with rk4_final_update_kernel
        f(x, k1, t);
        upd2(tmp, 1.0, x, 0.5 * dt, k1, n);
        f(tmp, k2, t + 0.5 * dt);
        upd2(tmp, 1.0, x, 0.5 * dt, k2, n);
        f(tmp, k3, t + 0.5 * dt);
        upd2(tmp, 1.0, x, dt, k3, n);
        f(tmp, k4, t + dt);
        const double d6 = dt / 6.0;
        for (size_t i = 0; i < n; ++i) x[i] += d6 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
    } else if (m == "rk5" || m == "rkck54" || m == "dopri5" || m == "rk54" || m == "rkf54") {
        // :141-185 (Cash-Karp) and :186-229 (Dormand-Prince), fixed step;
        // rk54 / rkf54 reuse dopri5 (:450-452)
        const bool ck = (m == "rk5" || m == "rkck54");
        double a21, a31, a32, a41, a42, a43, a51, a52, a53, a54, a61, a62, a63, a64, a65;
        double c2, c3, c4, c5, c6, b1, b3, b4, b5, b6;
        if (ck) {
            a21 = 1.0 / 5.0; a31 = 3.0 / 40.0; a32 = 9.0 / 40.0;
            a41 = 3.0 / 10.0; a42 = -9.0 / 10.0; a43 = 6.0 / 5.0;
            a51 = -11.0 / 54.0; a52 = 5.0 / 2.0; a53 = -70.0 / 27.0; a54 = 35.0 / 27.0;
            a61 = 1631.0 / 55296.0; a62 = 175.0 / 512.0; a63 = 575.0 / 13824.0;
            a64 = 44275.0 / 110592.0; a65 = 253.0 / 4096.0;
            c2 = 1.0 / 5.0; c3 = 3.0 / 10.0; c4 = 3.0 / 5.0; c5 = 1.0; c6 = 7.0 / 8.0;
            b1 = 37.0 / 378.0; b3 = 250.0 / 621.0; b4 = 125.0 / 594.0; b5 = 0.0; b6 = 512.0 / 1771.0;
        } else {
            a21 = 1.0 / 5.0; a31 = 3.0 / 40.0; a32 = 9.0 / 40.0;
            a41 = 44.0 / 45.0; a42 = -56.0 / 15.0; a43 = 32.0 / 9.0;
            a51 = 19372.0 / 6561.0; a52 = -25360.0 / 2187.0; a53 = 64448.0 / 6561.0; a54 = -212.0 / 729.0;
            a61 = 9017.0 / 3168.0; a62 = -355.0 / 33.0; a63 = 46732.0 / 5247.0; a64 = 49.0 / 176.0;
            a65 = -5103.0 / 18656.0;
            c2 = 1.0 / 5.0; c3 = 3.0 / 10.0; c4 = 4.0 / 5.0; c5 = 8.0 / 9.0; c6 = 1.0;
            b1 = 35.0 / 384.0; b3 = 500.0 / 1113.0; b4 = 125.0 / 192.0; b5 = -2187.0 / 6784.0; b6 = 11.0 / 84.0;
        }
        f(x, k1, t);
        upd2(tmp, 1.0, x, a21 * dt, k1, n);
        f(tmp, k2, t + c2 * dt);
        stage_upd(tmp, x, dt, {{a31, k1}, {a32, k2}}, n);
        f(tmp, k3, t + c3 * dt);
        stage_upd(tmp, x, dt, {{a41, k1}, {a42, k2}, {a43, k3}}, n);
        f(tmp, k4, t + c4 * dt);
        stage_upd(tmp, x, dt, {{a51, k1}, {a52, k2}, {a53, k3}, {a54, k4}}, n);
        f(tmp, k5, t + c5 * dt);
        stage_upd(tmp, x, dt, {{a61, k1}, {a62, k2}, {a63, k3}, {a64, k4}, {a65, k5}}, n);
        f(tmp, k6, t + c6 * dt);
        // dopri5_final_update_kernel; Cash-Karp passes k6 in the k5 slot with b5 = 0
        const double* k5s = ck ? k6 : k5;
        for (size_t i = 0; i < n; ++i)
            x[i] += dt * (b1 * k1[i] + b3 * k3[i] + b4 * k4[i] + b5 * k5s[i] + b6 * k6[i]);
    } else if (m == "rk78" || m == "rkf78") {  // :230-370
        double* K[11];
        for (int j = 0; j < 11; ++j) K[j] = w.k[j].data();
        // tmp = x + c*dt*K_a, then tmp += c*dt*K_b ... as the reference's
        // chains of update_arrays_kernel calls
        auto chain = [&](std::initializer_list<std::pair<int, double>> terms) {
            bool first = true;
            for (const auto& p : terms) {
                upd2(tmp, 1.0, first ? x : tmp, p.second * dt, K[p.first], n);
                first = false;
            }
        };
        const double c[11] = {0.0, 2.0 / 27.0, 1.0 / 9.0, 1.0 / 6.0, 5.0 / 12.0, 1.0 / 2.0,
                              5.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0, 1.0 / 3.0, 1.0};
        f(x, K[0], t);
        chain({{0, 2.0 / 27.0}});
        f(tmp, K[1], t + c[1] * dt);
        chain({{0, 1.0 / 36.0}, {1, 1.0 / 12.0}});
        f(tmp, K[2], t + c[2] * dt);
        chain({{0, 1.0 / 24.0}, {2, 1.0 / 8.0}});
        f(tmp, K[3], t + c[3] * dt);
        chain({{0, 5.0 / 12.0}, {3, -25.0 / 16.0}, {2, 25.0 / 16.0}});
        f(tmp, K[4], t + c[4] * dt);
        chain({{0, 1.0 / 20.0}, {3, 1.0 / 4.0}, {4, 1.0 / 5.0}});
        f(tmp, K[5], t + c[5] * dt);
        chain({{0, -25.0 / 108.0}, {3, 125.0 / 108.0}, {4, -65.0 / 27.0}, {5, 125.0 / 54.0}});
        f(tmp, K[6], t + c[6] * dt);
        chain({{0, 31.0 / 300.0}, {4, 61.0 / 225.0}, {5, -2.0 / 9.0}, {6, 13.0 / 900.0}});
        f(tmp, K[7], t + c[7] * dt);
        chain({{0, 2.0}, {3, -53.0 / 6.0}, {4, 704.0 / 45.0}, {5, -107.0 / 9.0}, {6, 67.0 / 90.0}, {7, 3.0}});
        f(tmp, K[8], t + c[8] * dt);
        chain({{0, -91.0 / 108.0}, {3, 23.0 / 108.0}, {4, -976.0 / 135.0}, {5, 311.0 / 54.0},
               {6, -19.0 / 60.0}, {7, 17.0 / 6.0}, {8, -1.0 / 12.0}});
        f(tmp, K[9], t + c[9] * dt);
        chain({{0, 2383.0 / 4100.0}, {3, -341.0 / 164.0}, {4, 4496.0 / 1025.0}, {5, -301.0 / 82.0},
               {6, 2133.0 / 4100.0}, {7, 45.0 / 82.0}, {8, 45.0 / 164.0}, {9, 18.0 / 41.0}});
        f(tmp, K[10], t + c[10] * dt);
        const std::pair<int, double> fin[7] = {{0, 41.0 / 840.0}, {5, 34.0 / 105.0}, {6, 9.0 / 35.0},
                                               {7, 9.0 / 35.0},   {8, 9.0 / 280.0},  {9, 9.0 / 280.0},
                                               {10, 41.0 / 840.0}};
        for (const auto& p : fin) upd2(x, 1.0, x, p.second * dt, K[p.first], n);
    } else if (m == "bulirsch_stoer" || m == "bs") {  // :371-405
        double* kk = k1;
        double* y_prev = k2;
        double* y_mid = k3;
        double* y_next = k4;
        const int n_sub = 4;
        const double hs = dt / n_sub;
        std::memcpy(y_prev, x, n * sizeof(double));
        f(x, kk, t);
        upd2(y_mid, 1.0, x, hs, kk, n);
        for (int i = 1; i < n_sub; ++i) {
            f(y_mid, kk, t + i * hs);
            upd2(y_next, 1.0, y_prev, 2.0 * hs, kk, n);
            std::memcpy(y_prev, y_mid, n * sizeof(double));
            std::memcpy(y_mid, y_next, n * sizeof(double));
        }
        f(y_mid, kk, t + dt);
        upd2(y_next, 0.5, y_mid, 0.5, y_prev, n);
        upd2(x, 1.0, y_next, 0.5 * hs, kk, n);
    } else if (m == "ssprk53") {  // :406-449
        constexpr double a30 = 0.355909775063327, a32 = 0.644090224936674;
        constexpr double a40 = 0.367933791638137, a43 = 0.632066208361863;
        constexpr double a52 = 0.237593836598569, a54 = 0.762406163401431;
        constexpr double b10 = 0.377268915331368, b21 = 0.377268915331368;
        constexpr double b32 = 0.242995220537396, b43 = 0.238458932846290;
        constexpr double b54 = 0.287632146308408;
        constexpr double c1 = 0.377268915331368, c2 = 0.754537830662736;
        constexpr double c3 = 0.728985661612188, c4 = 0.699226135931670;
        double* kk = k1;
        double* u = k2;
        f(x, kk, t);
        upd2(tmp, 1.0, x, b10 * dt, kk, n);
        f(tmp, kk, t + c1 * dt);
        upd2(u, 1.0, tmp, b21 * dt, kk, n);
        f(u, kk, t + c2 * dt);
        upd3(tmp, a30, x, a32, u, b32 * dt, kk, n);
        f(tmp, kk, t + c3 * dt);
        upd3(tmp, a40, x, a43, tmp, b43 * dt, kk, n);
        f(tmp, kk, t + c4 * dt);
        upd3(x, a52, u, a54, tmp, b54 * dt, kk, n);
    } else {
        return false;
    }
    return true;
}

// gpu::ode::integrate (integrator.cuh:470-487): t += dt while t < T_end, the
// observer every save_interval steps and once after the loop.
template <class F, class Obs>
void gpu_integrate(F&& f, double* x, size_t n, double T0, double T1, double dt, size_t save_interval,
                   const char* method, Obs&& obs) {
    GpuStepWork w(n);
    const char* m = method;
    double t = T0;
    size_t step = 0;
    while (t < T1) {
        if (step % save_interval == 0) obs(t);
        if (!gpu_step(f, x, n, t, dt, m, w)) {
            m = "ssprk53";
            gpu_step(f, x, n, t, dt, m, w);
        }
        t += dt;
        ++step;
    }
    obs(t);
}

}  // namespace

extern "C" {

OracleLattice* orc_lattice_new(int spin_dim, int n_atoms, int d1, int d2, int d3, int n_bonds,
                               const int* bond_src, const int* bond_partner, const int* bond_offset,
                               const double* bond_J, const double* field_atom,
                               const double* onsite_atom, const double* frames,
                               const double* afm_signs) {
    // a single periodic wrap (lattice.h:750-756) is only valid for offsets in [-L, L]
    for (int b = 0; b < n_bonds; ++b)
        for (int a = 0; a < 3; ++a) {
            const int L = a == 0 ? d1 : (a == 1 ? d2 : d3);
            if (bond_offset[3 * b + a] < -L || bond_offset[3 * b + a] > L) return nullptr;
        }
    auto* L = new OracleLattice();
    L->D = spin_dim;
    L->n_atoms = n_atoms;
    L->d1 = d1; L->d2 = d2; L->d3 = d3;
    L->N = size_t(n_atoms) * d1 * d2 * d3;
    const int D = spin_dim;
    const size_t D2 = size_t(D) * D;
    L->field.assign(L->N * D, 0.0);
    L->onsite.assign(L->N * D2, 0.0);
    L->bonds.assign(L->N, {});
    // Third pass of the Lattice constructor (lattice.h:744-807): sites in
    // flatten order, unit-cell bonds of the atom in insertion order; forward
    // J at the source, J^T at the partner.
    for (int i = 0; i < d1; ++i)
        for (int j = 0; j < d2; ++j)
            for (int k = 0; k < d3; ++k)
                for (int atom = 0; atom < n_atoms; ++atom) {
                    const size_t s = flatten(*L, i, j, k, atom);
                    for (int d = 0; d < D; ++d) L->field[s * D + d] = field_atom[atom * D + d];
                    for (size_t q = 0; q < D2; ++q) L->onsite[s * D2 + q] = onsite_atom[atom * D2 + q];
                    for (int b = 0; b < n_bonds; ++b) {
                        if (bond_src[b] != atom) continue;
                        const int pi = wrap(i + bond_offset[3 * b + 0], d1);
                        const int pj = wrap(j + bond_offset[3 * b + 1], d2);
                        const int pk = wrap(k + bond_offset[3 * b + 2], d3);
                        const size_t p = flatten(*L, pi, pj, pk, bond_partner[b]);
                        SiteBond fw{p, std::vector<double>(bond_J + b * D2, bond_J + (b + 1) * D2), b, false};
                        SiteBond rv{s, std::vector<double>(D2), b, true};
                        for (int r = 0; r < D; ++r)
                            for (int c = 0; c < D; ++c) rv.J[r * D + c] = bond_J[b * D2 + c * D + r];
                        L->bonds[s].push_back(std::move(fw));
                        L->bonds[p].push_back(std::move(rv));
                    }
                }
    build_flat(*L);
    L->frames.assign(frames, frames + n_atoms * D2);
    L->afm.assign(afm_signs, afm_signs + n_atoms);
    L->field_drive[0].assign(n_atoms * D, 0.0);
    L->field_drive[1].assign(n_atoms * D, 0.0);
    return L;
}

OracleLattice* orc_lattice_from_csr(int spin_dim, int n_atoms, int d1, int d2, int d3,
                                    const size_t* offset, const size_t* partner, const double* J,
                                    const double* field_site, const double* onsite_site,
                                    const double* frames, const double* afm_signs) {
    auto* L = new OracleLattice();
    L->D = spin_dim;
    L->n_atoms = n_atoms;
    L->d1 = d1; L->d2 = d2; L->d3 = d3;
    L->N = size_t(n_atoms) * d1 * d2 * d3;
    const size_t D2 = size_t(spin_dim) * spin_dim;
    L->field.assign(field_site, field_site + L->N * spin_dim);
    L->onsite.assign(onsite_site, onsite_site + L->N * D2);
    L->bonds.assign(L->N, {});
    for (size_t i = 0; i < L->N; ++i)
        for (size_t k = offset[i]; k < offset[i + 1]; ++k)
            L->bonds[i].push_back({partner[k], std::vector<double>(J + k * D2, J + (k + 1) * D2), -1, false});
    build_flat(*L);
    L->frames.assign(frames, frames + n_atoms * D2);
    L->afm.assign(afm_signs, afm_signs + n_atoms);
    L->field_drive[0].assign(n_atoms * spin_dim, 0.0);
    L->field_drive[1].assign(n_atoms * spin_dim, 0.0);
    return L;
}

void orc_lattice_free(OracleLattice* L) { delete L; }
size_t orc_lattice_size(const OracleLattice* L) { return L->N; }
size_t orc_total_bonds(const OracleLattice* L) { return L->bi_offset[L->N]; }

void orc_flat_tables(const OracleLattice* L, size_t* offset, size_t* partner, double* J) {
    std::copy(L->bi_offset.begin(), L->bi_offset.end(), offset);
    std::copy(L->bi_partner.begin(), L->bi_partner.end(), partner);
    std::copy(L->bi_J.begin(), L->bi_J.end(), J);
}

void orc_set_alpha(OracleLattice* L, double alpha) { L->alpha = alpha; }

void orc_set_pulse(OracleLattice* L, const double* f1, double t1, const double* f2, double t2,
                   double amp, double width, double freq) {
    set_pulse(*L, f1, t1, f2, t2, amp, width, freq);
}

void orc_clear_pulse(OracleLattice* L) {
    std::fill(L->field_drive[0].begin(), L->field_drive[0].end(), 0.0);
    std::fill(L->field_drive[1].begin(), L->field_drive[1].end(), 0.0);
    L->amp = 0.0;
}

void orc_local_field(const OracleLattice* L, const double* state, size_t site, double* H) {
    local_field(*L, state, site, H);
}

void orc_rhs(const OracleLattice* L, const double* state, double* dsdt, double t) {
    rhs(*L, state, dsdt, t);
}

double orc_energy(const OracleLattice* L, const double* x) {  // lattice.h:1485-1529
    const int D = L->D;
    double E = 0.0;
    for (size_t i = 0; i < L->N; ++i) {
        const double* Si = x + i * D;
        for (int d = 0; d < D; ++d) E -= L->field[i * D + d] * Si[d];
        const double* A = &L->onsite[i * D * D];
        for (int d = 0; d < D; ++d)
            for (int e = 0; e < D; ++e) E += Si[d] * A[d * D + e] * Si[e];
        for (const auto& b : L->bonds[i]) {
            if (b.partner > i) {
                const double* Sj = x + b.partner * D;
                for (int d = 0; d < D; ++d)
                    for (int e = 0; e < D; ++e) E += Si[d] * b.J[d * D + e] * Sj[e];
            }
        }
    }
    return E;
}

void orc_mags(const OracleLattice* L, const double* state, double* out) { mags(*L, state, out); }

void orc_rk4_step(const OracleLattice* L, double* state, double t, double dt) {
    const size_t n = L->N * L->D;
    RK4Work w(n);
    auto f = [&](const double* a, double* b, double tt) { rhs(*L, a, b, tt); };
    rk4_step_generic(f, state, n, t, dt, w);
}

size_t orc_boost_num_steps(double t0, double t1, double dt) {
    size_t k = 0;
    double t = t0;
    while (less_eq_with_sign(t + dt, t1)) {
        ++k;
        t = t0 + double(k) * dt;
    }
    return k;
}

size_t orc_integrate_md(const OracleLattice* L, double* state, double T0, double T1, double dt,
                        size_t save_interval, size_t max_rec, double* times, double* out_mags,
                        double* energy) {
    const size_t n = L->N * L->D;
    RK4Work w(n);
    size_t step_count = 0, nrec = 0;
    std::vector<double> m(3 * L->D);
    auto obs = [&](double t) {  // lattice_md.cpp:383-436 (renorm off, HDF5 omitted)
        if (step_count % save_interval == 0) {
            if (nrec < max_rec) {
                mags(*L, state, m.data());
                if (times) times[nrec] = t;
                if (out_mags) std::memcpy(out_mags + nrec * 3 * L->D, m.data(), sizeof(double) * 3 * L->D);
                if (energy) energy[nrec] = orc_energy(L, state) / double(L->N);
            }
            ++nrec;
        }
        ++step_count;
    };
    auto f = [&](const double* a, double* b, double tt) { rhs(*L, a, b, tt); };
    auto stp = [&](double t, double h) { rk4_step_generic(f, state, n, t, h, w); };
    integrate_const(stp, T0, T1, dt, obs);
    return nrec;
}

int orc_integrate_gpu(const OracleLattice* L, double* state, double T0, double T1, double dt,
                      size_t save_interval, const char* method, size_t max_rec, size_t* n_rec, double* times,
                      double* states, double* out_mags, double* energy) {
    const size_t n = L->N * L->D;
    GpuStepWork probe(0);
    double dummy = 0.0;
    const bool known = gpu_step([](const double*, double*, double) {}, &dummy, 0, 0.0, 1.0, method, probe);
    size_t nrec = 0;
    std::vector<double> m(3 * L->D);
    auto obs = [&](double t) {
        if (nrec < max_rec) {
            if (times) times[nrec] = t;
            if (states) std::memcpy(states + nrec * n, state, n * sizeof(double));
            if (out_mags) {
                mags(*L, state, m.data());
                std::memcpy(out_mags + nrec * 3 * L->D, m.data(), sizeof(double) * 3 * L->D);
            }
            if (energy) energy[nrec] = orc_energy(L, state) / double(L->N);
        }
        ++nrec;
    };
    auto f = [&](const double* a, double* b, double tt) { rhs(*L, a, b, tt); };
    gpu_integrate(f, state, n, T0, T1, dt, save_interval, method, obs);
    if (n_rec) *n_rec = nrec;
    return known ? 1 : 0;
}

long orc_integrate_dopri5(const OracleLattice* L, double* state, double T0, double T1, double dt,
                          int control, double abs_tol, double rel_tol, size_t save_interval,
                          const double* segs, size_t n_seg, size_t max_rec, double* times,
                          double* out_mags, double* energy, size_t* n_accepted, size_t* n_rejected) {
    const size_t n = L->N * L->D;
    size_t step_count = 0, nrec = 0;
    std::vector<double> m(3 * L->D);
    auto obs = [&](double t) {  // MD observer (lattice_md.cpp:383-436)
        if (step_count % save_interval == 0) {
            if (nrec < max_rec) {
                mags(*L, state, m.data());
                if (times) times[nrec] = t;
                if (out_mags) std::memcpy(out_mags + nrec * 3 * L->D, m.data(), sizeof(double) * 3 * L->D);
                if (energy) energy[nrec] = orc_energy(L, state) / double(L->N);
            }
            ++nrec;
        }
        ++step_count;
    };
    auto f = [L](const double* a, double* b, double tt) { rhs(*L, a, b, tt); };
    size_t acc = 0, rej = 0;
    const bool ok = dp5_controls(f, state, n, T0, T1, dt, control, abs_tol, rel_tol, segs, n_seg, obs, acc, rej);
    if (n_accepted) *n_accepted = acc;
    if (n_rejected) *n_rejected = rej;
    return ok ? long(nrec) : -1;
}

/* Scalar test problem for the Dormand-Prince pinning tests:
 * y' = lambda*y + cos(omega*t), integrated with the same stepper code. */
long orc_dopri5_scalar(double lambda, double omega, double y0, double T0, double T1, double dt,
                       int control, double abs_tol, double rel_tol, double* y_end, size_t* n_accepted,
                       size_t* n_rejected, double* err_first_step) {
    double y = y0;
    auto f = [lambda, omega](const double* a, double* b, double tt) { b[0] = lambda * a[0] + std::cos(omega * tt); };
    if (err_first_step) {  // embedded error estimate of one step of size dt from (T0, y0)
        DP5Work w(1);
        double k1;
        f(&y, &k1, T0);
        dp5_attempt(f, &y, &k1, T0, dt, 1, w);
        err_first_step[0] = w.xerr[0];
        err_first_step[1] = w.xnew[0];
    }
    size_t acc = 0, rej = 0;
    bool ok = true;
    if (control == 0) {
        DP5Work w(1);
        double k1;
        auto stp = [&](double t, double h) {
            f(&y, &k1, t);
            dp5_attempt(f, &y, &k1, t, h, 1, w);
            y = w.xnew[0];
            ++acc;
        };
        integrate_const(stp, T0, T1, dt, [](double) {});
    } else {
        DP5Run run(f, &y, 1, T0, abs_tol, rel_tol);
        double h = dt;
        ok = run.adaptive(T0, T1, h, [](double) {}, false);
        acc = run.accepted;
        rej = run.rejected;
    }
    *y_end = y;
    if (n_accepted) *n_accepted = acc;
    if (n_rejected) *n_rejected = rej;
    return ok ? 0 : -1;
}

size_t orc_single_pulse_drive(OracleLattice* L, const double* init, const double* field_in,
                              double t_B, double amp, double width, double freq, double T0,
                              double T1, double step, int chunking, size_t max_rec,
                              double* times, double* out_mags) {
    set_pulse(*L, field_in, t_B, nullptr, 0.0, amp, width, freq);  // lattice_md.cpp:512-513
    return pulse_drive(*L, init, {t_B}, T0, T1, step, chunking, max_rec, times, out_mags);
}

size_t orc_double_pulse_drive(OracleLattice* L, const double* init, const double* f1, double tB1,
                              const double* f2, double tB2, double amp, double width, double freq,
                              double T0, double T1, double step, int chunking, size_t max_rec,
                              double* times, double* out_mags) {
    set_pulse(*L, f1, tB1, f2, tB2, amp, width, freq);  // lattice_md.cpp:611-612
    return pulse_drive(*L, init, {tB1, tB2}, T0, T1, step, chunking, max_rec, times, out_mags);
}

double orc_max_dsdt_no_drive(const OracleLattice* L, const double* state) {
    OracleLattice tmp = *L;
    tmp.amp = 0.0;
    std::vector<double> d(L->N * L->D);
    rhs(tmp, state, d.data(), 0.0);
    double mx = 0.0;
    for (double v : d) mx = std::max(mx, std::abs(v));
    return mx;
}

int orc_pump_probe(const OracleLattice* Lc, const double* ground, const double* field_in,
                   double amp, double width, double freq, double tau_start, double tau_end,
                   double tau_step, double T0, double T1, double T_step, int reuse_m0,
                   double stat_tol, int n_outer, int chunking, size_t max_rec, size_t max_tau,
                   size_t* n_rec, double* times, double* M0, double* M1, double* M01,
                   double* tau_values, int* used_w1) {
    OracleLattice L = *Lc;
    const int MD3 = 3 * L.D;
    // M0 = single pulse at t = 0 (lattice_md.cpp:730)
    const size_t nr = orc_single_pulse_drive(&L, ground, field_in, 0.0, amp, width, freq, T0, T1,
                                             T_step, chunking, max_rec, times, M0);
    *n_rec = nr;
    // M_ground_arr (lattice_md.cpp:737-742): staggered uses alternating signs
    std::vector<double> mg(MD3);
    {
        const int D = L.D;
        for (int d = 0; d < MD3; ++d) mg[d] = 0.0;
        for (size_t i = 0; i < L.N; ++i) {
            const double sgn = std::pow(-1.0, double(i));  // magnetization_local_antiferro
            for (int d = 0; d < D; ++d) {
                mg[d] += ground[i * D + d] * sgn;
                mg[D + d] += ground[i * D + d];
            }
            const size_t atom = i % L.n_atoms;
            const double* R = &L.frames[atom * D * D];
            for (int mu = 0; mu < D; ++mu)
                for (int nu = 0; nu < D; ++nu) mg[2 * D + mu] += R[mu * D + nu] * ground[i * D + nu];
        }
        for (int d = 0; d < MD3; ++d) mg[d] /= double(L.N);
    }
    bool can_reuse = false;
    if (reuse_m0) can_reuse = orc_max_dsdt_no_drive(&L, ground) <= stat_tol;
    if (used_w1) *used_w1 = can_reuse ? 1 : 0;
    const int tau_steps = int(std::abs((tau_end - tau_start) / tau_step)) + 1;  // :768
    if (size_t(tau_steps) > max_tau) return -tau_steps;

    auto synth_M1 = [&](double tau, double* out) {  // synthesize_M1_from_M0 (:283-317)
        if (nr == 0) return;
        const double Ts = times[0];
        for (size_t q = 0; q < nr && q < max_rec; ++q) {
            const double ti = times[q];
            const double* src;
            if (ti < tau + Ts) {
                src = mg.data();
            } else {
                long idx = std::lround((ti - tau - Ts) / T_step);
                if (idx < 0) idx = 0;
                if (idx >= long(nr)) idx = long(nr) - 1;
                src = M0 + size_t(idx) * MD3;
            }
            std::memcpy(out + q * MD3, src, sizeof(double) * MD3);
        }
    };

    if (n_outer <= 1) {
        double cur = tau_start;  // serial path accumulates tau (:825-853)
        for (int i = 0; i < tau_steps; ++i) {
            tau_values[i] = cur;
            double* m1 = M1 + size_t(i) * max_rec * MD3;
            double* m01 = M01 + size_t(i) * max_rec * MD3;
            std::vector<double> tt(max_rec);
            if (can_reuse) synth_M1(cur, m1);
            else orc_single_pulse_drive(&L, ground, field_in, cur, amp, width, freq, T0, T1, T_step,
                                        chunking, max_rec, tt.data(), m1);
            orc_double_pulse_drive(&L, ground, field_in, 0.0, field_in, cur, amp, width, freq, T0,
                                   T1, T_step, chunking, max_rec, tt.data(), m01);
            cur += tau_step;
        }
    } else {
#ifdef _OPENMP
        const int saved = omp_get_max_active_levels();
        omp_set_max_active_levels(1);
#pragma omp parallel num_threads(n_outer)
#endif
        {
            OracleLattice local = L;  // per-thread deep clone (:863-868)
            std::vector<double> tt(max_rec);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int i = 0; i < tau_steps; ++i) {
                const double tau = tau_start + i * tau_step;
                tau_values[i] = tau;
                double* m1 = M1 + size_t(i) * max_rec * MD3;
                double* m01 = M01 + size_t(i) * max_rec * MD3;
                if (can_reuse) synth_M1(tau, m1);
                else orc_single_pulse_drive(&local, ground, field_in, tau, amp, width, freq, T0, T1,
                                            T_step, chunking, max_rec, tt.data(), m1);
                orc_double_pulse_drive(&local, ground, field_in, 0.0, field_in, tau, amp, width,
                                       freq, T0, T1, T_step, chunking, max_rec, tt.data(), m01);
            }
        }
#ifdef _OPENMP
        omp_set_max_active_levels(saved);
#endif
    }
    return tau_steps;
}

size_t orc_build_pulse_segments(double T0, double T1, const double* centers, size_t n_centers,
                                double window, double T_step, double free_dt_factor,
                                size_t max_seg, double* out) {
    auto segs = pulse_segments(T0, T1, std::vector<double>(centers, centers + n_centers), window,
                               T_step, free_dt_factor);
    for (size_t i = 0; i < segs.size() && i < max_seg; ++i) {
        out[3 * i + 0] = segs[i].t0;
        out[3 * i + 1] = segs[i].t1;
        out[3 * i + 2] = segs[i].dt;
    }
    return segs.size();
}

size_t orc_greedy_quench(const OracleLattice* L, double* state, const size_t* order, double spin_length,
                         double rel_tol, size_t max_sweeps) {
    const size_t N = L->N;
    const int D = L->D;
    std::vector<double> H(D);
    double E_prev = orc_energy(L, state);  // greedy_quench, lattice.h:3355
    size_t sweep = 0;
    while (sweep < max_sweeps) {
        // deterministic_sweep(1), lattice.h:3334-3349, sites in the given order
        for (size_t k = 0; k < N; ++k) {
            const size_t i = order[k];
            local_field(*L, state, i, H.data());
            double n2 = 0.0;
            for (int d = 0; d < D; ++d) n2 += H[d] * H[d];
            const double norm = std::sqrt(n2);
            if (norm < 1e-15) continue;
            for (int d = 0; d < D; ++d) state[i * D + d] = -H[d] / norm * spin_length;
        }
        ++sweep;
        const double E_curr = orc_energy(L, state);  // lattice.h:3360-3364
        if (std::abs(E_curr - E_prev) <= rel_tol * (std::abs(E_prev) + 1e-18)) break;
        E_prev = E_curr;
    }
    return sweep;
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ---------------- tangent-linear model (forward mode) ---------------- */

size_t orc_tangent_md(OracleLattice* L, double* state, double* dstate, const double* dJ,
                      const double* dB, double T0, double T1, double dt, size_t max_rec,
                      double* out_mags, double* out_dmags) {
    // Augmented system z = (x, dx): dz/dt = (f(x,p), f_x dx + f_p dp).  RK4 on
    // z is the exact derivative of the discrete RK4 map (linear stage sums).
    const int D = L->D;
    const size_t n = L->N * D;
    const size_t D2 = size_t(D) * D;
    if (D != 3) return 0;
    auto F = [&](const double* z, double* dz, double t) {
        const double* x = z;
        const double* dx = z + n;
        rhs(*L, x, dz, t);
        const bool drive = (L->amp != 0.0);
        double f1 = 0.0, f2 = 0.0;
        if (drive) drive_factors(*L, t, f1, f2);
        for (size_t i = 0; i < L->N; ++i) {
            double H[3], dH[3];
            local_field(*L, x, i, H);
            const size_t atom = i % L->n_atoms;
            if (drive)
                for (int d = 0; d < 3; ++d)
                    H[d] -= L->field_drive[0][atom * 3 + d] * f1 + L->field_drive[1][atom * 3 + d] * f2;
            for (int d = 0; d < 3; ++d) dH[d] = dB ? -dB[atom * 3 + d] : 0.0;
            const double* A = &L->onsite[i * D2];
            for (int d = 0; d < 3; ++d)
                for (int e = 0; e < 3; ++e) dH[d] += 2.0 * A[d * 3 + e] * dx[i * 3 + e];
            for (const auto& b : L->bonds[i]) {
                const double* Sp = x + b.partner * 3;
                const double* dSp = dx + b.partner * 3;
                for (int d = 0; d < 3; ++d)
                    for (int e = 0; e < 3; ++e) {
                        dH[d] += b.J[d * 3 + e] * dSp[e];
                        if (dJ) {
                            const double* g = dJ + size_t(b.ucb) * 9;
                            const double dj = b.transposed ? g[e * 3 + d] : g[d * 3 + e];
                            dH[d] += dj * Sp[e];
                        }
                    }
            }
            const double* S = x + i * 3;
            const double* dS = dx + i * 3;
            double o[3];
            // d(H x S) = dH x S + H x dS
            o[0] = dH[1] * S[2] - dH[2] * S[1] + H[1] * dS[2] - H[2] * dS[1];
            o[1] = dH[2] * S[0] - dH[0] * S[2] + H[2] * dS[0] - H[0] * dS[2];
            o[2] = dH[0] * S[1] - dH[1] * S[0] + H[0] * dS[1] - H[1] * dS[0];
            if (L->alpha > 0.0) {
                // term = (alpha/|S|) (S (S.H) - H |S|^2)
                const double S2 = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
                const double nS = std::sqrt(S2);
                const double SH = S[0] * H[0] + S[1] * H[1] + S[2] * H[2];
                const double SdS = S[0] * dS[0] + S[1] * dS[1] + S[2] * dS[2];
                const double dSH = dS[0] * H[0] + dS[1] * H[1] + dS[2] * H[2] +
                                   S[0] * dH[0] + S[1] * dH[1] + S[2] * dH[2];
                const double g = L->alpha / nS;
                const double dg = -L->alpha * SdS / (nS * S2);
                for (int d = 0; d < 3; ++d) {
                    const double u = S[d] * SH - H[d] * S2;
                    const double du = dS[d] * SH + S[d] * dSH - dH[d] * S2 - H[d] * 2.0 * SdS;
                    o[d] += dg * u + g * du;
                }
            }
            for (int d = 0; d < 3; ++d) dz[n + i * 3 + d] = o[d];
        }
    };
    std::vector<double> z(2 * n);
    std::copy(state, state + n, z.begin());
    std::copy(dstate, dstate + n, z.begin() + n);
    RK4Work w(2 * n);
    size_t nrec = 0;
    auto obs = [&](double) {
        if (nrec < max_rec) {
            if (out_mags) mags(*L, z.data(), out_mags + nrec * 9);
            if (out_dmags) mags(*L, z.data() + n, out_dmags + nrec * 9);
        }
        ++nrec;
    };
    auto stp = [&](double t, double h) { rk4_step_generic(F, z.data(), 2 * n, t, h, w); };
    integrate_const(stp, T0, T1, dt, obs);
    std::copy(z.begin(), z.begin() + n, state);
    std::copy(z.begin() + n, z.end(), dstate);
    return nrec;
}

}  // extern "C"

/* ------------------------------------------------------------------------
 * Mixed SU(2) + SU(3) lattice (MixedLattice, mixed_lattice.h / _md.cpp).
 * State [N2*3 | N3*8] (mixed_lattice.h:3878-3900).  Supported terms: fields,
 * onsite 2A, SU(2)-SU(2) / SU(3)-SU(3) bilinears, SU(2)-SU(3) bilinears (J at
 * the SU(2) site, J^T at the SU(3) site), two-pulse drives per species,
 * SU(2) Gilbert damping and SU(3) Bloch damping.
 * ---------------------------------------------------------------------- */
struct MixedOracle {
    OracleLattice* s2 = nullptr;
    OracleLattice* s3 = nullptr;
    size_t N2 = 0, N3 = 0;
    std::vector<size_t> m2_off, m2_par, m3_off, m3_par;
    std::vector<double> m2_J, m3_J;  // [nnz][3][8], [nnz][8][3]
    double alpha = 0.0;
    double gamma[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<double> eq;          // [N3*8]
};

namespace {

// build_mixed_interactions, bilinear part (mixed_lattice.h:773-800)
void mixed_bonds(MixedOracle& M, int d1, int d2, int d3, int A2, int A3, int nb, const int* src,
                 const int* par, const int* off, const double* J) {
    std::vector<std::vector<std::pair<size_t, const double*>>> b2(M.N2), b3(M.N3);
    for (int i = 0; i < d1; ++i)
        for (int j = 0; j < d2; ++j)
            for (int k = 0; k < d3; ++k)
                for (int atom = 0; atom < A2; ++atom) {
                    const size_t s = ((size_t(i) * d2 + j) * d3 + k) * A2 + atom;
                    for (int b = 0; b < nb; ++b) {
                        if (src[b] != atom) continue;
                        const int pi = wrap(i + off[3 * b], d1), pj = wrap(j + off[3 * b + 1], d2),
                                  pk = wrap(k + off[3 * b + 2], d3);
                        const size_t p = ((size_t(pi) * d2 + pj) * d3 + pk) * A3 + par[b];
                        b2[s].push_back({p, J + b * 24});
                        b3[p].push_back({s, J + b * 24});
                    }
                }
    M.m2_off.assign(M.N2 + 1, 0);
    M.m3_off.assign(M.N3 + 1, 0);
    for (size_t s = 0; s < M.N2; ++s) {
        M.m2_off[s + 1] = M.m2_off[s] + b2[s].size();
        for (auto& e : b2[s]) {
            M.m2_par.push_back(e.first);
            M.m2_J.insert(M.m2_J.end(), e.second, e.second + 24);
        }
    }
    for (size_t s = 0; s < M.N3; ++s) {
        M.m3_off[s + 1] = M.m3_off[s] + b3[s].size();
        for (auto& e : b3[s]) {
            M.m3_par.push_back(e.first);
            for (int r = 0; r < 8; ++r)  // J^T (8x3)
                for (int c = 0; c < 3; ++c) M.m3_J.push_back(e.second[c * 8 + r]);
        }
    }
}

// get_local_field_SU2_flat_into without trilinears / assisted bonds
// (mixed_lattice_md.cpp:716-879); no drive.
void mixed_field_su2(const MixedOracle& M, const double* x, size_t site, double* H) {
    const OracleLattice& L = *M.s2;
    const double* S = x + site * 3;
    const double* x3 = x + M.N2 * 3;
    for (int a = 0; a < 3; ++a) H[a] = -L.field[site * 3 + a];
    const double* A = &L.onsite[site * 9];
    for (int a = 0; a < 3; ++a) {
        double acc = 0.0;
        for (int b = 0; b < 3; ++b) acc += A[a * 3 + b] * S[b];
        H[a] += 2.0 * acc;
    }
    for (size_t k = L.bi_offset[site]; k < L.bi_offset[site + 1]; ++k) {
        const double* Sp = x + L.bi_partner[k] * 3;
        const double* J = &L.bi_J[k * 9];
        for (int a = 0; a < 3; ++a) {
            double acc = 0.0;
            for (int b = 0; b < 3; ++b) acc += J[a * 3 + b] * Sp[b];
            H[a] += acc;
        }
    }
    for (size_t k = M.m2_off[site]; k < M.m2_off[site + 1]; ++k) {
        const double* lp = x3 + M.m2_par[k] * 8;
        const double* J = &M.m2_J[k * 24];
        for (int a = 0; a < 3; ++a) {
            double acc = 0.0;
            for (int c = 0; c < 8; ++c) acc += J[a * 8 + c] * lp[c];
            H[a] += acc;
        }
    }
}

// get_local_field_SU3_flat_into (mixed_lattice_md.cpp:967-1105); no drive.
void mixed_field_su3(const MixedOracle& M, const double* x, size_t site, double* H) {
    const OracleLattice& L = *M.s3;
    const double* x3 = x + M.N2 * 3;
    const double* S = x3 + site * 8;
    for (int a = 0; a < 8; ++a) H[a] = -L.field[site * 8 + a];
    const double* A = &L.onsite[site * 64];
    for (int a = 0; a < 8; ++a) {
        double acc = 0.0;
        for (int b = 0; b < 8; ++b) acc += A[a * 8 + b] * S[b];
        H[a] += 2.0 * acc;
    }
    for (size_t k = L.bi_offset[site]; k < L.bi_offset[site + 1]; ++k) {
        const double* Sp = x3 + L.bi_partner[k] * 8;
        const double* J = &L.bi_J[k * 64];
        for (int a = 0; a < 8; ++a) {
            double acc = 0.0;
            for (int b = 0; b < 8; ++b) acc += J[a * 8 + b] * Sp[b];
            H[a] += acc;
        }
    }
    for (size_t k = M.m3_off[site]; k < M.m3_off[site + 1]; ++k) {
        const double* Sp = x + M.m3_par[k] * 3;
        const double* J = &M.m3_J[k * 24];
        for (int a = 0; a < 8; ++a) {
            double acc = 0.0;
            for (int b = 0; b < 3; ++b) acc += J[a * 3 + b] * Sp[b];
            H[a] += acc;
        }
    }
}

// MixedLattice::landau_lifshitz, both species standard (mixed_lattice_md.cpp:433-605)
void mixed_rhs(const MixedOracle& M, const double* x, double* dxdt, double t) {
    const OracleLattice& L2 = *M.s2;
    const OracleLattice& L3 = *M.s3;
    double f21 = 0.0, f22 = 0.0, f31 = 0.0, f32 = 0.0;
    if (L2.amp != 0.0) drive_factors(L2, t, f21, f22);   // drive_envelopes_SU2 (:265-286)
    if (L3.amp != 0.0) drive_factors(L3, t, f31, f32);   // drive_envelopes_SU3 (:288-313)
    const long n2 = long(M.N2), n3 = long(M.N3);
#ifdef _OPENMP
#pragma omp parallel if (n2 + n3 >= 64)
#endif
    {
#ifdef _OPENMP
#pragma omp for schedule(static) nowait
#endif
        for (long si = 0; si < n2; ++si) {
            const size_t s = size_t(si);
            double H[3];
            mixed_field_su2(M, x, s, H);
            if (f21 != 0.0 || f22 != 0.0) {
                const size_t atom = s % L2.n_atoms;
                const double* fd0 = &L2.field_drive[0][atom * 3];
                const double* fd1 = &L2.field_drive[1][atom * 3];
                for (int a = 0; a < 3; ++a) H[a] -= fd0[a] * f21 + fd1[a] * f22;
            }
            const double* S = x + s * 3;
            double* o = dxdt + s * 3;
            o[0] = H[1] * S[2] - H[2] * S[1];
            o[1] = H[2] * S[0] - H[0] * S[2];
            o[2] = H[0] * S[1] - H[1] * S[0];
            if (M.alpha != 0.0) {  // (:570-578)
                const double S2 = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
                const double SdotH = S[0] * H[0] + S[1] * H[1] + S[2] * H[2];
                const double inv_S = (S2 > 0.0) ? M.alpha / std::sqrt(S2) : 0.0;
                for (int a = 0; a < 3; ++a) o[a] += inv_S * (S[a] * SdotH - H[a] * S2);
            }
        }
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (long si = 0; si < n3; ++si) {
            const size_t s = size_t(si);
            double H[8];
            mixed_field_su3(M, x, s, H);
            if (f31 != 0.0 || f32 != 0.0) {
                const size_t atom = s % L3.n_atoms;
                const double* fd0 = &L3.field_drive[0][atom * 8];
                const double* fd1 = &L3.field_drive[1][atom * 8];
                for (int a = 0; a < 8; ++a) H[a] -= fd0[a] * f31 + fd1[a] * f32;
            }
            const double* lam = x + M.N2 * 3 + s * 8;
            double* o = dxdt + M.N2 * 3 + s * 8;
            su3_cross(H, lam, o);
            for (int a = 0; a < 8; ++a)  // Bloch damping (:596-605)
                if (M.gamma[a] != 0.0) o[a] -= M.gamma[a] * (lam[a] - M.eq[s * 8 + a]);
        }
    }
}

// total_energy_flat without trilinears (mixed_lattice.h:1783-1900)
double mixed_energy(const MixedOracle& M, const double* x) {
    const OracleLattice& L2 = *M.s2;
    const OracleLattice& L3 = *M.s3;
    const double* x3 = x + M.N2 * 3;
    double e = 0.0;
    for (size_t i = 0; i < M.N2; ++i) {
        const double* s = x + i * 3;
        for (int d = 0; d < 3; ++d) e -= s[d] * L2.field[i * 3 + d];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) e += s[a] * L2.onsite[i * 9 + a * 3 + b] * s[b];
        for (size_t k = L2.bi_offset[i]; k < L2.bi_offset[i + 1]; ++k) {
            const double* p = x + L2.bi_partner[k] * 3;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) e += 0.5 * s[a] * L2.bi_J[k * 9 + a * 3 + b] * p[b];
        }
        for (size_t k = M.m2_off[i]; k < M.m2_off[i + 1]; ++k) {
            const double* p = x3 + M.m2_par[k] * 8;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 8; ++b) e += 0.5 * s[a] * M.m2_J[k * 24 + a * 8 + b] * p[b];
        }
    }
    for (size_t i = 0; i < M.N3; ++i) {
        const double* s = x3 + i * 8;
        for (int d = 0; d < 8; ++d) e -= s[d] * L3.field[i * 8 + d];
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) e += s[a] * L3.onsite[i * 64 + a * 8 + b] * s[b];
        for (size_t k = L3.bi_offset[i]; k < L3.bi_offset[i + 1]; ++k) {
            const double* p = x3 + L3.bi_partner[k] * 8;
            for (int a = 0; a < 8; ++a)
                for (int b = 0; b < 8; ++b) e += 0.5 * s[a] * L3.bi_J[k * 64 + a * 8 + b] * p[b];
        }
        for (size_t k = M.m3_off[i]; k < M.m3_off[i + 1]; ++k) {
            const double* p = x + M.m3_par[k] * 3;
            for (int a = 0; a < 8; ++a)
                for (int b = 0; b < 3; ++b) e += 0.5 * s[a] * M.m3_J[k * 24 + a * 3 + b] * p[b];
        }
    }
    return e;
}

// MD observer record (mixed_lattice_md.cpp:1249-1281):
// [M2_anti, M2_local, M2_global (3 each), M3_anti, M3_local, M3_global (8 each), E/(N2+N3)]
void mixed_record(const MixedOracle& M, const double* x, double* out) {
    const OracleLattice& L2 = *M.s2;
    const OracleLattice& L3 = *M.s3;
    double ml2[3] = {0, 0, 0}, ma2[3] = {0, 0, 0}, mg2[3] = {0, 0, 0};
    for (size_t i = 0; i < M.N2; ++i)  // compute_sublattice_magnetizations_from_flat
        for (int d = 0; d < 3; ++d) ml2[d] += x[i * 3 + d];
    for (size_t i = 0; i < M.N2; ++i) {  // compute_magnetization_global_SU2_from_flat
        const double* R = &L2.frames[(i % L2.n_atoms) * 9];
        for (int mu = 0; mu < 3; ++mu)
            for (int nu = 0; nu < 3; ++nu) mg2[mu] += R[mu * 3 + nu] * x[i * 3 + nu];
    }
    for (size_t i = 0; i < M.N2; ++i) {  // compute_magnetization_staggered_SU2_from_flat
        const size_t atom = i % L2.n_atoms;
        const double* R = &L2.frames[atom * 9];
        const double sg = L2.afm[atom];
        for (int mu = 0; mu < 3; ++mu)
            for (int nu = 0; nu < 3; ++nu) ma2[mu] += sg * R[mu * 3 + nu] * x[i * 3 + nu];
    }
    const double* x3 = x + M.N2 * 3;
    double ml3[8] = {0}, ma3[8] = {0}, mg3[8] = {0};
    for (size_t i = 0; i < M.N3; ++i) {
        const double sg = (i % 2 == 0) ? 1.0 : -1.0;
        for (int d = 0; d < 8; ++d) {
            ml3[d] += x3[i * 8 + d];
            ma3[d] += x3[i * 8 + d] * sg;
        }
    }
    for (size_t i = 0; i < M.N3; ++i) {  // compute_magnetization_global_SU3_from_flat
        const double* R = &L3.frames[(i % L3.n_atoms) * 64];
        for (int mu = 0; mu < 8; ++mu)
            for (int nu = 0; nu < 8; ++nu) mg3[mu] += R[mu * 8 + nu] * x3[i * 8 + nu];
    }
    const double n2 = double(M.N2), n3 = double(M.N3);
    for (int d = 0; d < 3; ++d) {
        out[d] = ma2[d] / n2;
        out[3 + d] = ml2[d] / n2;
        out[6 + d] = mg2[d] / n2;
    }
    for (int d = 0; d < 8; ++d) {
        out[9 + d] = ma3[d] / n3;
        out[17 + d] = ml3[d] / n3;
        out[25 + d] = mg3[d] / n3;
    }
    out[33] = mixed_energy(M, x) / (n2 + n3);
}

}  // namespace

extern "C" {

MixedOracle* orc_mixed_new(int d1, int d2, int d3, int A2, const double* field2, const double* onsite2,
                           const double* frames2, const double* afm2, int nb2, const int* src2,
                           const int* par2, const int* off2, const double* J2, int A3,
                           const double* field3, const double* onsite3, const double* frames3, int nb3,
                           const int* src3, const int* par3, const int* off3, const double* J3, int nbm,
                           const int* srcm, const int* parm, const int* offm, const double* Jm) {
    for (int b = 0; b < nbm; ++b)
        for (int a = 0; a < 3; ++a) {
            const int L = a == 0 ? d1 : (a == 1 ? d2 : d3);
            if (offm[3 * b + a] < -L || offm[3 * b + a] > L) return nullptr;
        }
    std::vector<double> afm3(A3, 1.0);
    OracleLattice* l2 = orc_lattice_new(3, A2, d1, d2, d3, nb2, src2, par2, off2, J2, field2, onsite2,
                                        frames2, afm2);
    OracleLattice* l3 = orc_lattice_new(8, A3, d1, d2, d3, nb3, src3, par3, off3, J3, field3, onsite3,
                                        frames3, afm3.data());
    if (!l2 || !l3) {
        orc_lattice_free(l2);
        orc_lattice_free(l3);
        return nullptr;
    }
    auto* M = new MixedOracle();
    M->s2 = l2;
    M->s3 = l3;
    M->N2 = l2->N;
    M->N3 = l3->N;
    M->eq.assign(M->N3 * 8, 0.0);
    mixed_bonds(*M, d1, d2, d3, A2, A3, nbm, srcm, parm, offm, Jm);
    return M;
}

void orc_mixed_free(MixedOracle* M) {
    if (!M) return;
    orc_lattice_free(M->s2);
    orc_lattice_free(M->s3);
    delete M;
}

void orc_mixed_sizes(const MixedOracle* M, size_t* N2, size_t* N3, size_t* nnz22, size_t* nnz33,
                     size_t* nnz23) {
    *N2 = M->N2;
    *N3 = M->N3;
    *nnz22 = M->s2->bi_offset[M->N2];
    *nnz33 = M->s3->bi_offset[M->N3];
    *nnz23 = M->m2_off[M->N2];
}

void orc_mixed_tables(const MixedOracle* M, size_t* off22, size_t* par22, double* J22, size_t* off33,
                      size_t* par33, double* J33, size_t* off23, size_t* par23, double* J23, size_t* off32,
                      size_t* par32, double* J32) {
    std::copy(M->s2->bi_offset.begin(), M->s2->bi_offset.end(), off22);
    std::copy(M->s2->bi_partner.begin(), M->s2->bi_partner.end(), par22);
    std::copy(M->s2->bi_J.begin(), M->s2->bi_J.end(), J22);
    std::copy(M->s3->bi_offset.begin(), M->s3->bi_offset.end(), off33);
    std::copy(M->s3->bi_partner.begin(), M->s3->bi_partner.end(), par33);
    std::copy(M->s3->bi_J.begin(), M->s3->bi_J.end(), J33);
    std::copy(M->m2_off.begin(), M->m2_off.end(), off23);
    std::copy(M->m2_par.begin(), M->m2_par.end(), par23);
    std::copy(M->m2_J.begin(), M->m2_J.end(), J23);
    std::copy(M->m3_off.begin(), M->m3_off.end(), off32);
    std::copy(M->m3_par.begin(), M->m3_par.end(), par32);
    std::copy(M->m3_J.begin(), M->m3_J.end(), J32);
}

void orc_mixed_set_damping(MixedOracle* M, double alpha, const double* gamma8, const double* eq) {
    M->alpha = alpha;
    for (int a = 0; a < 8; ++a) M->gamma[a] = gamma8 ? gamma8[a] : 0.0;
    if (eq) std::copy(eq, eq + M->N3 * 8, M->eq.begin());
    else std::fill(M->eq.begin(), M->eq.end(), 0.0);
}

// set_pulse_SU2 / set_pulse_SU3 (mixed_lattice_md.cpp:114-263): drive
// directions are F^T * B_global per atom.
void orc_mixed_set_pulse(MixedOracle* M, int species, const double* field_in1, double t1,
                         const double* field_in2, double t2, double amp, double width, double freq) {
    OracleLattice& L = species == 3 ? *M->s3 : *M->s2;
    const int D = L.D;
    for (size_t a = 0; a < L.n_atoms; ++a) {
        const double* R = &L.frames[a * D * D];
        for (int mu = 0; mu < D; ++mu) {
            double v1 = 0.0, v2 = 0.0;
            for (int nu = 0; nu < D; ++nu) {
                v1 += R[nu * D + mu] * field_in1[a * D + nu];
                v2 += R[nu * D + mu] * field_in2[a * D + nu];
            }
            L.field_drive[0][a * D + mu] = v1;
            L.field_drive[1][a * D + mu] = v2;
        }
    }
    L.t_pulse[0] = t1;
    L.t_pulse[1] = t2;
    L.amp = amp;
    L.width = width;
    L.freq = freq;
}

void orc_mixed_local_field(const MixedOracle* M, const double* x, int species, size_t site, double* H) {
    if (species == 3) mixed_field_su3(*M, x, site, H);
    else mixed_field_su2(*M, x, site, H);
}

void orc_mixed_rhs(const MixedOracle* M, const double* x, double* dxdt, double t) { mixed_rhs(*M, x, dxdt, t); }

double orc_mixed_energy(const MixedOracle* M, const double* x) { return mixed_energy(*M, x); }

void orc_mixed_record(const MixedOracle* M, const double* x, double* out34) { mixed_record(*M, x, out34); }

// Boost runge_kutta4 + integrate_const over the flat mixed state, one
// record (34 doubles) per observed time.
size_t orc_mixed_integrate_md(const MixedOracle* M, double* state, double T0, double T1, double dt,
                              size_t max_rec, double* times, double* recs) {
    const size_t n = M->N2 * 3 + M->N3 * 8;
    RK4Work w(n);
    size_t nrec = 0;
    auto obs = [&](double t) {
        if (nrec < max_rec) {
            if (times) times[nrec] = t;
            if (recs) mixed_record(*M, state, recs + nrec * 34);
        }
        ++nrec;
    };
    auto F = [&](const double* x, double* dx, double t) { mixed_rhs(*M, x, dx, t); };
    auto stp = [&](double t, double h) { rk4_step_generic(F, state, n, t, h, w); };
    integrate_const(stp, T0, T1, dt, obs);
    return nrec;
}

int orc_mixed_integrate_gpu(const MixedOracle* M, double* state, double T0, double T1, double dt,
                            size_t save_interval, const char* method, size_t max_rec, size_t* n_rec,
                            double* times, double* states, double* recs) {
    const size_t n = M->N2 * 3 + M->N3 * 8;
    GpuStepWork probe(0);
    double dummy = 0.0;
    const bool known = gpu_step([](const double*, double*, double) {}, &dummy, 0, 0.0, 1.0, method, probe);
    size_t nrec = 0;
    auto obs = [&](double t) {
        if (nrec < max_rec) {
            if (times) times[nrec] = t;
            if (states) std::memcpy(states + nrec * n, state, n * sizeof(double));
            if (recs) mixed_record(*M, state, recs + nrec * 34);
        }
        ++nrec;
    };
    auto F = [M](const double* x, double* dx, double t) { mixed_rhs(*M, x, dx, t); };
    gpu_integrate(F, state, n, T0, T1, dt, save_interval, method, obs);
    if (n_rec) *n_rec = nrec;
    return known ? 1 : 0;
}

// Dormand-Prince 5(4) over the flat mixed state (MixedLattice::
// integrate_ode_system with "dopri5", mixed_lattice.h:3960-4010): the
// controls of orc_integrate_dopri5, a record (34 doubles) every
// save_interval-th observer call.
long orc_mixed_integrate_dopri5(const MixedOracle* M, double* state, double T0, double T1, double dt,
                                int control, double abs_tol, double rel_tol, size_t save_interval,
                                const double* segs, size_t n_seg, size_t max_rec, double* times, double* recs,
                                size_t* n_accepted, size_t* n_rejected) {
    const size_t n = M->N2 * 3 + M->N3 * 8;
    size_t calls = 0, nrec = 0;
    auto obs = [&](double t) {
        if (calls % save_interval == 0) {
            if (nrec < max_rec) {
                if (times) times[nrec] = t;
                if (recs) mixed_record(*M, state, recs + nrec * 34);
            }
            ++nrec;
        }
        ++calls;
    };
    auto F = [M](const double* x, double* dx, double t) { mixed_rhs(*M, x, dx, t); };
    size_t acc = 0, rej = 0;
    const bool ok = dp5_controls(F, state, n, T0, T1, dt, control, abs_tol, rel_tol, segs, n_seg, obs, acc, rej);
    if (n_accepted) *n_accepted = acc;
    if (n_rejected) *n_rejected = rej;
    return ok ? long(nrec) : -1;
}

}  // extern "C"
