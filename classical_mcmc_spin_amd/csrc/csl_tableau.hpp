// csl_tableau.hpp — the fixed-step explicit Runge–Kutta methods of the
// reference GPU stepper (gpu::ode::step, integrator.cuh:91-458) as stage
// tables for the generic stage engine.
//
// The reference writes every method as a sequence of element-wise updates of
// named buffers (update_arrays_kernel out = a1*in1 + a2*in2 and its three- to
// six-term relatives, gpu_common_helpers.cu:9-120).  Each buffer it ever
// forms is a linear combination  g*x + dt * sum_j w_j k_j  of the step's
// start state x and the stage derivatives k_j, so a method is fully
// described by, per stage s, its time fraction c_s and the combination that
// forms its input, plus the combination that forms the result.  The builders
// below replay the reference's update sequences on such combinations (class
// Lin), which gives the table with the reference's coefficients; the device
// then forms each stage input in one pass, g*x + sum_j (w_j dt) k_j.
//
// "rk4" keeps its dedicated kernels (bit-level Boost / GPU RK4 order); the
// table here exists for it too so the generic engine can be checked against
// them.
#pragma once

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace cslh {

constexpr int kMaxStages = 13;

// g * x + dt * sum_j k[j] * k_j
struct Lin {
    double g = 0.0;
    double k[kMaxStages] = {};
    static Lin state() {
        Lin l;
        l.g = 1.0;
        return l;
    }
    static Lin deriv(int j) {
        Lin l;
        l.k[j] = 1.0;
        return l;
    }
};

// out = a1*in1 + a2*in2 (update_arrays_kernel); a2 multiplies a derivative
// buffer in units of dt when in2 is a k_j
inline Lin comb(double a1, const Lin& in1, double a2, const Lin& in2) {
    Lin o;
    o.g = a1 * in1.g + a2 * in2.g;
    for (int j = 0; j < kMaxStages; ++j) o.k[j] = a1 * in1.k[j] + a2 * in2.k[j];
    return o;
}

struct Tableau {
    std::string name;
    int S = 0;                       // RHS evaluations per step
    double c[kMaxStages] = {};       // stage s is evaluated at t + c[s]*dt
    Lin in[kMaxStages];              // input of stage s (in[0] = x)
    Lin out;                         // the step's result
};

namespace tab_detail {

struct Builder {
    Tableau T;
    Lin x = Lin::state();
    Lin K(int j) const { return Lin::deriv(j); }
    // evaluate stage number T.S at t + c*dt on input y
    void stage(const Lin& y, double c) {
        T.in[T.S] = y;
        T.c[T.S] = c;
        ++T.S;
    }
};

// integrator.cuh:91-99
inline Tableau euler() {
    Builder b;
    b.stage(b.x, 0.0);
    b.T.out = comb(1.0, b.x, 1.0, b.K(0));
    return b.T;
}

// integrator.cuh:100-114 (modified midpoint)
inline Tableau rk2() {
    Builder b;
    b.stage(b.x, 0.0);
    b.stage(comb(1.0, b.x, 0.5, b.K(0)), 0.5);
    b.T.out = comb(1.0, b.x, 1.0, b.K(1));
    return b.T;
}

// integrator.cuh:115-139
inline Tableau rk4() {
    Builder b;
    b.stage(b.x, 0.0);
    b.stage(comb(1.0, b.x, 0.5, b.K(0)), 0.5);
    b.stage(comb(1.0, b.x, 0.5, b.K(1)), 0.5);
    b.stage(comb(1.0, b.x, 1.0, b.K(2)), 1.0);
    Lin o = b.x;
    const double w[4] = {1.0 / 6.0, 2.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0};
    for (int j = 0; j < 4; ++j) o = comb(1.0, o, w[j], b.K(j));
    b.T.out = o;
    return b.T;
}

// y = x + sum_i a_i k_{j_i} (the rk_stage_update_n kernels)
inline Lin row(const Lin& x, std::initializer_list<std::pair<int, double>> terms) {
    Lin o = x;
    for (const auto& t : terms) o = comb(1.0, o, t.second, Lin::deriv(t.first));
    return o;
}

// integrator.cuh:141-185 (Cash–Karp; the final update is the Dormand–Prince
// kernel with b5 = 0 on k6)
inline Tableau rkck54() {
    Builder b;
    const Lin& x = b.x;
    b.stage(x, 0.0);
    b.stage(row(x, {{0, 1.0 / 5.0}}), 1.0 / 5.0);
    b.stage(row(x, {{0, 3.0 / 40.0}, {1, 9.0 / 40.0}}), 3.0 / 10.0);
    b.stage(row(x, {{0, 3.0 / 10.0}, {1, -9.0 / 10.0}, {2, 6.0 / 5.0}}), 3.0 / 5.0);
    b.stage(row(x, {{0, -11.0 / 54.0}, {1, 5.0 / 2.0}, {2, -70.0 / 27.0}, {3, 35.0 / 27.0}}), 1.0);
    b.stage(row(x, {{0, 1631.0 / 55296.0}, {1, 175.0 / 512.0}, {2, 575.0 / 13824.0}, {3, 44275.0 / 110592.0},
                    {4, 253.0 / 4096.0}}),
            7.0 / 8.0);
    b.T.out = row(x, {{0, 37.0 / 378.0}, {2, 250.0 / 621.0}, {3, 125.0 / 594.0}, {5, 0.0}, {5, 512.0 / 1771.0}});
    return b.T;
}

// integrator.cuh:186-229
inline Tableau dopri5() {
    Builder b;
    const Lin& x = b.x;
    b.stage(x, 0.0);
    b.stage(row(x, {{0, 1.0 / 5.0}}), 1.0 / 5.0);
    b.stage(row(x, {{0, 3.0 / 40.0}, {1, 9.0 / 40.0}}), 3.0 / 10.0);
    b.stage(row(x, {{0, 44.0 / 45.0}, {1, -56.0 / 15.0}, {2, 32.0 / 9.0}}), 4.0 / 5.0);
    b.stage(row(x, {{0, 19372.0 / 6561.0}, {1, -25360.0 / 2187.0}, {2, 64448.0 / 6561.0}, {3, -212.0 / 729.0}}),
            8.0 / 9.0);
    b.stage(row(x, {{0, 9017.0 / 3168.0}, {1, -355.0 / 33.0}, {2, 46732.0 / 5247.0}, {3, 49.0 / 176.0},
                    {4, -5103.0 / 18656.0}}),
            1.0);
    b.T.out = row(x, {{0, 35.0 / 384.0}, {2, 500.0 / 1113.0}, {3, 125.0 / 192.0}, {4, -2187.0 / 6784.0},
                      {5, 11.0 / 84.0}});
    return b.T;
}

// integrator.cuh:230-370 (Fehlberg 7(8) stages, the 7th-order weights on
// k1, k6..k11; 11 RHS evaluations)
inline Tableau rkf78() {
    Builder b;
    const Lin& x = b.x;
    b.stage(x, 0.0);
    b.stage(row(x, {{0, 2.0 / 27.0}}), 2.0 / 27.0);
    b.stage(row(x, {{0, 1.0 / 36.0}, {1, 1.0 / 12.0}}), 1.0 / 9.0);
    b.stage(row(x, {{0, 1.0 / 24.0}, {2, 1.0 / 8.0}}), 1.0 / 6.0);
    b.stage(row(x, {{0, 5.0 / 12.0}, {3, -25.0 / 16.0}, {2, 25.0 / 16.0}}), 5.0 / 12.0);
    b.stage(row(x, {{0, 1.0 / 20.0}, {3, 1.0 / 4.0}, {4, 1.0 / 5.0}}), 1.0 / 2.0);
    b.stage(row(x, {{0, -25.0 / 108.0}, {3, 125.0 / 108.0}, {4, -65.0 / 27.0}, {5, 125.0 / 54.0}}), 5.0 / 6.0);
    b.stage(row(x, {{0, 31.0 / 300.0}, {4, 61.0 / 225.0}, {5, -2.0 / 9.0}, {6, 13.0 / 900.0}}), 1.0 / 6.0);
    b.stage(row(x, {{0, 2.0}, {3, -53.0 / 6.0}, {4, 704.0 / 45.0}, {5, -107.0 / 9.0}, {6, 67.0 / 90.0},
                    {7, 3.0}}),
            2.0 / 3.0);
    b.stage(row(x, {{0, -91.0 / 108.0}, {3, 23.0 / 108.0}, {4, -976.0 / 135.0}, {5, 311.0 / 54.0},
                    {6, -19.0 / 60.0}, {7, 17.0 / 6.0}, {8, -1.0 / 12.0}}),
            1.0 / 3.0);
    b.stage(row(x, {{0, 2383.0 / 4100.0}, {3, -341.0 / 164.0}, {4, 4496.0 / 1025.0}, {5, -301.0 / 82.0},
                    {6, 2133.0 / 4100.0}, {7, 45.0 / 82.0}, {8, 45.0 / 164.0}, {9, 18.0 / 41.0}}),
            1.0);
    b.T.out = row(x, {{0, 41.0 / 840.0}, {5, 34.0 / 105.0}, {6, 9.0 / 35.0}, {7, 9.0 / 35.0}, {8, 9.0 / 280.0},
                      {9, 9.0 / 280.0}, {10, 41.0 / 840.0}});
    return b.T;
}

// integrator.cuh:371-405: modified midpoint with four substeps h = dt/4,
// y_prev / y_mid / y_next rotated through the loop, final
// 0.5 (y_mid + y_prev) + (h/2) f(t + dt, y_mid)
inline Tableau bulirsch_stoer() {
    Builder b;
    const double h = 0.25;
    Lin y_prev = b.x;
    b.stage(b.x, 0.0);
    Lin y_mid = comb(1.0, b.x, h, b.K(0));
    for (int i = 1; i < 4; ++i) {
        b.stage(y_mid, i * h);
        const Lin y_next = comb(1.0, y_prev, 2.0 * h, b.K(i));
        y_prev = y_mid;
        y_mid = y_next;
    }
    b.stage(y_mid, 1.0);
    const Lin y_next = comb(0.5, y_mid, 0.5, y_prev);
    b.T.out = comb(1.0, y_next, 0.5 * h, b.K(4));
    return b.T;
}

// integrator.cuh:406-449: SSPRK(5,3) in the Shu–Osher form of the reference
// (k reused, u = second stage value kept for the last update)
inline Tableau ssprk53() {
    constexpr double a30 = 0.355909775063327, a32 = 0.644090224936674;
    constexpr double a40 = 0.367933791638137, a43 = 0.632066208361863;
    constexpr double a52 = 0.237593836598569, a54 = 0.762406163401431;
    constexpr double b10 = 0.377268915331368, b21 = 0.377268915331368;
    constexpr double b32 = 0.242995220537396, b43 = 0.238458932846290;
    constexpr double b54 = 0.287632146308408;
    constexpr double c1 = 0.377268915331368, c2 = 0.754537830662736;
    constexpr double c3 = 0.728985661612188, c4 = 0.699226135931670;
    Builder b;
    b.stage(b.x, 0.0);
    const Lin tmp1 = comb(1.0, b.x, b10, b.K(0));
    b.stage(tmp1, c1);
    const Lin u = comb(1.0, tmp1, b21, b.K(1));
    b.stage(u, c2);
    // update_arrays_three_kernel(out, in1, a1, in2, a2, in3, a3)
    auto three = [](double a1, const Lin& i1, double a2, const Lin& i2, double a3, const Lin& i3) {
        return comb(1.0, comb(a1, i1, a2, i2), a3, i3);
    };
    const Lin tmp3 = three(a30, b.x, a32, u, b32, b.K(2));
    b.stage(tmp3, c3);
    const Lin tmp4 = three(a40, b.x, a43, tmp3, b43, b.K(3));
    b.stage(tmp4, c4);
    b.T.out = three(a52, u, a54, tmp4, b54, b.K(4));
    return b.T;
}

}  // namespace tab_detail

// Method name -> table, with the reference's names and fallbacks
// (integrator.cuh:450-458: rk54 / rkf54 reuse dopri5, anything unknown is
// reported on stderr and runs ssprk53).  *known is false for the fallback.
inline Tableau tableau_for(const char* method, bool* known = nullptr) {
    const std::string m = method ? method : "rk4";
    using namespace tab_detail;
    Tableau T;
    bool ok = true;
    if (m == "euler") T = euler();
    else if (m == "rk2" || m == "midpoint") T = rk2();
    else if (m == "rk4") T = rk4();
    else if (m == "rk5" || m == "rkck54") T = rkck54();
    else if (m == "dopri5" || m == "rk54" || m == "rkf54") T = dopri5();
    else if (m == "rk78" || m == "rkf78") T = rkf78();
    else if (m == "bulirsch_stoer" || m == "bs") T = bulirsch_stoer();
    else if (m == "ssprk53") T = ssprk53();
    else {
        ok = false;
        T = ssprk53();
    }
    T.name = ok ? m : "ssprk53";
    if (known) *known = ok;
    return T;
}

// Non-zero terms of a combination as (derivative index, weight * dt) pairs.
inline int lin_terms(const Lin& l, int upto, double dt, int* idx, double* w) {
    int m = 0;
    for (int j = 0; j < upto; ++j)
        if (l.k[j] != 0.0) {
            idx[m] = j;
            w[m] = l.k[j] * dt;
            ++m;
        }
    return m;
}

}  // namespace cslh
