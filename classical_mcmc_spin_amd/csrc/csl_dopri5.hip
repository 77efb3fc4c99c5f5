// csl_dopri5.hip — Dormand-Prince 5(4) integration of batched trajectories.
//
// One workgroup per trajectory, the lattice state in LDS, every thread owning
// SPT sites whose state x and stage derivatives k1..k5 stay in VGPRs.  Each
// trajectory runs its own step-size sequence inside one launch: the error
// norm of an attempt is a workgroup max-reduction, and the accept / reject
// decision and the new step size are computed identically by every thread.
//
// Integration semantics (the reference's CPU integrate_ode_system with
// make_controlled<runge_kutta_dopri5>(abs, rel), lattice.h:4072-4085):
//   control 0  fixed steps dt from each point of the observation grid (the
//              reference GPU stepper's "dopri5", integrator.cuh:186-228),
//   control 1  integrate_adaptive: observer at T0 and after every accepted
//              step, the last step cut to land on T1 (molecular_dynamics_cpu,
//              lattice_md.cpp:452-453),
//   control 2  integrate_const with the controlled stepper: observer on the
//              grid, adaptive steps inside each interval, the step size
//              carried from interval to interval (pulse drivers,
//              lattice_md.cpp:570-575); optionally one integrate_const call
//              per pulse-window segment of the trajectory, each starting
//              from dt again (W3 chunking, lattice_md.cpp:562-571).
// Error norm (Boost.Odeint default_error_checker, a_x = a_dxdt = 1):
//   max_i |xerr_i| / (abs + rel (|x_i| + |dt| |dxdt_i|));
// step control (default_step_adjuster): reject if err > 1 and scale dt by
// max(0.9 err^(-1/3), 0.2); after an accepted step with err < 1/2 scale dt
// by 0.9 max(err, 5^-5)^(-1/5).  The tableau is Dormand & Prince (1980), used
// first-same-as-last: the derivative at an accepted state starts the next
// step.
//
// LDS: two state buffers (stage inputs alternate between them, one barrier
// per stage), the observable partials and the per-wave error maxima.
#include "csl_device.hpp"
#include "csl_lds.hpp"

namespace csl {

namespace {

template <int SPT>
constexpr int dopri5_max_block() {
    return SPT == 3 ? 384 : 512;
}

// Dormand-Prince 5(4) tableau
constexpr double kC2 = 1.0 / 5.0, kC3 = 3.0 / 10.0, kC4 = 4.0 / 5.0, kC5 = 8.0 / 9.0;
constexpr double kA21 = 1.0 / 5.0;
constexpr double kA31 = 3.0 / 40.0, kA32 = 9.0 / 40.0;
constexpr double kA41 = 44.0 / 45.0, kA42 = -56.0 / 15.0, kA43 = 32.0 / 9.0;
constexpr double kA51 = 19372.0 / 6561.0, kA52 = -25360.0 / 2187.0, kA53 = 64448.0 / 6561.0,
                 kA54 = -212.0 / 729.0;
constexpr double kA61 = 9017.0 / 3168.0, kA62 = -355.0 / 33.0, kA63 = 46732.0 / 5247.0, kA64 = 49.0 / 176.0,
                 kA65 = -5103.0 / 18656.0;
constexpr double kB1 = 35.0 / 384.0, kB3 = 500.0 / 1113.0, kB4 = 125.0 / 192.0, kB5 = -2187.0 / 6784.0,
                 kB6 = 11.0 / 84.0;
// 5th minus embedded 4th-order weights
constexpr double kE1 = 35.0 / 384.0 - 5179.0 / 57600.0, kE3 = 500.0 / 1113.0 - 7571.0 / 16695.0,
                 kE4 = 125.0 / 192.0 - 393.0 / 640.0, kE5 = -2187.0 / 6784.0 + 92097.0 / 339200.0,
                 kE6 = 11.0 / 84.0 - 187.0 / 2100.0, kE7 = -1.0 / 40.0;

constexpr double kEps = 2.220446049250313e-16;  // numeric_limits<double>::epsilon()
constexpr int kMaxRejections = 500;

// Boost.Odeint less_with_sign / less_eq_with_sign for dt > 0
__device__ __forceinline__ bool less_with_sign(double t1, double t2) { return t2 - t1 > kEps; }
__device__ __forceinline__ bool less_eq_with_sign(double t1, double t2) { return t1 - t2 <= kEps; }
// integrate_const's observation time start + k * dt, rounded after the
// product as on the host (no fused multiply-add)
__device__ __forceinline__ double grid_time(double start, long long k, double dt) {
#pragma clang fp contract(off)
    return start + double(k) * dt;
}
// std::max(a, b) semantics (b when a < b, else a)
__device__ __forceinline__ double smax(double a, double b) { return a < b ? b : a; }

// Drive envelope of one pulse (the formula of drive_factor), expanded in
// place: it runs where few values are live, so the FP64 exp/cos need no call
// and no spill.
__device__ __forceinline__ double drive_factor_inline(double amp, double t, double tp, double width,
                                                      double freq) {
    const double d = t - tp;
    const double a = d / (2.0 * width);
    return amp * exp(-(a * a)) * cos(freq * d);
}

// Value of lane `src` (wave-uniform) of a double, as a scalar.
__device__ __forceinline__ double readlane_f64(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}

}  // namespace

// ATOM_MAJOR: every wave owns SPT * 64 cells of one sublattice (site m of
// lane l of wave w at atom(w) * npad + (w_in_atom * SPT + m) * 64 + l) when
// the block tiles the sublattices, so a wave's sites share one atom's bond
// matrices.
template <int SPT, int ZMAX, bool UNI, bool ATOM_MAJOR = false>
__global__ void __launch_bounds__(dopri5_max_block<SPT>()) k_dopri5(LatDev Lk, Dopri5Args A) {
    LatDev L0 = Lk;
    L0.mats += size_t(blockIdx.x) * A.mats_stride;
    L0.field += size_t(blockIdx.x) * A.field_stride;
    extern __shared__ double smem[];
    const int NP = L0.NP;
    const LdsState bufA = lds_state(smem, NP);
    const LdsState bufB = lds_state(smem + 3 * NP, NP);
    const int b = blockIdx.x;
    const int T = blockDim.x;
    const int tid = threadIdx.x;
    const int wave = uniform(tid >> 6);
    const int lane = tid & (kWave - 1);
    const int nwaves = T >> 6;
    double* red = smem + 6 * NP;     // [nwaves][10] observable partials
    double* emx = red + 10 * nwaves;  // [nwaves] error maxima

    const double tp1 = A.t_pulse[2 * b], tp2 = A.t_pulse[2 * b + 1];
    const double am1 = A.amp[2 * b], am2 = A.amp[2 * b + 1];
    const bool drive_on = (am1 != 0.0) || (am2 != 0.0);
    const double alpha = L0.alpha;
    const bool want_e = A.energy != nullptr;
    const double inv_n = 1.0 / double(L0.N);
    const int control = A.control;
    const int si = A.save_interval;

    // initial state -> bufA (device order)
    const double* init = A.init + size_t(b) * A.init_stride;
    for (int p = tid; p < NP; p += T) {
        const int atom = p / L0.npad, cell = p - atom * L0.npad;
        double v[3] = {0.0, 0.0, 0.0};
        if (cell < L0.n_cells) {
            const size_t i = size_t(cell) * L0.n_atoms + atom;
            v[0] = init[3 * i];
            v[1] = init[3 * i + 1];
            v[2] = init[3 * i + 2];
        }
        bufA.store(p, v);
    }

    const int wpa = ATOM_MAJOR ? L0.npad / (kWave * SPT) : 1;  // waves per atom
    const int atom_w = ATOM_MAJOR ? wave / wpa : 0;
    const int pbase = ATOM_MAJOR ? atom_w * L0.npad + (wave - atom_w * wpa) * (SPT * kWave) + lane : tid;
    auto site = [&](int m) { return ATOM_MAJOR ? pbase + m * kWave : tid + m * T; };
    double x[SPT][3], k1[SPT][3], k2[SPT][3], k3[SPT][3], k4[SPT][3], k5[SPT][3];
    double xn[SPT][3], k7[SPT][3];
    int nb[SPT][ZMAX];
    int cnt[SPT], atm[SPT];
    bool act[SPT];
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        const int p = site(m);
        const int pw = ATOM_MAJOR ? atom_w * L0.npad : wave * kWave + m * T;  // wave-uniform
        atm[m] = ATOM_MAJOR ? atom_w : uniform(pw < NP ? pw / L0.npad : 0);
        act[m] = active_site(L0, p);
        const int c = UNI ? uniform(pw < NP ? as_const(L0.cnt_atom)[atm[m]] : 0) : (p < NP ? L0.cnt[p] : 0);
        cnt[m] = c;
#pragma unroll
        for (int k = 0; k < ZMAX; ++k) nb[m][k] = (act[m] && k < c) ? L0.col[size_t(k) * NP + p] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        const int p = site(m);
        if (act[m]) {
            bufA.load(p, x[m]);
        } else {
            x[m][0] = x[m][1] = x[m][2] = 0.0;
        }
    }

    // dS/dt of the owned sites from the stage input in `cur` with drive
    // factors f1, f2 (zero when the drive is off); with obs, the observable
    // partials of that state go to red[] (all lanes of every wave must take
    // part).
    auto rhs = [&](const LdsState& cur, double f1, double f2, double (&k)[SPT][3], bool obs)
                   __attribute__((always_inline)) {
        const LatDev L = refresh(L0);
        // LDS indices offset by an opaque zero: the address arithmetic is
        // redone in every stage instead of being hoisted out of the time loop
        // and held in VGPRs for both buffers (about 16 VGPRs per owned site)
        const int zz = opaque_zero();
        double part[10];
#pragma unroll
        for (int j = 0; j < 10; ++j) part[j] = 0.0;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = site(m) + zz;
            k[m][0] = k[m][1] = k[m][2] = 0.0;
            if (act[m]) {
                const int atom = ATOM_MAJOR ? atom_w : atm[m];
                int q[ZMAX];
#pragma unroll
                for (int j = 0; j < ZMAX; ++j) q[j] = nb[m][j] + zz;
                double S[3];
                cur.load(p, S);
                double H[3];
                local_field_lds<ZMAX, UNI>(L, cur, p, atom, q, cnt[m], S, H);
                if (obs) site_obs(L, atom, p, S, H, part);
                if (drive_on && (f1 != 0.0 || f2 != 0.0)) sub_drive(L, atom, f1, f2, H);
                ll_su2(H, S, alpha, k[m]);
            }
        }
        if (obs) {
#pragma unroll
            for (int j = 0; j < 10; ++j) {
                if (j == 9 && !want_e) break;
                const double v = wave_sum_lane63(part[j]);
                if (lane == kWave - 1) red[wave * 10 + j] = v;
            }
        }
    };
    // record ri of the partials in red[] (after a barrier), time t
    auto publish = [&](long long ri, double t) {
        if (ri < A.max_rec && wave == 0 && lane < 10) {
            double v = 0.0;
            for (int w = 0; w < nwaves; ++w) v += red[w * 10 + lane];
            const size_t r = size_t(b) * size_t(A.max_rec) + size_t(ri);
            if (lane < 9) {
                if (A.mags) A.mags[r * 9 + lane] = v * inv_n;
            } else if (A.energy) {
                A.energy[r] = v * inv_n;
            }
            if (lane == 0 && A.times) A.times[r] = t;
        }
    };
    // Drive factors of up to 32 (time, pulse) pairs, one per lane: lane l
    // evaluates pulse l & 1 at time ts (the caller's per-lane time); read back
    // with readlane_f64(fl, 2 * i + pulse).
    auto lane_drive = [&](double ts) __attribute__((always_inline)) -> double {
        const bool second = lane & 1;
        return drive_factor_inline(second ? am2 : am1, ts, second ? tp2 : tp1, L0.width, L0.freq);
    };
    auto store_stage = [&](const LdsState& dst, const double (&y)[SPT][3]) {
#pragma unroll
        for (int m = 0; m < SPT; ++m)
            if (act[m]) dst.store(site(m), y[m]);
    };

    // One attempt of size h from (x, k1) at t: stages 2..7 (inputs alternate
    // bufB / bufA; the accepted candidate xn ends in bufA), k7 = f(xn, t7),
    // the error norm (control != 0) and, with obs, the observable partials of
    // xn.  Ends with a barrier; returns the workgroup error norm.
    auto attempt = [&](double t, double h, double t7, bool obs) __attribute__((always_inline)) -> double {
        double y[SPT][3];
        // drive factors of the stage times t + c2 h, t + c3 h, t + c4 h,
        // t + c5 h, t + h and t7 (lanes 0..11), evaluated while only x and k1
        // are live
        double fl = 0.0;
        if (drive_on) {
            const int st = lane >> 1;
            const double c = st == 0 ? kC2 : (st == 1 ? kC3 : (st == 2 ? kC4 : (st == 3 ? kC5 : 1.0)));
            const double ts = st == 5 ? t7 : t + c * h;
            fl = lane < 12 ? lane_drive(ts) : 0.0;
        }
        auto f = [&](int i, int pl) { return drive_on ? readlane_f64(fl, 2 * i + pl) : 0.0; };
        // stage 2
#pragma unroll
        for (int m = 0; m < SPT; ++m)
#pragma unroll
            for (int d = 0; d < 3; ++d) y[m][d] = x[m][d] + h * kA21 * k1[m][d];
        store_stage(bufB, y);
        __syncthreads();
        rhs(bufB, f(0, 0), f(0, 1), k2, false);
        // stage 3
#pragma unroll
        for (int m = 0; m < SPT; ++m)
#pragma unroll
            for (int d = 0; d < 3; ++d) y[m][d] = x[m][d] + h * (kA31 * k1[m][d] + kA32 * k2[m][d]);
        store_stage(bufA, y);
        __syncthreads();
        rhs(bufA, f(1, 0), f(1, 1), k3, false);
        // stage 4
#pragma unroll
        for (int m = 0; m < SPT; ++m)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                y[m][d] = x[m][d] + h * (kA41 * k1[m][d] + kA42 * k2[m][d] + kA43 * k3[m][d]);
        store_stage(bufB, y);
        __syncthreads();
        rhs(bufB, f(2, 0), f(2, 1), k4, false);
        // stage 5
#pragma unroll
        for (int m = 0; m < SPT; ++m)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                y[m][d] = x[m][d] + h * (kA51 * k1[m][d] + kA52 * k2[m][d] + kA53 * k3[m][d] + kA54 * k4[m][d]);
        store_stage(bufA, y);
        __syncthreads();
        rhs(bufA, f(3, 0), f(3, 1), k5, false);
        // stage 6 (k6 lives in y's registers after the store)
#pragma unroll
        for (int m = 0; m < SPT; ++m)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                y[m][d] = x[m][d] + h * (kA61 * k1[m][d] + kA62 * k2[m][d] + kA63 * k3[m][d] +
                                         kA64 * k4[m][d] + kA65 * k5[m][d]);
        store_stage(bufB, y);
        __syncthreads();
        double k6[SPT][3];
        rhs(bufB, f(4, 0), f(4, 1), k6, false);
        // candidate and the error weights without k7
        double ep[SPT][3];
#pragma unroll
        for (int m = 0; m < SPT; ++m)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                xn[m][d] = x[m][d] + h * (kB1 * k1[m][d] + kB3 * k3[m][d] + kB4 * k4[m][d] + kB5 * k5[m][d] +
                                          kB6 * k6[m][d]);
                ep[m][d] = kE1 * k1[m][d] + kE3 * k3[m][d] + kE4 * k4[m][d] + kE5 * k5[m][d] + kE6 * k6[m][d];
            }
        store_stage(bufA, xn);
        __syncthreads();
        rhs(bufA, f(5, 0), f(5, 1), k7, obs);
        double e = 0.0;
        if (control != 0) {
            const double ah = fabs(h);
#pragma unroll
            for (int m = 0; m < SPT; ++m)
                if (act[m]) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const double xe = h * (ep[m][d] + kE7 * k7[m][d]);
                        const double r =
                            fabs(xe) / (A.abs_tol + A.rel_tol * (fabs(x[m][d]) + ah * fabs(k1[m][d])));
                        e = fmax(e, r);
                    }
                }
            e = wave_max(e);
            if (lane == 0) emx[wave] = e;
        }
        __syncthreads();
        double err = 0.0;
        if (control != 0)
            for (int w = 0; w < nwaves; ++w) err = fmax(err, emx[w]);
        return err;
    };
    auto accept = [&]() {
#pragma unroll
        for (int m = 0; m < SPT; ++m)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                x[m][d] = xn[m][d];
                k1[m][d] = k7[m][d];
            }
    };

    long long obs_calls = 0, saved = 0, n_acc = 0, n_rej = 0, attempts = 0;
    long long status = kDopri5Ok;

    // One loop for the three controls so that the attempt is expanded once.
    // control 1 runs from T0 to T1.  control 0 walks the observation grid
    // interval by interval (gi).  control 2 makes one integrate_const call
    // per segment [s0, s1] (the whole window, or the trajectory's own list):
    // the observer at s0 and a fresh stepper there (k1 = f(x, s0), step hint
    // dt), then interval k from s0 + k dt to (s0 + k dt) + dt, observed at
    // s0 + (k + 1) dt, with the step size carried from interval to interval
    // and k1 taken again at the interval start (a new stepper copy per
    // interval; FSAL's value is kept when the accepted time already equals
    // it or the drive is off, where both are the same number).
    const int nseg = control == 2 ? (A.segs ? A.n_segs[b] : 1) : 0;
    int sg = 0;
    double s0 = A.T0, s1 = A.T1;
    if (control == 2 && A.segs) {
        s0 = A.segs[(size_t(b) * A.max_seg) * 2];
        s1 = A.segs[(size_t(b) * A.max_seg) * 2 + 1];
    }
    double t = s0, h = A.dt, tcur = s0, tend = A.T1;
    long long kk = 0;  // control 2: interval index within the segment
    int gi = 0, fails = 0;
    // pending evaluation k1 = f(x, te) of the state in bufA, with an
    // observer call when eobs
    bool need_eval = true, eobs = true;
    double te = s0;
    bool fresh = true;  // control 2: at the start of interval kk
    while (true) {
        if (need_eval) {
            const bool ob = eobs && (obs_calls % si) == 0;
            const double fl = drive_on && lane < 2 ? lane_drive(te) : 0.0;
            rhs(bufA, drive_on ? readlane_f64(fl, 0) : 0.0, drive_on ? readlane_f64(fl, 1) : 0.0, k1, ob);
            __syncthreads();
            if (ob) publish(saved++, te);
            if (eobs) ++obs_calls;
            need_eval = false;
        }
        double t7, tobs;
        bool ends;
        if (control == 0) {
            if (gi + 1 >= A.n_grid) break;
            t = A.tgrid[gi];
            h = A.dt;
            t7 = A.tgrid[gi + 1];
            tobs = t7;
            ends = true;
        } else {
            if (control == 1) {
                if (!less_with_sign(t, A.T1)) break;
                tend = A.T1;
            } else if (fresh) {
                if (!less_eq_with_sign(tcur + A.dt, s1)) {
                    // segment done: the next integrate_const call
                    if (++sg >= nseg) break;
                    s0 = A.segs[(size_t(b) * A.max_seg + sg) * 2];
                    s1 = A.segs[(size_t(b) * A.max_seg + sg) * 2 + 1];
                    kk = 0;
                    t = tcur = te = s0;
                    h = A.dt;
                    need_eval = eobs = true;
                    continue;
                }
                tend = tcur + A.dt;
                if (!less_with_sign(tcur, tend)) {
                    // an interval too short to step: observed unchanged
                    ++kk;
                    t = tcur = te = grid_time(s0, kk, A.dt);
                    need_eval = eobs = true;
                    continue;
                }
                fresh = false;
                fails = 0;
                if (drive_on && t != tcur) {
                    t = te = tcur;
                    need_eval = true;
                    eobs = false;
                    continue;
                }
                t = tcur;
            }
            if (fails == 0 && less_with_sign(tend, t + h)) h = tend - t;
            if (!(h > 0.0) || !(t + h > t)) {
                status = kDopri5BadStep;
                break;
            }
            t7 = t + h;
            tobs = control == 1 ? t7 : grid_time(s0, kk + 1, A.dt);
            // the observer fires when this step, if accepted, ends the run / interval
            ends = control == 1 || !less_with_sign(t + h, tend);
        }
        if (attempts >= A.max_attempts) {
            status = kDopri5AttemptCap;
            break;
        }
        ++attempts;
        const bool obs = ends && (obs_calls % si) == 0;
        const double err = attempt(t, h, t7, obs);
        if (control != 0 && !(err <= 1.0)) {
            h = h * smax(0.9 * pow(err, -1.0 / 3.0), 0.2);
            ++n_rej;
            if (++fails >= kMaxRejections) {
                status = kDopri5StepFailed;
                break;
            }
            continue;
        }
        fails = 0;
        accept();
        if (control != 0) {
            t = t + h;
            if (err < 0.5) h = h * (0.9 * pow(smax(1.0 / 3125.0, err), -1.0 / 5.0));
        }
        ++n_acc;
        if (ends) {
            if (obs) publish(saved++, tobs);
            ++obs_calls;
            if (control == 0) ++gi;
            if (control == 2) {
                ++kk;
                tcur = tobs;
                fresh = true;
            }
        }
    }

    if (tid == 0) {
        long long* c = A.counters + size_t(b) * 4;
        c[0] = saved;
        c[1] = n_acc;
        c[2] = n_rej;
        c[3] = status;
    }
    if (A.final_states) {
        double* out = A.final_states + size_t(b) * L0.N * 3;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = site(m);
            if (act[m]) {
                const int cell = p - atm[m] * L0.npad;
                const size_t i = size_t(cell) * L0.n_atoms + atm[m];
                out[3 * i] = x[m][0];
                out[3 * i + 1] = x[m][1];
                out[3 * i + 2] = x[m][2];
            }
        }
    }
}

namespace {

template <int SPT, int ZMAX, bool UNI, bool ATOM_MAJOR = false>
hipError_t launch_d(const LatDev& L, const Dopri5Args& A, int block, size_t lds, hipStream_t st) {
    auto kern = k_dopri5<SPT, ZMAX, UNI, ATOM_MAJOR>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(A.B), dim3(block), lds, st, L, A);
    return hipGetLastError();
}

template <int SPT, int ZMAX>
hipError_t launch_d_uni(const LatDev& L, const Dopri5Args& A, int block, size_t lds, hipStream_t st) {
    if (L.uniform && L.npad % (kWave * SPT) == 0 && block == L.n_atoms * (L.npad / (kWave * SPT)) * kWave)
        return launch_d<SPT, ZMAX, true, true>(L, A, block, lds, st);
    return L.uniform ? launch_d<SPT, ZMAX, true>(L, A, block, lds, st)
                     : launch_d<SPT, ZMAX, false>(L, A, block, lds, st);
}

template <int SPT>
hipError_t launch_d_z(const LatDev& L, const Dopri5Args& A, int block, size_t lds, hipStream_t st) {
    if (L.zmax <= 3) return launch_d_uni<SPT, 3>(L, A, block, lds, st);
    if (L.zmax <= 6) return launch_d_uni<SPT, 6>(L, A, block, lds, st);
    if (L.zmax <= 12) return launch_d_uni<SPT, 12>(L, A, block, lds, st);
    return hipErrorInvalidValue;
}

}  // namespace

int dopri5_sites_per_thread(const LatDev& L) {
    if (L.NP <= dopri5_max_block<1>()) return 1;
    if ((L.NP + 1) / 2 <= dopri5_max_block<2>()) return 2;
    if ((L.NP + 2) / 3 <= dopri5_max_block<3>()) return 3;
    if ((L.NP + 3) / 4 <= dopri5_max_block<4>()) return 4;
    return 0;
}

int dopri5_block(const LatDev& L) {
    const int spt = dopri5_sites_per_thread(L);
    if (spt == 0) return 0;
    return ((L.NP + spt - 1) / spt + kWave - 1) / kWave * kWave;
}

size_t dopri5_lds_bytes(const LatDev& L, int block) {
    return sizeof(double) * (size_t(6) * L.NP + size_t(block / kWave) * 11);
}

hipError_t launch_dopri5(const LatDev& L, const Dopri5Args& A, hipStream_t st) {
    const int spt = dopri5_sites_per_thread(L);
    const int block = dopri5_block(L);
    if (spt == 0 || block == 0 || L.zmax > 12) return hipErrorInvalidValue;
    const size_t lds = dopri5_lds_bytes(L, block);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    switch (spt) {
        case 1: return launch_d_z<1>(L, A, block, lds, st);
        case 2: return launch_d_z<2>(L, A, block, lds, st);
        case 3: return launch_d_z<3>(L, A, block, lds, st);
        case 4: return launch_d_z<4>(L, A, block, lds, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace csl
