// csl_persistent.hip — the persistent batched RK4 kernel.
//
// Compiled once per sites-per-thread value (CSL_SPT, set by the build) so the
// template instantiations build in parallel; each object defines
// launch_persistent_spt<CSL_SPT>.
//
// One workgroup per trajectory; the whole lattice state of that trajectory
// lives in LDS, each thread owns SPT sites and keeps their RK4 state x and
// accumulator in VGPRs; all steps run inside one launch; M(t) and E(t) are
// reduced in-kernel (wave64 DPP sums + LDS) every save_interval steps.
// Replaces the per-step launch sequence of integrate_gpu_batched_internal
// (lattice_gpu.cu:455-577).
//
// FIELD_ONLY: the on-site part of the local field is a per-sublattice field
// with no anisotropy matrix and every site has ZMAX bonds (uniform tables,
// site_uniform, !has_onsite, full_coord).  The runtime branches between
// per-site / per-sublattice fields, the onsite product and the per-bond
// count tests then disappear at compile time, with the scalar registers and
// lane masks they held (319 -> 251 ms on the 36x36 Kitaev batch).  Chosen by
// the launcher for the double-buffered, stage-unrolled shapes.
//
// LDS state layout (one buffer; two when DBUF): the x,y components of every
// site packed as a 16-byte pair, then the z plane, so a neighbour gather is
// one ds_read_b128 plus one ds_read_b64:
//     [ (x,y)[0..NP) | z[0..NP) ]          3*NP doubles per buffer
#include <type_traits>

#include "csl_device.hpp"
#include "csl_lds.hpp"

#ifndef CSL_SPT
#error "CSL_SPT (sites per thread) must be defined by the build"
#endif

namespace csl {

namespace {
// e_i = 1/2 S.(H - B) with the field of the site's sublattice
__device__ __forceinline__ double site_energy_atom(const LatDev& L, int atom, const double S[3],
                                                   const double H[3]) {
    cptr<double> Bu = as_const(L.field) + atom * 3;
    return 0.5 * (S[0] * (H[0] - Bu[0]) + S[1] * (H[1] - Bu[1]) + S[2] * (H[2] - Bu[2]));
}
}  // namespace

// ATOM_MAJOR (FIELD_ONLY shapes of n_atoms * waves_per_atom waves whose
// waves can hold the npad / 64 units of 64 cells of a sublattice, at most
// SPT each): wave w owns units [start, start + units) of sublattice atom(w)
// (atom_major_tile), site m < units of lane l at atom(w) * npad + (start +
// m) * 64 + l, so all owned sites share one atom and its bond matrices are
// loaded once per bond for all of them (instead of once per site slot).
// MULTI: every trajectory carries its own lattice (A.lats[A.lat_of[b]]),
// so the LDS layout and the site ownership are set up per piece.
template <int SPT, int ZMAX, bool UNI, bool DBUF, bool UNROLL, bool FIELD_ONLY, bool ATOM_MAJOR = false,
          bool MULTI = false>
__global__ void __launch_bounds__(persistent_max_block<SPT>()) k_persistent(LatDev Lk, PersistArgs A) {
    extern __shared__ double smem[];
    int NP = Lk.NP;
    LdsState bufA = lds_state(smem, NP);
    LdsState bufB = DBUF ? lds_state(smem + 3 * NP, NP) : bufA;
    double* drv = smem + (DBUF ? 6 : 3) * NP;  // [2 parity][3 stage times][2 pulses]
    double* red = drv + 12;                    // [nwaves][10]
    // PAIR (the FIELD_ONLY + ATOM_MAJOR shapes, always double-buffered): the
    // two buffers of a site share one 48-byte record [x y | x' y' | z z'],
    // so every access to site q of either buffer is one address (48 q)
    // plus an immediate offset: one address register per gathered
    // neighbour instead of four.  Other shapes keep the planar LdsState
    // buffers.
    constexpr bool PAIR = FIELD_ONLY && ATOM_MAJOR && DBUF;
    auto lds_ld = [&](int which, int q, double S[3]) {
        if constexpr (PAIR) {
            const double* r = smem + 6 * q;
            const double2 v = *reinterpret_cast<const double2*>(r + 2 * which);
            S[0] = v.x;
            S[1] = v.y;
            S[2] = r[4 + which];
        } else {
            (which ? bufB : bufA).load(q, S);
        }
    };
    auto lds_st = [&](int which, int q, const double S[3]) {
        if constexpr (PAIR) {
            double* r = smem + 6 * q;
            *reinterpret_cast<double2*>(r + 2 * which) = make_double2(S[0], S[1]);
            r[4 + which] = S[2];
        } else {
            (which ? bufB : bufA).store(q, S);
        }
    };

    const int T = blockDim.x;
    const int tid = threadIdx.x;
    const int wave = uniform(tid >> 6);
    const int lane = tid & (kWave - 1);
    const int nwaves = T >> 6;
    const double dt = A.dt;

    // per-thread ownership: sites p = tid + m*T, or the atom-major mapping
    // (the same for every trajectory)
    const int wpa = ATOM_MAJOR ? nwaves / Lk.n_atoms : 1;  // waves per atom
    const int atom_w = ATOM_MAJOR ? wave / wpa : 0;
    // atom-major: this wave owns units [wstart, wstart + wunits) of its
    // sublattice (slots m >= wunits own nothing)
    int wstart = 0, wunits = SPT;
    if constexpr (ATOM_MAJOR) {
        atom_major_tile(nwaves, Lk.n_atoms, Lk.npad / kWave, wave, wstart, wunits);
        wstart = uniform(wstart);
        wunits = uniform(wunits);
    }
    const int pbase = ATOM_MAJOR ? atom_w * Lk.npad + wstart * kWave + lane : tid;
    auto site = [&](int m) { return ATOM_MAJOR ? pbase + m * kWave : tid + m * T; };
    double x[SPT][3], acc[SPT][3];
    int nb[SPT][ZMAX];
    int cnt[SPT], atm[SPT];
    bool act[SPT];
    auto own = [&](const LatDev& Lo) {
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = site(m);
            const int pw = ATOM_MAJOR ? atom_w * Lo.npad : wave * kWave + m * T;  // wave-uniform
            atm[m] = ATOM_MAJOR ? atom_w : uniform(pw < NP ? pw / Lo.npad : 0);
            act[m] = m < wunits && active_site(Lo, p);
            const int c = UNI ? uniform(pw < NP ? as_const(Lo.cnt_atom)[atm[m]] : 0) : (p < NP ? Lo.cnt[p] : 0);
            cnt[m] = c;
#pragma unroll
            for (int k = 0; k < ZMAX; ++k) nb[m][k] = (act[m] && k < c) ? Lo.col[size_t(k) * NP + p] : 0;
        }
    };
    if (!MULTI) own(Lk);

    const int n_pieces = A.sched ? A.max_pieces : 1;
    for (int piece = 0; piece < n_pieces; ++piece) {
    int b = blockIdx.x, s_begin = 0, s_end = A.n_steps, slot_in = -1, slot_out = -1;
    if (A.sched) {
        const int* r = A.sched + (size_t(blockIdx.x) * A.max_pieces + piece) * 5;
        b = r[0];
        s_begin = r[1];
        s_end = r[2];
        slot_in = r[3];
        slot_out = r[4];
        if (b < 0) continue;  // uniform
    }
    // this trajectory's Hamiltonian tables
    LatDev L0 = MULTI ? A.lats[A.lat_of[b]] : Lk;
    if (!MULTI) {
        L0.mats += size_t(b) * A.mats_stride;
        L0.field += size_t(b) * A.field_stride;
    }
    const LatDev& L = L0;
    if (MULTI) {
        __syncthreads();  // the previous piece is done with the LDS buffers and red[]
        NP = L.NP;
        bufA = lds_state(smem, NP);
        bufB = DBUF ? lds_state(smem + 3 * NP, NP) : bufA;
        drv = smem + (DBUF ? 6 : 3) * NP;
        red = drv + 12;
        own(L);
    }

    const double tp1 = A.t_pulse[2 * b], tp2 = A.t_pulse[2 * b + 1];
    const double am1 = A.amp[2 * b], am2 = A.amp[2 * b + 1];
    const bool drive_on = (am1 != 0.0) || (am2 != 0.0);
    const double alpha = L.alpha;

    __syncthreads();  // the previous piece is done with the LDS buffers
    if (slot_in < 0) {
        // initial state -> LDS (device order)
        const double* init = A.init + (MULTI ? A.init_off[b] : size_t(b) * A.init_stride);
        for (int p = tid; p < NP; p += T) {
            const int atom = p / L.npad, cell = p - atom * L.npad;
            double v[3] = {0.0, 0.0, 0.0};
            if (cell < L.n_cells) {
                const size_t i = size_t(cell) * L.n_atoms + atom;
                v[0] = init[3 * i];
                v[1] = init[3 * i + 1];
                v[2] = init[3 * i + 2];
            }
            lds_st(0, p, v);
        }
    } else {
        // the state at step s_begin, saved by the workgroup that ran the
        // trajectory's first steps (it does so before anything else, so the
        // wait is short); a bounded wait that gives up flags the launch
        if (tid == 0) {
            unsigned ok = 0;
            for (long it = 0; it < (1L << 24); ++it) {
                ok = __hip_atomic_load(A.flags + slot_in, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                if (ok) break;
                __builtin_amdgcn_s_sleep(8);
            }
            if (!ok) __hip_atomic_store(A.status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        __atomic_thread_fence(__ATOMIC_ACQUIRE);
        const double* sv = A.saved + size_t(slot_in) * 3 * size_t(Lk.NP);
        for (int p = tid; p < NP; p += T) {
            const double v[3] = {sv[3 * p], sv[3 * p + 1], sv[3 * p + 2]};
            lds_st(0, p, v);
        }
    }
#pragma unroll
    for (int m = 0; m < SPT; ++m) acc[m][0] = acc[m][1] = acc[m][2] = 0.0;
    __syncthreads();
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        const int p = site(m);
        if (act[m]) {
            lds_ld(0, p, x[m]);
        } else {
            x[m][0] = x[m][1] = x[m][2] = 0.0;
        }
    }

    const double b16 = (1.0 / 6.0) * dt, b13 = (1.0 / 3.0) * dt;
    const double inv_n = 1.0 / double(L.N);
    const bool want_e = A.energy != nullptr;
    // The last wave owns the fewest real sites (padding sits at the end), so
    // it evaluates the drive envelopes.
    const int drive_wave = nwaves - 1;

    // Drive envelopes of the three stage times of step n (t, t+dt/2, t+dt)
    // for both pulses, written by 6 lanes of one wave into parity slot n&1.
    // Slot n&1 is written at the top of step n-1 (step 0: before the loop)
    // and read during step n; the barriers of step n-1 separate the two.
    auto write_drive = [&](int n) {
        if (drive_on && wave == drive_wave && lane < 6) {
            const int st = lane >> 1, pl = lane & 1;
            const double t = A.tgrid[n];
            const double ts = st == 0 ? t : (st == 1 ? t + 0.5 * dt : t + 1.0 * dt);
            drv[(n & 1) * 6 + lane] = drive_factor(pl ? am2 : am1, ts, pl ? tp2 : tp1, L.width, L.freq);
        }
    };
    // wave-reduce the observables of the owned sites and publish them in red[]
    auto reduce_obs = [&](const LatDev& Lr, const double* e) {
        if constexpr (ATOM_MAJOR) {
            // every owned site of the wave is in sublattice atom_w: the frame
            // rotation and the staggered sign act once on the wave's sum
            double sx[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                if (act[m]) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) sx[d] += x[m][d];
                    sx[3] += e[m];
                }
            }
            double v[4];
#pragma unroll
            for (int d = 0; d < 3; ++d) v[d] = wave_sum_lane63(sx[d]);
            v[3] = want_e ? wave_sum_lane63(sx[3]) : 0.0;
            if (lane == kWave - 1) {
                cptr<double> R = as_const(Lr.frames) + atom_w * 9;
                const double sg = as_const(Lr.afm)[atom_w];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double g = R[d * 3 + 0] * v[0] + R[d * 3 + 1] * v[1] + R[d * 3 + 2] * v[2];
                    red[wave * 10 + d] = sg * g;
                    red[wave * 10 + 3 + d] = v[d];
                    red[wave * 10 + 6 + d] = g;
                }
                red[wave * 10 + 9] = v[3];
            }
            return;
        }
        double part[10];
#pragma unroll
        for (int j = 0; j < 10; ++j) part[j] = 0.0;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            if (act[m]) {
                const int am = ATOM_MAJOR ? atom_w : atm[m];
                cptr<double> R = as_const(Lr.frames) + am * 9;
                const double sg = as_const(Lr.afm)[am];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double g = R[d * 3 + 0] * x[m][0] + R[d * 3 + 1] * x[m][1] + R[d * 3 + 2] * x[m][2];
                    part[d] += sg * g;
                    part[3 + d] += x[m][d];
                    part[6 + d] += g;
                }
                part[9] += e[m];
            }
        }
#pragma unroll
        for (int j = 0; j < 10; ++j) {
            if (j == 9 && !want_e) break;
            const double v = wave_sum_lane63(part[j]);
            if (lane == kWave - 1) red[wave * 10 + j] = v;
        }
    };
    auto publish = [&](int ri) {
        if (wave == 0 && lane < 10) {
            double v = 0.0;
            for (int w = 0; w < nwaves; ++w) v += red[w * 10 + lane];
            if (lane < 9) {
                if (A.mags) A.mags[(size_t(ri) * A.B + b) * 9 + lane] = v * inv_n;
            } else if (A.energy) {
                A.energy[size_t(ri) * A.B + b] = v * inv_n;
            }
        }
    };

    write_drive(s_begin);
    __syncthreads();

    // One RK4 stage of step n.  Called with a runtime stage index (one copy of
    // the code, selects on s) or with compile-time constants (UNROLL: the four
    // stages are specialised and the selects disappear).
    auto stage = [&](auto s_, int par, bool rec, int ri) {
        const int s = static_cast<int>(s_);
        if constexpr (PAIR) {
            // Every owned site of the wave is in sublattice atom_w and has
            // ZMAX bonds with the same matrices: the sites are advanced
            // together, bond by bond, so each bond matrix is read once per
            // wave and stage (scalar loads shared by the SPT sites) and the
            // gathers of the SPT sites are in flight at once.  There is no
            // per-site branch: padding slots (act = false) hold zero spins,
            // gather site 0 and store their (zero) stage values into padding
            // slots that no real site reads; only the records and the state
            // outputs skip them.  The arithmetic per site is the general
            // path's: H = -B + sum_k J_k S_k (one FMA chain per component),
            // the energy from the undriven field, then H -= f1 d1 + f2 d2.
            const LatDev L = refresh(L0);
            const int cur = s & 1, nxt = 1 - (s & 1);
            const int st = s == 0 ? 0 : (s == 3 ? 2 : 1);
            const double f1 = drive_on ? drv[par * 6 + st * 2 + 0] : 0.0;
            const double f2 = drive_on ? drv[par * 6 + st * 2 + 1] : 0.0;
            const double cdt = (s == 2) ? 1.0 * dt : 0.5 * dt;
            const double bdt = (s == 0 || s == 3) ? b16 : b13;
            const bool obs = (s == 0) && rec;
            cptr<double> Bu = as_const(L.field) + atom_w * 3;
            // drive vector of the wave's sublattice: the same sum per
            // component as sub_drive, formed once per wave
            const bool drive_now = drive_on && (f1 != 0.0 || f2 != 0.0);
            double D[3] = {0.0, 0.0, 0.0};
            if (drive_now) {
                cptr<double> d0 = as_const(L.fdrive) + atom_w * 3;
                cptr<double> d1 = as_const(L.fdrive) + (L.n_atoms + atom_w) * 3;
#pragma unroll
                for (int d = 0; d < 3; ++d) D[d] = d0[d] * f1 + d1[d] * f2;
            }
            // bond by bond: one bond matrix (scalar loads) live at a time and
            // the gathers of the wave's sites for that bond in flight
            // together; every site's field is still -B + J_0 S_0 + J_1 S_1
            // + ... in this order (one FMA chain per component)
            double H[SPT][3];
#pragma unroll
            for (int m = 0; m < SPT; ++m)
#pragma unroll
                for (int d = 0; d < 3; ++d) H[m][d] = -Bu[d];
#pragma unroll
            for (int k = 0; k < ZMAX; ++k) {
                cptr<double> J = as_const(L.mats) + (atom_w * L.zmax + k) * 9;
                double Jk[9];
#pragma unroll
                for (int q = 0; q < 9; ++q) Jk[q] = J[q];
#pragma unroll
                for (int m = 0; m < SPT; ++m) {
                    if (m >= wunits) continue;  // wave-uniform: the slot owns no sites
                    double Sp[3];
                    lds_ld(cur, nb[m][k], Sp);
#pragma unroll
                    for (int d = 0; d < 3; ++d)
#pragma unroll
                        for (int e = 0; e < 3; ++e) H[m][d] = fma(Jk[d * 3 + e], Sp[e], H[m][d]);
                }
            }
            double e_obs[SPT];
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                e_obs[m] = 0.0;
                if (m >= wunits) continue;  // wave-uniform: the slot owns no sites
                double S[3];
                // stage 0 reads x itself (the LDS copy holds the same bits)
                if (s == 0) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) S[d] = x[m][d];
                } else {
                    lds_ld(cur, site(m), S);
                }
                e_obs[m] = (obs && want_e) ? site_energy_atom(L, atom_w, S, H[m]) : 0.0;
                if (drive_now) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) H[m][d] -= D[d];
                }
                double k[3];
                ll_su2_guarded(H[m], S, alpha, k);
                double y[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    if (s == 0) acc[m][d] = x[m][d] + bdt * k[d];
                    else if (s < 3) acc[m][d] = acc[m][d] + bdt * k[d];
                    else x[m][d] = acc[m][d] + bdt * k[d];
                    y[d] = (s < 3) ? x[m][d] + cdt * k[d] : x[m][d];
                }
                lds_st(nxt, site(m), y);
            }
            if (obs) reduce_obs(L, e_obs);
            __syncthreads();
            if (obs) publish(ri);
        } else {
            const LatDev L = refresh(L0);
            const LdsState cur = (DBUF && (s & 1)) ? bufB : bufA;
            const LdsState nxt = DBUF ? ((s & 1) ? bufA : bufB) : bufA;
            const int st = s == 0 ? 0 : (s == 3 ? 2 : 1);
            const double f1 = drive_on ? drv[par * 6 + st * 2 + 0] : 0.0;
            const double f2 = drive_on ? drv[par * 6 + st * 2 + 1] : 0.0;
            const double cdt = (s == 2) ? 1.0 * dt : 0.5 * dt;
            const double bdt = (s == 0 || s == 3) ? b16 : b13;
            const bool obs = (s == 0) && rec;
            double e[SPT];
            double yn[DBUF ? 1 : SPT][3];
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int p = site(m);
                e[m] = 0.0;
                if (act[m]) {
                    const int atom = ATOM_MAJOR ? atom_w : atm[m];
                    double S[3];
                    cur.load(p, S);
                    double H[3];
                    local_field_lds<ZMAX, UNI, FIELD_ONLY>(L, cur, p, atom, nb[m], cnt[m], S, H);
                    if (obs && want_e)
                        e[m] = FIELD_ONLY ? site_energy_atom(L, atom, S, H) : site_energy(L, atom, p, S, H);
                    if (drive_on && (f1 != 0.0 || f2 != 0.0)) sub_drive(L, atom, f1, f2, H);
                    double k[3];
                    ll_su2(H, S, alpha, k);
                    double y[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        if (s == 0) acc[m][d] = x[m][d] + bdt * k[d];
                        else if (s < 3) acc[m][d] = acc[m][d] + bdt * k[d];
                        else x[m][d] = acc[m][d] + bdt * k[d];
                        y[d] = (s < 3) ? x[m][d] + cdt * k[d] : x[m][d];
                    }
                    if (DBUF) {
                        nxt.store(p, y);
                    } else {
#pragma unroll
                        for (int d = 0; d < 3; ++d) yn[DBUF ? 0 : m][d] = y[d];
                    }
                }
            }
            // stage 0: x still holds the state at t_n (the record's state)
            if (obs) reduce_obs(L, e);
            if (!DBUF) {
                __syncthreads();
#pragma unroll
                for (int m = 0; m < SPT; ++m) {
                    const int p = site(m);
                    if (act[m]) nxt.store(p, yn[DBUF ? 0 : m]);
                }
            }
            __syncthreads();
            if (obs) publish(ri);
        }
    };

    for (int n = s_begin; n < s_end; ++n) {
        const int par = n & 1;
        const bool rec = (n % A.save_interval) == 0;
        const int ri = n / A.save_interval;
        if (n + 1 < s_end) write_drive(n + 1);
        if constexpr (UNROLL) {
            stage(std::integral_constant<int, 0>{}, par, rec, ri);
            stage(std::integral_constant<int, 1>{}, par, rec, ri);
            stage(std::integral_constant<int, 2>{}, par, rec, ri);
            stage(std::integral_constant<int, 3>{}, par, rec, ri);
        } else {
#pragma unroll 1
            for (int s = 0; s < 4; ++s) stage(s, par, rec, ri);
        }
    }

    if (slot_out >= 0) {
        // hand the state at s_end to the piece that continues the trajectory
        double* sv = A.saved + size_t(slot_out) * 3 * size_t(Lk.NP);
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = site(m);
            if (p < NP && m < wunits) {
#pragma unroll
                for (int d = 0; d < 3; ++d) sv[3 * p + d] = act[m] ? x[m][d] : 0.0;
            }
        }
        __threadfence();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(A.flags + slot_out, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    // final record: state after the last step (lattice_gpu.cu:567-568)
    if (s_end == A.n_steps) {
        const LatDev L = refresh(L0);
        double e[SPT];
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = site(m);
            e[m] = 0.0;
            if (act[m]) {
                double H[3];
                if constexpr (PAIR) {
                    cptr<double> Bu = as_const(L.field) + atom_w * 3;
#pragma unroll
                    for (int d = 0; d < 3; ++d) H[d] = -Bu[d];
#pragma unroll
                    for (int k = 0; k < ZMAX; ++k) {
                        double Sp[3];
                        lds_ld(0, nb[m][k], Sp);
                        add_bond_fused(as_const(L.mats) + (atom_w * L.zmax + k) * 9, Sp, H);
                    }
                } else {
                    local_field_lds<ZMAX, UNI, FIELD_ONLY>(L, bufA, p, atm[m], nb[m], cnt[m], x[m], H);
                }
                e[m] = FIELD_ONLY ? site_energy_atom(L, atm[m], x[m], H) : site_energy(L, atm[m], p, x[m], H);
            }
        }
        reduce_obs(L, e);
        __syncthreads();
        publish(A.n_rec - 1);
        if (A.final_states) {
            double* out = A.final_states + size_t(b) * L.N * 3;
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int p = site(m);
                if (act[m]) {
                    const int cell = p - atm[m] * L.npad;
                    const size_t i = size_t(cell) * L.n_atoms + atm[m];
                    out[3 * i] = x[m][0];
                    out[3 * i + 1] = x[m][1];
                    out[3 * i + 2] = x[m][2];
                }
            }
        }
    }
    }  // pieces
}

namespace {

template <int SPT, int ZMAX, bool UNI, bool DBUF, bool UNROLL, bool FIELD_ONLY = false, bool ATOM_MAJOR = false,
          bool MULTI = false>
hipError_t launch_p(const LatDev& L, const PersistArgs& A, int block, size_t lds, hipStream_t st, int* occ) {
    auto kern = k_persistent<SPT, ZMAX, UNI, DBUF, UNROLL, FIELD_ONLY, ATOM_MAJOR, MULTI>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    if (occ) return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, kern, block, lds);
    hipLaunchKernelGGL(kern, dim3(A.sched ? A.grid : A.B), dim3(block), lds, st, L, A);
    return hipGetLastError();
}

template <int SPT, int ZMAX, bool UNI>
hipError_t launch_p_db(const LatDev& L, const PersistArgs& A, int block, int flags, size_t lds,
                       hipStream_t st, int* occ) {
    const bool dbuf = flags & kPersistDoubleBuffer, unroll = flags & kPersistUnrollStages;
    if (flags & kPersistMulti) return launch_p<SPT, ZMAX, UNI, true, true, false, false, true>(L, A, block, lds, st, occ);
    const int variant = persistent_variant(L, block, SPT, flags);
    if (UNI && (variant & kVariantFieldOnly)) {
        if (variant & kVariantAtomMajor) return launch_p<SPT, ZMAX, UNI, true, true, true, true>(L, A, block, lds, st, occ);
        return launch_p<SPT, ZMAX, UNI, true, true, true>(L, A, block, lds, st, occ);
    }
    if (unroll)
        return dbuf ? launch_p<SPT, ZMAX, UNI, true, true>(L, A, block, lds, st, occ)
                    : launch_p<SPT, ZMAX, UNI, false, true>(L, A, block, lds, st, occ);
    return dbuf ? launch_p<SPT, ZMAX, UNI, true, false>(L, A, block, lds, st, occ)
                : launch_p<SPT, ZMAX, UNI, false, false>(L, A, block, lds, st, occ);
}

template <int SPT, int ZMAX>
hipError_t launch_p_uni(const LatDev& L, const PersistArgs& A, int block, int flags, size_t lds,
                        hipStream_t st, int* occ) {
    return L.uniform ? launch_p_db<SPT, ZMAX, true>(L, A, block, flags, lds, st, occ)
                     : launch_p_db<SPT, ZMAX, false>(L, A, block, flags, lds, st, occ);
}

}  // namespace

template <>
hipError_t launch_persistent_spt<CSL_SPT>(const LatDev& L, const PersistArgs& A, int block, int flags,
                                          size_t lds, hipStream_t st, int* occ) {
    if (L.zmax <= 3) return launch_p_uni<CSL_SPT, 3>(L, A, block, flags, lds, st, occ);
    if (L.zmax <= 6) return launch_p_uni<CSL_SPT, 6>(L, A, block, flags, lds, st, occ);
    if (L.zmax <= 12) return launch_p_uni<CSL_SPT, 12>(L, A, block, flags, lds, st, occ);
    return hipErrorInvalidValue;
}

}  // namespace csl
