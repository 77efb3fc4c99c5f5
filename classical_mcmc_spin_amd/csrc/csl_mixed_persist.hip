// csl_mixed_persist.hip — persistent batched RK4 kernel of the mixed
// SU(2) + SU(3) lattice: one workgroup integrates one trajectory for the
// whole time grid (mixed_lattice_gpu.cu:278-347 launches one kernel per RK
// stage instead), so the per-stage state never crosses HBM.
//
// On-chip budget of one trajectory of the 8x8x8 TmFeO3-like cell (2048 SU(2)
// + 2048 SU(3) sites, 176 KB of stage values): RK4 needs the step-start
// state x, the accumulator and the stage values.  The stage values are what
// the neighbours gather, so they go to LDS — the SU(2) planes and SU(3)
// components 0..5 (144 KB); SU(3) components 6 and 7 (32 KB) go to a
// per-workgroup global buffer with two parities that stays in L2.  The
// accumulators of a thread's sites live in VGPRs; x lives in the trajectory's
// own slot of the global state array (read in stages 1 and 2, written in
// stage 3: three 176 KB passes per step instead of the stage kernels' sixteen
// plane passes).
//
// Per stage (stage values y_s in LDS + G[s&1]):  every thread evaluates its
// SU(2) and SU(3) sites from y_s, updates the accumulators, writes SU(3)
// components 6, 7 of y_{s+1} straight to G[(s+1)&1] and keeps the rest of
// y_{s+1} in VGPRs -> barrier (every read of y_s done) -> y_{s+1} into LDS ->
// barrier.  Threads own sites p = tid + j*512 of each species (sublattice
// major, npad a multiple of 64), so a wave covers 64 sites of one sublattice
// and reads that sublattice's bond matrices, field, onsite matrix, frame and
// drive directions with scalar loads (translation-invariant tables only).
// Records (both species' magnetisations and E/(N2+N3), the 34 values of the
// stage kernels' record) are reduced inside the workgroup and written
// directly, at stage 0 of every save_interval-th step and after the last step.
#include <algorithm>
#include <type_traits>

#include "csl_device.hpp"
#include "csl_mixed_device.hpp"

namespace csl {

namespace {

// a double of lane l of this wave, in every lane
__device__ __forceinline__ double lane_double(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// The lattice description lives in device memory and is read through a
// pointer laundered by an opaque zero where it is used (once per site and
// phase): its fields and the table pointers are fetched with scalar loads at
// the point of use instead of occupying SGPRs for the whole kernel (the
// description is 416 bytes, more than the SGPR file).
using MP = cptr<MixDev>;
__device__ __forceinline__ MP fresh_mix(const MixDev* Mg) { return as_const(Mg) + opaque_zero(); }

// Stage values of one trajectory: SU(2) planes F[3][NP2] and SU(3)
// components 0..5 TL[6][NP3] in LDS, components 6, 7 in Gc[2][NP3].
struct MixStageView {
    const double* F;
    const double* TL;
    const double* Gc;
    int NP2, NP3;
    __device__ __forceinline__ double tm(int c, int q) const {
        return c < 6 ? TL[c * NP3 + q] : Gc[(c - 6) * NP3 + q];
    }
    __device__ __forceinline__ double fe(int c, int q) const { return F[c * NP2 + q]; }
};

// Undriven local field of SU(2) site p (sublattice a) from the stage values:
// -B + 2A S + sum J22 S_p + sum J23 l_p, the terms in the order of the stage
// kernels' field_su2 (uniform tables, per-sublattice field and onsite).
__device__ __forceinline__ void fe_field(const MixDev* Mg, const MixStageView& V, int a, int p, const double S[3],
                                         double H[3]) {
    const MP M = fresh_mix(Mg);
    const cptr<double> f = as_const(M->field2) + a * 3;
#pragma unroll
    for (int d = 0; d < 3; ++d) H[d] = -f[d];
    if (M->has_onsite2) {
        const cptr<double> A2 = as_const(M->onsite2) + a * 9;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double r = 0.0;
#pragma unroll
            for (int e = 0; e < 3; ++e) r += A2[d * 3 + e] * S[e];
            H[d] += r;
        }
    }
    // partner indices of every bond first (independent loads in flight),
    // then the gathers and products
    const int c22 = uniform(as_const(M->e22.cnt_atom)[a]);
    const int c23 = uniform(as_const(M->e23.cnt_atom)[a]);
    int q22[kMixZ22], q23[kMixZ23];
#pragma unroll
    for (int k = 0; k < kMixZ22; ++k)
        if (k < c22) q22[k] = M->e22.col[size_t(k) * V.NP2 + fresh(p)];
#pragma unroll
    for (int k = 0; k < kMixZ23; ++k)
        if (k < c23) q23[k] = M->e23.col[size_t(k) * V.NP2 + fresh(p)];
#pragma unroll
    for (int k = 0; k < kMixZ22; ++k) {
        if (k < c22) {
            double s[3];
#pragma unroll
            for (int e = 0; e < 3; ++e) s[e] = V.fe(e, q22[k]);
            const cptr<double> J = as_const(M->e22.mats) + (a * M->e22.zmax + k) * 9;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int e = 0; e < 3; ++e) H[d] = fma(J[d * 3 + e], s[e], H[d]);
        }
    }
#pragma unroll
    for (int k = 0; k < kMixZ23; ++k) {
        if (k < c23) {
            double s[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) s[e] = V.tm(e, q23[k]);
            const cptr<double> J = as_const(M->e23.mats) + (a * M->e23.zmax + k) * 24;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int e = 0; e < 8; ++e) H[d] = fma(J[d * 8 + e], s[e], H[d]);
        }
    }
}

// The same for an SU(3) site: -B + 2A l + sum J33 l_p + sum J32 S_p
// (field_su3); TT: the lattice has SU(3)-SU(3) bonds.
template <bool TT>
__device__ __forceinline__ void tm_field(const MixDev* Mg, const MixStageView& V, int a, int p, const double S[8],
                                         double H[8]) {
    const MP M = fresh_mix(Mg);
    const cptr<double> f = as_const(M->field3) + a * 8;
#pragma unroll
    for (int d = 0; d < 8; ++d) H[d] = -f[d];
    if (M->has_onsite3) {
        const cptr<double> A3 = as_const(M->onsite3) + a * 64;
#pragma unroll
        for (int d = 0; d < 8; ++d) {
            double r = 0.0;
#pragma unroll
            for (int e = 0; e < 8; ++e) r += A3[d * 8 + e] * S[e];
            H[d] += r;
        }
    }
    const int c33 = TT ? uniform(as_const(M->e33.cnt_atom)[a]) : 0;
    const int c32 = uniform(as_const(M->e32.cnt_atom)[a]);
    int q33[TT ? kMixZ33 : 1], q32[kMixZ32];
    if constexpr (TT) {
#pragma unroll
        for (int k = 0; k < kMixZ33; ++k)
            if (k < c33) q33[k] = M->e33.col[size_t(k) * V.NP3 + fresh(p)];
    }
#pragma unroll
    for (int k = 0; k < kMixZ32; ++k)
        if (k < c32) q32[k] = M->e32.col[size_t(k) * V.NP3 + fresh(p)];
    if constexpr (TT) {
#pragma unroll
        for (int k = 0; k < kMixZ33; ++k) {
            if (k < c33) {
                double s[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) s[e] = V.tm(e, q33[k]);
                const cptr<double> J = as_const(M->e33.mats) + (a * M->e33.zmax + k) * 64;
#pragma unroll
                for (int d = 0; d < 8; ++d)
#pragma unroll
                    for (int e = 0; e < 8; ++e) H[d] = fma(J[d * 8 + e], s[e], H[d]);
            }
        }
    }
    // the SU(3)->SU(2) matrices as transposes of the SU(2)->SU(3) ones when
    // the tables pair up (half the bond-matrix footprint in the scalar
    // cache: the SU(2) sites read the same matrices); same values, same order
    const int* jo = M->j32off;
    if (jo) {
        int off[kMixZ32];
#pragma unroll
        for (int k = 0; k < kMixZ32; ++k) off[k] = as_const(jo)[a * 8 + k];
#pragma unroll
        for (int k = 0; k < kMixZ32; ++k) {
            if (k < c32) {
                double s[3];
#pragma unroll
                for (int e = 0; e < 3; ++e) s[e] = V.fe(e, q32[k]);
                const cptr<double> J = as_const(M->e23.mats) + uniform(off[k]);
#pragma unroll
                for (int d = 0; d < 8; ++d)
#pragma unroll
                    for (int e = 0; e < 3; ++e) H[d] = fma(J[e * 8 + d], s[e], H[d]);
            }
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < kMixZ32; ++k) {
        if (k < c32) {
            double s[3];
#pragma unroll
            for (int e = 0; e < 3; ++e) s[e] = V.fe(e, q32[k]);
            const cptr<double> J = as_const(M->e32.mats) + (a * M->e32.zmax + k) * 24;
#pragma unroll
            for (int d = 0; d < 8; ++d)
#pragma unroll
                for (int e = 0; e < 3; ++e) H[d] = fma(J[d * 3 + e], s[e], H[d]);
        }
    }
}

// 0.5 S.(H - B) of a site: its share of the energy (H undriven)
template <int D>
__device__ __forceinline__ double site_energy_mixed(const double* field, int a, const double* S, const double* H) {
    const cptr<double> B = as_const(field) + a * D;
    double e = 0.0;
#pragma unroll
    for (int mu = 0; mu < D; ++mu) e += S[mu] * (H[mu] - B[mu]);
    return 0.5 * e;
}

// Magnetisation sums (record entries 0..32) of the stage values of this
// thread's sites, summed over the wave into red[wave][0..32].
template <int SPT, int T>
__device__ __forceinline__ void observe(MP M, const MixStageView& V, int tid, int lane, int wave,
                                        const int* at2, const int* at3, unsigned act2, unsigned act3, double* red) {
    double pt[33];
#pragma unroll
    for (int i = 0; i < 33; ++i) pt[i] = 0.0;
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        if (!((act2 >> j) & 1u)) continue;
        const int p = fresh(tid + j * T), a = at2[j];
        const double S[3] = {V.fe(0, p), V.fe(1, p), V.fe(2, p)};
        const cptr<double> R = as_const(M->frames2) + a * 9;
        const double sg = as_const(M->afm2)[a];
#pragma unroll
        for (int mu = 0; mu < 3; ++mu) {
            const double g = R[mu * 3] * S[0] + R[mu * 3 + 1] * S[1] + R[mu * 3 + 2] * S[2];
            pt[mu] += sg * g;
            pt[3 + mu] += S[mu];
            pt[6 + mu] += g;
        }
    }
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        if (!((act3 >> j) & 1u)) continue;
        const int p = fresh(tid + j * T), a = at3[j];
        const int cell = p - a * M->npad;
        double S[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) S[c] = V.tm(c, p);
        const cptr<double> R = as_const(M->frames3) + a * 64;
        const double sg = ((cell * M->A3 + a) & 1) ? -1.0 : 1.0;
#pragma unroll
        for (int mu = 0; mu < 8; ++mu) {
            double g = 0.0;
#pragma unroll
            for (int nu = 0; nu < 8; ++nu) g += R[mu * 8 + nu] * S[nu];
            pt[9 + mu] += sg * S[mu];
            pt[17 + mu] += S[mu];
            pt[25 + mu] += g;
        }
    }
#pragma unroll
    for (int i = 0; i < 33; ++i) {
        const double v = wave_sum(pt[i]);
        if (lane == 0) red[wave * kRec + i] = v;
    }
}

}  // namespace

template <int SPT, int T, bool TT>
__global__ void __launch_bounds__(T) k_mixed_persist(const MixDev* __restrict__ Mg, MixPersistArgs A) {
    const MP M0 = fresh_mix(Mg);
    extern __shared__ double smem[];
    constexpr int nwaves = T / kWave;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = uniform(tid >> 6);
    const int NP2 = M0->NP2, NP3 = M0->NP3;
    double* F = smem;                        // [3][NP2] SU(2) stage values
    double* TL = F + 3 * size_t(NP2);        // [6][NP3] SU(3) stage values, components 0..5
    double* red = TL + 6 * size_t(NP3);      // [nwaves][kRec] record sums of the waves
    double* Gw = A.G + size_t(blockIdx.x) * 4 * size_t(NP3);  // [2 parities][2][NP3]
    const size_t SZ = 3 * size_t(NP2) + 8 * size_t(NP3);
    const size_t o3 = 3 * size_t(NP2);
    const double dt = A.dt;
    const double b16 = (1.0 / 6.0) * dt, b13 = (1.0 / 3.0) * dt;

    // own sites of each species: p = tid + j*T, sublattice p / npad
    int at2[SPT], at3[SPT];
    unsigned act2 = 0, act3 = 0;
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        const int pw = wave * kWave + j * T;
        at2[j] = uniform(min(pw, NP2 - 1) / M0->npad);
        at3[j] = uniform(min(pw, NP3 - 1) / M0->npad);
        const int p = tid + j * T;
        if (p < NP2 && p - at2[j] * M0->npad < M0->n_cells) act2 |= 1u << j;
        if (p < NP3 && p - at3[j] * M0->npad < M0->n_cells) act3 |= 1u << j;
    }

    for (int piece = 0;; ++piece) {
        // this workgroup's next trajectory (steps [s_begin, s_end))
        int b = int(blockIdx.x) + piece * int(gridDim.x), s_begin = 0, s_end = A.n_steps, wait = -1, post = -1;
        if (A.sched) {
            if (piece >= A.max_pieces) break;
            const int* r = A.sched + (size_t(blockIdx.x) * A.max_pieces + piece) * 5;
            b = r[0];
            s_begin = r[1];
            s_end = r[2];
            wait = r[3];
            post = r[4];
            if (b < 0) break;  // unused rows come last
        } else if (b >= A.B) {
            break;
        }
        b += A.b0;
        double* X = A.X + size_t(b) * SZ;
        const bool drive = A.tp != nullptr;
        const double tpa = drive ? A.tp[2 * size_t(b)] : 0.0, tpb = drive ? A.tp[2 * size_t(b) + 1] : 0.0;
        const double tqa = drive ? A.tp3[2 * size_t(b)] : 0.0, tqb = drive ? A.tp3[2 * size_t(b) + 1] : 0.0;
        const double sca = (drive && A.scale) ? A.scale[2 * size_t(b)] : 1.0;
        const double scb = (drive && A.scale) ? A.scale[2 * size_t(b) + 1] : 1.0;

        if (wait >= 0) {
            // the state at s_begin, stored in X by the workgroup that ran
            // the trajectory's first steps as its first piece; a bounded
            // wait that gives up flags the launch
            if (tid == 0) {
                unsigned ok = 0;
                for (long it = 0; it < (1L << 24); ++it) {
                    ok = __hip_atomic_load(A.flags + wait, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                    if (ok) break;
                    __builtin_amdgcn_s_sleep(8);
                }
                if (!ok) __hip_atomic_store(A.status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
            __atomic_thread_fence(__ATOMIC_ACQUIRE);
        }
        // initial state: the stage values of stage 0 (LDS and G[0])
        __syncthreads();  // the previous trajectory is done with LDS and G
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
            const int p = tid + j * T;
            if ((act2 >> j) & 1u)
#pragma unroll
                for (int c = 0; c < 3; ++c) F[c * NP2 + p] = X[size_t(c) * NP2 + p];
            if ((act3 >> j) & 1u) {
#pragma unroll
                for (int c = 0; c < 6; ++c) TL[c * NP3 + p] = X[o3 + size_t(c) * NP3 + p];
                Gw[p] = X[o3 + size_t(6) * NP3 + p];
                Gw[NP3 + p] = X[o3 + size_t(7) * NP3 + p];
            }
        }
        __syncthreads();

        double acc2[SPT][3], acc3[SPT][8];
        for (int n = s_begin; n < s_end; ++n) {
            const bool rec = (n % A.save_interval) == 0;
            const int ri = n / A.save_interval;
            const double tn = A.tgrid[n];
            // RK4 stages at t, t + dt/2, t + dt/2, t + dt
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const MP M = fresh_mix(Mg);
                const MixStageView V{F, TL, Gw + (s & 1) * 2 * NP3, NP2, NP3};  // y_s
                double* Gn = Gw + (1 - (s & 1)) * 2 * NP3;                      // components 6, 7 of y_{s+1}
                const double cdt = (s == 2) ? 1.0 * dt : 0.5 * dt;
                const double bdt = (s == 0 || s == 3) ? b16 : b13;
                const bool obs = (s == 0) && rec;
                // drive factors of this stage time (lanes 0..3: species x pulse)
                double f21 = 0.0, f22 = 0.0, f31 = 0.0, f32 = 0.0;
                if (drive) {
                    const double ts = tn + (s == 0 ? 0.0 : (s == 3 ? 1.0 * dt : 0.5 * dt));
                    double v = 0.0;
                    if (lane < 4) {
                        const int sp = lane >> 1, kk = lane & 1;
                        const double amp = sp ? M->amp3 : M->amp2;
                        if (amp != 0.0) {
                            const double tc = sp ? (kk ? tqb : tqa) : (kk ? tpb : tpa);
                            v = drive_factor(amp, ts, tc, sp ? M->width3 : M->width2, sp ? M->freq3 : M->freq2);
                            v *= kk ? scb : sca;
                        }
                    }
                    f21 = lane_double(v, 0);
                    f22 = lane_double(v, 1);
                    f31 = lane_double(v, 2);
                    f32 = lane_double(v, 3);
                }
                if (obs) observe<SPT, T>(M, V, tid, lane, wave, at2, at3, act2, act3, red);
                double e_acc = 0.0;

                // SU(3) sites: field from y_s, RK4 update; the new stage
                // values go to LDS / G once nothing reads y_s any more
                double y3[TT ? SPT : 1][6];
                auto tm_sites = [&](auto in_place) {
#pragma unroll
                    for (int j = 0; j < SPT; ++j) {
                        if (!((act3 >> j) & 1u)) continue;
                        const int p = fresh(tid + j * T), a = at3[j];
                        const MP M = fresh_mix(Mg);
                        double S[8];
#pragma unroll
                        for (int c = 0; c < 8; ++c) S[c] = V.tm(c, p);
                        double H[8];
                        tm_field<TT>(Mg, V, a, p, S, H);
                        if (obs) e_acc += site_energy_mixed<8>(M->field3, a, S, H);
                        if (f31 != 0.0 || f32 != 0.0) {
                            const cptr<double> d0 = as_const(M->fdrive3) + a * 8;
                            const cptr<double> d1 = as_const(M->fdrive3) + (M->A3 + a) * 8;
#pragma unroll
                            for (int d = 0; d < 8; ++d) H[d] -= d0[d] * f31 + d1[d] * f32;
                        }
                        double k[8];
                        su3_cross(H, S, k);
                        if (M->has_bloch) {  // mixed_lattice_md.cpp:596-605
#pragma unroll
                            for (int d = 0; d < 8; ++d)
                                if (M->gamma[d] != 0.0) k[d] -= M->gamma[d] * (S[d] - M->eq3[size_t(d) * NP3 + p]);
                        }
#pragma unroll
                        for (int d = 0; d < 8; ++d) {
                            const size_t q = o3 + size_t(d) * NP3 + p;
                            double y;
                            if (s == 0) {
                                acc3[j][d] = S[d] + bdt * k[d];
                                y = S[d] + cdt * k[d];
                            } else if (s < 3) {
                                acc3[j][d] = acc3[j][d] + bdt * k[d];
                                y = X[q] + cdt * k[d];
                            } else {
                                y = acc3[j][d] + bdt * k[d];
                                X[q] = y;
                            }
                            if (d >= 6) Gn[(d - 6) * NP3 + p] = y;
                            else if constexpr (decltype(in_place)::value) TL[d * NP3 + p] = y;
                            else y3[TT ? j : 0][d] = y;
                        }
                    }
                };

                // SU(2) sites: field from y_s, RK4 update into y2
                double y2[SPT][3];
#pragma unroll
                for (int j = 0; j < SPT; ++j) {
                    if (!((act2 >> j) & 1u)) continue;
                    const int p = fresh(tid + j * T), a = at2[j];
                    const MP M = fresh_mix(Mg);
                    const double S[3] = {V.fe(0, p), V.fe(1, p), V.fe(2, p)};
                    double H[3];
                    fe_field(Mg, V, a, p, S, H);
                    if (obs) e_acc += site_energy_mixed<3>(M->field2, a, S, H);
                    if (f21 != 0.0 || f22 != 0.0) {
                        const cptr<double> d0 = as_const(M->fdrive2) + a * 3;
                        const cptr<double> d1 = as_const(M->fdrive2) + (M->A2 + a) * 3;
#pragma unroll
                        for (int d = 0; d < 3; ++d) H[d] -= d0[d] * f21 + d1[d] * f22;
                    }
                    double k[3] = {H[1] * S[2] - H[2] * S[1], H[2] * S[0] - H[0] * S[2],
                                   H[0] * S[1] - H[1] * S[0]};
                    if (M->alpha != 0.0) {  // mixed_lattice_md.cpp:570-578
                        const double S2 = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
                        const double SH = S[0] * H[0] + S[1] * H[1] + S[2] * H[2];
                        const double g = (S2 > 0.0) ? M->alpha / sqrt(S2) : 0.0;
#pragma unroll
                        for (int d = 0; d < 3; ++d) k[d] += g * (S[d] * SH - H[d] * S2);
                    }
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const size_t q = size_t(d) * NP2 + p;
                        if (s == 0) {
                            acc2[j][d] = S[d] + bdt * k[d];
                            y2[j][d] = S[d] + cdt * k[d];
                        } else if (s < 3) {
                            acc2[j][d] = acc2[j][d] + bdt * k[d];
                            y2[j][d] = X[q] + cdt * k[d];
                        } else {
                            const double xn = acc2[j][d] + bdt * k[d];
                            X[q] = xn;
                            y2[j][d] = xn;
                        }
                    }
                }
                if constexpr (TT) {
                    // SU(3) sites read each other's stage values: every site
                    // of both species before any new value is stored
                    tm_sites(std::false_type{});
                    if (obs) {
                        const double v = wave_sum(e_acc);
                        if (lane == 0) red[wave * kRec + 33] = v;
                    }
                    __syncthreads();  // every read of y_s is done
#pragma unroll
                    for (int j = 0; j < SPT; ++j) {
                        const int p = fresh(tid + j * T);
                        if ((act3 >> j) & 1u)
#pragma unroll
                            for (int d = 0; d < 6; ++d) TL[d * NP3 + p] = y3[j][d];
                    }
                } else {
                    // no SU(3)-SU(3) bonds: an SU(3) site's own stage value
                    // is read only by its thread, so once the SU(2) sites
                    // are done with y_s the SU(3) sites update in place
                    __syncthreads();  // the SU(2) sites' reads of y_s are done
                    tm_sites(std::true_type{});
                    if (obs) {
                        const double v = wave_sum(e_acc);
                        if (lane == 0) red[wave * kRec + 33] = v;
                    }
                    __syncthreads();  // the SU(3) sites' reads of the SU(2) stage values are done
                }
#pragma unroll
                for (int j = 0; j < SPT; ++j) {
                    const int p = fresh(tid + j * T);
                    if ((act2 >> j) & 1u)
#pragma unroll
                        for (int d = 0; d < 3; ++d) F[d * NP2 + p] = y2[j][d];
                }
                if (obs && wave == 0 && lane < kRec) {
                    double t = 0.0;
                    for (int w = 0; w < nwaves; ++w) t += red[w * kRec + lane];
                    const double sc = lane < 9 ? A.inv2 : (lane < 33 ? A.inv3 : A.invE);
                    A.recs[(size_t(ri) * A.B_rec + b) * kRec + lane] = t * sc;
                }
                __syncthreads();  // y_{s+1} complete (LDS and G)
            }
        }
        if (post >= 0) {
            // hand the state at s_end (in X since the last step's stage 3)
            // to the workgroup that continues the trajectory
            __threadfence();
            __syncthreads();
            if (tid == 0) __hip_atomic_store(A.flags + post, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        // final record: the state after the last step (the stage values of
        // the next step's stage 0: LDS and G[0])
        if (s_end == A.n_steps) {
            const MP M = fresh_mix(Mg);
            const MixStageView V{F, TL, Gw, NP2, NP3};
            observe<SPT, T>(M, V, tid, lane, wave, at2, at3, act2, act3, red);
            double e_acc = 0.0;
#pragma unroll
            for (int j = 0; j < SPT; ++j) {
                if (!((act2 >> j) & 1u)) continue;
                const int p = fresh(tid + j * T), a = at2[j];
                const double S[3] = {V.fe(0, p), V.fe(1, p), V.fe(2, p)};
                double H[3];
                fe_field(Mg, V, a, p, S, H);
                e_acc += site_energy_mixed<3>(M->field2, a, S, H);
            }
#pragma unroll
            for (int j = 0; j < SPT; ++j) {
                if (!((act3 >> j) & 1u)) continue;
                const int p = fresh(tid + j * T), a = at3[j];
                double S[8];
#pragma unroll
                for (int c = 0; c < 8; ++c) S[c] = V.tm(c, p);
                double H[8];
                tm_field<TT>(Mg, V, a, p, S, H);
                e_acc += site_energy_mixed<8>(M->field3, a, S, H);
            }
            const double v = wave_sum(e_acc);
            if (lane == 0) red[wave * kRec + 33] = v;
            __syncthreads();
            if (wave == 0 && lane < kRec) {
                double t = 0.0;
                for (int w = 0; w < nwaves; ++w) t += red[w * kRec + lane];
                const double sc = lane < 9 ? A.inv2 : (lane < 33 ? A.inv3 : A.invE);
                A.recs[(size_t(A.n_rec - 1) * A.B_rec + b) * kRec + lane] = t * sc;
            }
        }
    }
}

namespace {

template <int SPT, int T, bool TT>
hipError_t launch_mp(const MixDev* M, const MixPersistArgs& A, int grid, size_t lds, hipStream_t st, int* occ) {
    auto kern = k_mixed_persist<SPT, T, TT>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    if (occ) return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, kern, T, lds);
    hipLaunchKernelGGL(kern, dim3(unsigned(grid)), dim3(T), lds, st, M, A);
    return hipGetLastError();
}

}  // namespace

// Without SU(3)-SU(3) bonds: 1024 threads (4 waves per SIMD) and up to two
// sites of each species per thread — the SU(3) sites update in place, so a
// thread holds only its accumulators and its SU(2) sites' next values.  With
// them: 512 threads and up to four sites (the SU(3) next values are held too).
MixPersistShape mixed_persist_shape(int NP2, int NP3, bool tt) {
    const int np = std::max(NP2, NP3);
    if (!tt) {
        if (np <= 1024) return {1024, 1, false};
        if (np <= 2048) return {1024, 2, false};
        return {0, 0, false};
    }
    const int spt = (np + 511) / 512;
    if (spt <= 4) return {512, spt, true};
    return {0, 0, true};
}

size_t mixed_persist_lds_bytes(int NP2, int NP3, int block) {
    return sizeof(double) * (3 * size_t(NP2) + 6 * size_t(NP3) + size_t(block / kWave) * kRec);
}

hipError_t launch_mixed_persist(const MixDev& Mh, const MixDev* M, const MixPersistArgs& A, const MixPersistShape& sh,
                                int grid, hipStream_t st, int* occ) {
    const size_t lds = mixed_persist_lds_bytes(Mh.NP2, Mh.NP3, sh.block);
    if (!sh.tt && sh.block == 1024) {
        if (sh.spt == 1) return launch_mp<1, 1024, false>(M, A, grid, lds, st, occ);
        if (sh.spt == 2) return launch_mp<2, 1024, false>(M, A, grid, lds, st, occ);
    }
    if (sh.tt && sh.block == 512) {
        switch (sh.spt) {
            case 1: return launch_mp<1, 512, true>(M, A, grid, lds, st, occ);
            case 2: return launch_mp<2, 512, true>(M, A, grid, lds, st, occ);
            case 3: return launch_mp<3, 512, true>(M, A, grid, lds, st, occ);
            case 4: return launch_mp<4, 512, true>(M, A, grid, lds, st, occ);
            default: break;
        }
    }
    return hipErrorInvalidValue;
}

}  // namespace csl
