// csl_kernels.hip — gfx950 kernels of the Landau–Lifshitz RK4 path.
//
//   k_persistent   one workgroup per trajectory; the whole lattice state of
//                  that trajectory lives in LDS (SoA), each thread owns SPT
//                  sites and keeps their RK4 state x and accumulator in
//                  VGPRs; all steps run inside one launch; M(t) and E(t)
//                  are reduced in-kernel (wave64 shuffles + LDS) every
//                  save_interval steps.  Replaces the per-step launch
//                  sequence of integrate_gpu_batched_internal
//                  (lattice_gpu.cu:455-577).
//   k_stage        grid-wide fused RK4 stage: local field + H x S + stage
//                  update + accumulator update for one stage of every site,
//                  one launch per stage (the "plain" path and large
//                  lattices).  Replaces LLG_flat_kernel + update_arrays_kernel
//                  + rk4_final_update_kernel (lattice_gpu.cu:73-117,
//                  gpu_common_helpers.cu:9-52).
//   k_eval         local field / dS/dt / energy terms of one state.
//   k_reduce_obs   block partials -> one record (M, E).
//   k_perm_*       reference AoS order <-> device SoA order.
#include "csl_kernels.hpp"

namespace csl {

// ---------------------------------------------------------------------------
// helpers

// A wave-uniform zero the compiler cannot see through.  Table pointers are
// offset by it once per RK stage so that the wave-uniform table reads (bond
// matrices, field, onsite matrix, frames, drive directions) are re-issued as
// scalar loads inside the time loop instead of being hoisted out of it and
// pinned in VGPRs for the whole kernel.
__device__ __forceinline__ int opaque_zero() {
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    return z;
}

__device__ __forceinline__ LatDev refresh(const LatDev& L0) {
    LatDev L = L0;
    const int z = opaque_zero();
    L.mats += z;
    L.field += z;
    L.onsite += z;
    L.frames += z;
    L.afm += z;
    L.fdrive += z;
    return L;
}

__device__ __forceinline__ int atom_of(const LatDev& L, int p) { return p / L.npad; }
__device__ __forceinline__ bool active_site(const LatDev& L, int p) {
    return p < L.NP && (p - (p / L.npad) * L.npad) < L.n_cells;
}

// Local field of site p (no drive) from an SoA state plane set.
template <int ZMAX, bool UNI>
__device__ __forceinline__ void local_field_soa(const LatDev& L, const double* __restrict__ y,
                                                int NP, int p, int atom, const int* nb, int cnt,
                                                const double S[3], double H[3]) {
    field_onsite(L, atom, p, S, H);
#pragma unroll
    for (int k = 0; k < ZMAX; ++k) {
        if (k < cnt) {
            const int q = nb[k];
            const double Sp[3] = {y[q], y[NP + q], y[2 * NP + q]};
            if (UNI)
                add_bond(as_const(L.mats) + (atom * L.zmax + k) * 9, Sp, H);
            else
                add_bond(L.mats + (size_t(k) * NP + p) * 9, Sp, H);
        }
    }
}

// Partial observables of one site: [M_anti(3), M_local(3), M_global(3), e_i]
// e_i = 1/2 S.(H - B) with H the undriven local field, which sums to the
// reference total energy  sum(-B.S + S.A.S) + sum_pairs S_i J S_j
// (lattice.h:1485-1529) because every bond appears once with J and once with
// J^T (lattice.h:790-803).
__device__ __forceinline__ void site_obs(const LatDev& L, int atom, int p, const double S[3],
                                         const double H[3], double part[10]) {
    cptr<double> R = as_const(L.frames) + atom * 9;
    const double sg = as_const(L.afm)[atom];
    const double* B = L.site_uniform ? (const double*)(L.field + atom * 3) : L.field + size_t(p) * 3;
    double Bv[3];
    if (L.site_uniform) {
        cptr<double> Bu = as_const(L.field) + atom * 3;
        Bv[0] = Bu[0]; Bv[1] = Bu[1]; Bv[2] = Bu[2];
    } else {
        Bv[0] = B[0]; Bv[1] = B[1]; Bv[2] = B[2];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double g = R[d * 3 + 0] * S[0] + R[d * 3 + 1] * S[1] + R[d * 3 + 2] * S[2];
        part[d] += sg * g;
        part[3 + d] += S[d];
        part[6 + d] += g;
    }
    part[9] += 0.5 * (S[0] * (H[0] - Bv[0]) + S[1] * (H[1] - Bv[1]) + S[2] * (H[2] - Bv[2]));
}

// e_i = 1/2 S.(H - B) of one site (see site_obs)
__device__ __forceinline__ double site_energy(const LatDev& L, int atom, int p, const double S[3],
                                              const double H[3]) {
    double B0, B1, B2;
    if (L.site_uniform) {
        cptr<double> Bu = as_const(L.field) + atom * 3;
        B0 = Bu[0]; B1 = Bu[1]; B2 = Bu[2];
    } else {
        const double* Bp = L.field + size_t(p) * 3;
        B0 = Bp[0]; B1 = Bp[1]; B2 = Bp[2];
    }
    return 0.5 * (S[0] * (H[0] - B0) + S[1] * (H[1] - B1) + S[2] * (H[2] - B2));
}

__device__ __forceinline__ void sub_drive(const LatDev& L, int atom, double f1, double f2,
                                          double H[3]) {
    cptr<double> d0 = as_const(L.fdrive) + atom * 3;
    cptr<double> d1 = as_const(L.fdrive) + (L.n_atoms + atom) * 3;
#pragma unroll
    for (int d = 0; d < 3; ++d) H[d] -= d0[d] * f1 + d1[d] * f2;
}

// ---------------------------------------------------------------------------
// persistent batched kernel

// Largest workgroup a persistent instantiation accepts.  Smaller bounds give
// each wave a larger VGPR budget (512 / waves-per-SIMD): FP64 state and
// accumulator cost 12 VGPRs per owned site, so threads that own more sites
// run in smaller workgroups.
template <int SPT>
constexpr int persistent_max_block() {
    return SPT <= 2 ? 1024 : (SPT <= 4 ? 768 : 512);
}

template <int SPT, int ZMAX, bool UNI, bool DBUF>
__global__ void __launch_bounds__(persistent_max_block<SPT>()) k_persistent(LatDev L0, PersistArgs A) {
    const LatDev& L = L0;
    extern __shared__ double smem[];
    const int NP = L.NP;
    double* bufA = smem;
    double* bufB = DBUF ? smem + 3 * NP : smem;
    double* drv = smem + (DBUF ? 6 : 3) * NP;  // [2 parity][3 stage times][2 pulses]
    double* red = drv + 12;                    // [nwaves][10]

    const int b = blockIdx.x;
    const int T = blockDim.x;
    const int tid = threadIdx.x;
    const int wave = uniform(tid >> 6);
    const int lane = tid & (kWave - 1);
    const int nwaves = T >> 6;

    const double tp1 = A.t_pulse[2 * b], tp2 = A.t_pulse[2 * b + 1];
    const double am1 = A.amp[2 * b], am2 = A.amp[2 * b + 1];
    const bool drive_on = (am1 != 0.0) || (am2 != 0.0);
    const double dt = A.dt;
    const double alpha = L.alpha;

    // initial state -> LDS (device order)
    const double* init = A.init + size_t(b) * A.init_stride;
    for (int p = tid; p < NP; p += T) {
        const int atom = p / L.npad, cell = p - atom * L.npad;
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
        if (cell < L.n_cells) {
            const size_t i = size_t(cell) * L.n_atoms + atom;
            v0 = init[3 * i];
            v1 = init[3 * i + 1];
            v2 = init[3 * i + 2];
        }
        bufA[p] = v0;
        bufA[NP + p] = v1;
        bufA[2 * NP + p] = v2;
    }

    // per-thread ownership: sites p = tid + m*T
    double x[SPT][3], acc[SPT][3];
    int nb[SPT][ZMAX];
    int cnt[SPT], atm[SPT];
    bool act[SPT];
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        const int p = tid + m * T;
        const int pw = wave * kWave + m * T;  // wave-uniform
        atm[m] = uniform(pw < NP ? pw / L.npad : 0);
        act[m] = active_site(L, p);
        const int c = UNI ? uniform(pw < NP ? as_const(L.cnt_atom)[atm[m]] : 0) : (p < NP ? L.cnt[p] : 0);
        cnt[m] = c;
#pragma unroll
        for (int k = 0; k < ZMAX; ++k) nb[m][k] = (act[m] && k < c) ? L.col[size_t(k) * NP + p] : 0;
        acc[m][0] = acc[m][1] = acc[m][2] = 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        const int p = tid + m * T;
        if (act[m]) {
            x[m][0] = bufA[p];
            x[m][1] = bufA[NP + p];
            x[m][2] = bufA[2 * NP + p];
        } else {
            x[m][0] = x[m][1] = x[m][2] = 0.0;
        }
    }

    const double b16 = (1.0 / 6.0) * dt, b13 = (1.0 / 3.0) * dt;
    const double inv_n = 1.0 / double(L.N);

    // Drive envelopes of the three stage times of step n (t, t+dt/2, t+dt)
    // for both pulses, written by 6 lanes of wave 0 into parity slot n&1.
    // Slot n&1 is written at the top of step n-1 (step 0: before the loop)
    // and read during step n; the barriers of step n-1 separate the two.
    auto write_drive = [&](int n) {
        if (drive_on && wave == 0 && lane < 6) {
            const int st = lane >> 1, pl = lane & 1;
            const double t = A.tgrid[n];
            const double ts = st == 0 ? t : (st == 1 ? t + 0.5 * dt : t + 1.0 * dt);
            drv[(n & 1) * 6 + lane] = drive_factor(pl ? am2 : am1, ts, pl ? tp2 : tp1, L.width, L.freq);
        }
    };
    // wave-reduce the observables of the owned sites and publish them in red[]
    auto reduce_obs = [&](const LatDev& Lr, const double* e) {
        double part[10];
#pragma unroll
        for (int j = 0; j < 10; ++j) part[j] = 0.0;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            if (act[m]) {
                cptr<double> R = as_const(Lr.frames) + atm[m] * 9;
                const double sg = as_const(Lr.afm)[atm[m]];
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
            const double v = wave_sum(part[j]);
            if (lane == 0) red[wave * 10 + j] = v;
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

    write_drive(0);
    __syncthreads();

    for (int n = 0; n < A.n_steps; ++n) {
        const int par = n & 1;
        const bool rec = (n % A.save_interval) == 0;
        const int ri = n / A.save_interval;
        if (n + 1 < A.n_steps) write_drive(n + 1);
#pragma unroll 1
        for (int s = 0; s < 4; ++s) {
            const LatDev L = refresh(L0);
            const double* cur = (DBUF && (s & 1)) ? bufB : bufA;
            double* nxt = DBUF ? ((s & 1) ? bufA : bufB) : bufA;
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
                const int p = tid + m * T;
                e[m] = 0.0;
                if (act[m]) {
                    const int atom = atm[m];
                    const double S[3] = {cur[p], cur[NP + p], cur[2 * NP + p]};
                    double H[3];
                    local_field_soa<ZMAX, UNI>(L, cur, NP, p, atom, nb[m], cnt[m], S, H);
                    if (obs) e[m] = site_energy(L, atom, p, S, H);
                    if (drive_on) sub_drive(L, atom, f1, f2, H);
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
                        nxt[p] = y[0];
                        nxt[NP + p] = y[1];
                        nxt[2 * NP + p] = y[2];
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
                    const int p = tid + m * T;
                    if (act[m]) {
                        nxt[p] = yn[DBUF ? 0 : m][0];
                        nxt[NP + p] = yn[DBUF ? 0 : m][1];
                        nxt[2 * NP + p] = yn[DBUF ? 0 : m][2];
                    }
                }
            }
            __syncthreads();
            if (obs) publish(ri);
        }
    }

    // final record: state after the last step (lattice_gpu.cu:567-568)
    {
        const LatDev L = refresh(L0);
        double e[SPT];
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = tid + m * T;
            e[m] = 0.0;
            if (act[m]) {
                double H[3];
                local_field_soa<ZMAX, UNI>(L, bufA, NP, p, atm[m], nb[m], cnt[m], x[m], H);
                e[m] = site_energy(L, atm[m], p, x[m], H);
            }
        }
        reduce_obs(L, e);
        __syncthreads();
        publish(A.n_rec - 1);
        if (A.final_states) {
            double* out = A.final_states + size_t(b) * L.N * 3;
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int p = tid + m * T;
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
}

// ---------------------------------------------------------------------------
// grid-wide fused stage kernel (plain path, any lattice size, B in grid.y)

template <int ZMAX, bool UNI>
__global__ void __launch_bounds__(256) k_stage(LatDev L, StageArgs A) {
    __shared__ double red[4 * 10];
    const int NP = L.NP;
    const int b = blockIdx.y;
    const int tid = threadIdx.x;
    const int p = blockIdx.x * blockDim.x + tid;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & (kWave - 1);
    const int pw = blockIdx.x * blockDim.x + wave * kWave;
    const int atom = uniform(pw < NP ? pw / L.npad : 0);
    const size_t off = size_t(b) * 3 * NP;
    const double* cur = A.ycur + off;
    double* nxt = A.ynext + off;
    double* x = A.x + off;
    double* acc = A.acc + off;
    const int s = A.stage;
    const double f1 = A.fac ? A.fac[(size_t(b) * 2) + 0] : 0.0;
    const double f2 = A.fac ? A.fac[(size_t(b) * 2) + 1] : 0.0;
    const bool drive_on = A.fac != nullptr;
    const double dt = A.dt;
    const double cdt = (s == 2) ? 1.0 * dt : 0.5 * dt;
    const double bdt = (s == 0 || s == 3) ? (1.0 / 6.0) * dt : (1.0 / 3.0) * dt;
    double part[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) part[j] = 0.0;
    if (active_site(L, p)) {
        const int c = UNI ? L.cnt_atom[atom] : L.cnt[p];
        int nb[ZMAX];
#pragma unroll
        for (int k = 0; k < ZMAX; ++k) nb[k] = k < c ? L.col[size_t(k) * NP + p] : 0;
        const double S[3] = {cur[p], cur[NP + p], cur[2 * NP + p]};
        double H[3];
        local_field_soa<ZMAX, UNI>(L, cur, NP, p, atom, nb, c, S, H);
        if (A.rec) site_obs(L, atom, p, S, H, part);
        if (drive_on) sub_drive(L, atom, f1, f2, H);
        double k[3];
        ll_su2(H, S, L.alpha, k);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const size_t q = size_t(d) * NP + p;
            const double xd = x[q];
            double xn = xd;
            if (s == 0) acc[q] = xd + bdt * k[d];
            else if (s < 3) acc[q] = acc[q] + bdt * k[d];
            else {
                xn = acc[q] + bdt * k[d];
                x[q] = xn;
            }
            if (s < 3) nxt[q] = xd + cdt * k[d];
        }
    }
    if (A.rec) {
#pragma unroll
        for (int j = 0; j < 10; ++j) {
            const double v = wave_sum(part[j]);
            if (lane == 0) red[wave * 10 + j] = v;
        }
        __syncthreads();
        if (tid < 10) {
            double v = 0.0;
            for (int w = 0; w < int(blockDim.x >> 6); ++w) v += red[w * 10 + tid];
            A.partials[(size_t(b) * gridDim.x + blockIdx.x) * 10 + tid] = v;
        }
    }
}

// observables of a state without stepping (final record of the plain path)
template <int ZMAX, bool UNI>
__global__ void __launch_bounds__(256) k_observe(LatDev L, const double* __restrict__ y,
                                                 double* __restrict__ partials) {
    __shared__ double red[4 * 10];
    const int NP = L.NP;
    const int b = blockIdx.y;
    const int tid = threadIdx.x;
    const int p = blockIdx.x * blockDim.x + tid;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & (kWave - 1);
    const int pw = blockIdx.x * blockDim.x + wave * kWave;
    const int atom = uniform(pw < NP ? pw / L.npad : 0);
    const double* cur = y + size_t(b) * 3 * NP;
    double part[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) part[j] = 0.0;
    if (active_site(L, p)) {
        const int c = UNI ? L.cnt_atom[atom] : L.cnt[p];
        int nb[ZMAX];
#pragma unroll
        for (int k = 0; k < ZMAX; ++k) nb[k] = k < c ? L.col[size_t(k) * NP + p] : 0;
        const double S[3] = {cur[p], cur[NP + p], cur[2 * NP + p]};
        double H[3];
        local_field_soa<ZMAX, UNI>(L, cur, NP, p, atom, nb, c, S, H);
        site_obs(L, atom, p, S, H, part);
    }
#pragma unroll
    for (int j = 0; j < 10; ++j) {
        const double v = wave_sum(part[j]);
        if (lane == 0) red[wave * 10 + j] = v;
    }
    __syncthreads();
    if (tid < 10) {
        double v = 0.0;
        for (int w = 0; w < int(blockDim.x >> 6); ++w) v += red[w * 10 + tid];
        partials[(size_t(b) * gridDim.x + blockIdx.x) * 10 + tid] = v;
    }
}

// one block per trajectory: sum block partials into record ri
__global__ void __launch_bounds__(64) k_reduce_obs(const double* __restrict__ partials, int nblk,
                                                   int B, int ri, double inv_n,
                                                   double* __restrict__ mags,
                                                   double* __restrict__ energy) {
    const int b = blockIdx.x;
    const int j = threadIdx.x;
    if (j >= 10) return;
    double v = 0.0;
    for (int k = 0; k < nblk; ++k) v += partials[(size_t(b) * nblk + k) * 10 + j];
    if (j < 9) {
        if (mags) mags[(size_t(ri) * B + b) * 9 + j] = v * inv_n;
    } else if (energy) {
        energy[size_t(ri) * B + b] = v * inv_n;
    }
}

// local field (no drive) and dS/dt at time t for one state (SoA in, SoA out)
template <int ZMAX, bool UNI>
__global__ void __launch_bounds__(256) k_eval(LatDev L, const double* __restrict__ y,
                                              double* __restrict__ Hout,
                                              double* __restrict__ dsdt, double f1, double f2,
                                              int drive_on) {
    const int NP = L.NP;
    const int tid = threadIdx.x;
    const int p = blockIdx.x * blockDim.x + tid;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int pw = blockIdx.x * blockDim.x + wave * kWave;
    const int atom = uniform(pw < NP ? pw / L.npad : 0);
    if (!active_site(L, p)) return;
    const int c = UNI ? L.cnt_atom[atom] : L.cnt[p];
    int nb[ZMAX];
#pragma unroll
    for (int k = 0; k < ZMAX; ++k) nb[k] = k < c ? L.col[size_t(k) * NP + p] : 0;
    const double S[3] = {y[p], y[NP + p], y[2 * NP + p]};
    double H[3];
    local_field_soa<ZMAX, UNI>(L, y, NP, p, atom, nb, c, S, H);
    if (Hout) {
        Hout[p] = H[0];
        Hout[NP + p] = H[1];
        Hout[2 * NP + p] = H[2];
    }
    if (dsdt) {
        if (drive_on) sub_drive(L, atom, f1, f2, H);
        double k[3];
        ll_su2(H, S, L.alpha, k);
        dsdt[p] = k[0];
        dsdt[NP + p] = k[1];
        dsdt[2 * NP + p] = k[2];
    }
}

// reference AoS [B][N][3]  <->  device SoA [B][3][NP]
__global__ void k_perm_to_soa(LatDev L, const double* __restrict__ aos, size_t aos_stride,
                              double* __restrict__ soa) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (p >= L.NP) return;
    const int atom = p / L.npad, cell = p - atom * L.npad;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (cell < L.n_cells) {
        const double* a = aos + size_t(b) * aos_stride + (size_t(cell) * L.n_atoms + atom) * 3;
        v0 = a[0];
        v1 = a[1];
        v2 = a[2];
    }
    double* o = soa + size_t(b) * 3 * L.NP;
    o[p] = v0;
    o[L.NP + p] = v1;
    o[2 * L.NP + p] = v2;
}

__global__ void k_perm_to_aos(LatDev L, const double* __restrict__ soa, double* __restrict__ aos) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (!active_site(L, p)) return;
    const int atom = p / L.npad, cell = p - atom * L.npad;
    const double* s = soa + size_t(b) * 3 * L.NP;
    double* a = aos + (size_t(b) * L.N + size_t(cell) * L.n_atoms + atom) * 3;
    a[0] = s[p];
    a[1] = s[L.NP + p];
    a[2] = s[2 * L.NP + p];
}

// S <- S * l / |S| when |S| > 1e-10 (normalize_spins_kernel,
// gpu_common_helpers.cu:138-160), on the reference AoS state
__global__ void k_normalize(double* __restrict__ s, int N, double len) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double* v = s + size_t(i) * 3;
    const double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (n > 1e-10) {
        const double f = len / n;
        v[0] *= f;
        v[1] *= f;
        v[2] *= f;
    }
}

// drive factors of every stage time of every step for B trajectories:
// fac[((n*3 + st)*B + b)*2 + pulse]
__global__ void k_drive_table(const double* __restrict__ tgrid, int n_steps, double dt, int B,
                              const double* __restrict__ t_pulse, const double* __restrict__ amp,
                              double width, double freq, double* __restrict__ fac) {
    const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t total = size_t(n_steps) * 3 * B * 2;
    if (id >= total) return;
    const int pl = int(id & 1);
    const size_t r = id >> 1;
    const int b = int(r % B);
    const size_t ns = r / B;
    const int st = int(ns % 3);
    const size_t n = ns / 3;
    const double t = tgrid[n];
    const double ts = st == 0 ? t : (st == 1 ? t + 0.5 * dt : t + dt);
    fac[id] = drive_factor(amp[2 * b + pl], ts, t_pulse[2 * b + pl], width, freq);
}

// ---------------------------------------------------------------------------
// launchers

namespace {

template <int SPT, int ZMAX, bool UNI, bool DBUF>
hipError_t launch_p(const LatDev& L, const PersistArgs& A, int block, size_t lds,
                    hipStream_t st) {
    auto kern = k_persistent<SPT, ZMAX, UNI, DBUF>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(A.B), dim3(block), lds, st, L, A);
    return hipGetLastError();
}

template <int SPT, int ZMAX, bool UNI>
hipError_t launch_p_db(const LatDev& L, const PersistArgs& A, int block, bool dbuf, size_t lds,
                       hipStream_t st) {
    return dbuf ? launch_p<SPT, ZMAX, UNI, true>(L, A, block, lds, st)
                : launch_p<SPT, ZMAX, UNI, false>(L, A, block, lds, st);
}

template <int SPT, int ZMAX>
hipError_t launch_p_uni(const LatDev& L, const PersistArgs& A, int block, bool dbuf, size_t lds,
                        hipStream_t st) {
    return L.uniform ? launch_p_db<SPT, ZMAX, true>(L, A, block, dbuf, lds, st)
                     : launch_p_db<SPT, ZMAX, false>(L, A, block, dbuf, lds, st);
}

template <int SPT>
hipError_t launch_p_z(const LatDev& L, const PersistArgs& A, int block, bool dbuf, size_t lds,
                      hipStream_t st) {
    if (L.zmax <= 3) return launch_p_uni<SPT, 3>(L, A, block, dbuf, lds, st);
    if (L.zmax <= 6) return launch_p_uni<SPT, 6>(L, A, block, dbuf, lds, st);
    if (L.zmax <= 12) return launch_p_uni<SPT, 12>(L, A, block, dbuf, lds, st);
    return hipErrorInvalidValue;
}

template <int ZMAX>
hipError_t launch_s_uni(const LatDev& L, const StageArgs& A, int nblk, hipStream_t st) {
    if (L.uniform)
        hipLaunchKernelGGL((k_stage<ZMAX, true>), dim3(nblk, A.B), dim3(256), 0, st, L, A);
    else
        hipLaunchKernelGGL((k_stage<ZMAX, false>), dim3(nblk, A.B), dim3(256), 0, st, L, A);
    return hipGetLastError();
}

template <int ZMAX>
hipError_t launch_o_uni(const LatDev& L, const double* y, double* partials, int nblk, int B,
                        hipStream_t st) {
    if (L.uniform)
        hipLaunchKernelGGL((k_observe<ZMAX, true>), dim3(nblk, B), dim3(256), 0, st, L, y, partials);
    else
        hipLaunchKernelGGL((k_observe<ZMAX, false>), dim3(nblk, B), dim3(256), 0, st, L, y,
                           partials);
    return hipGetLastError();
}

template <int ZMAX>
hipError_t launch_e_uni(const LatDev& L, const double* y, double* H, double* dsdt, double f1,
                        double f2, int drive_on, hipStream_t st) {
    const int nblk = (L.NP + 255) / 256;
    if (L.uniform)
        hipLaunchKernelGGL((k_eval<ZMAX, true>), dim3(nblk), dim3(256), 0, st, L, y, H, dsdt, f1,
                           f2, drive_on);
    else
        hipLaunchKernelGGL((k_eval<ZMAX, false>), dim3(nblk), dim3(256), 0, st, L, y, H, dsdt, f1,
                           f2, drive_on);
    return hipGetLastError();
}

}  // namespace

size_t persistent_lds_bytes(const LatDev& L, int block, bool dbuf) {
    return sizeof(double) * (size_t(dbuf ? 6 : 3) * L.NP + 12 + size_t(block / kWave) * 10);
}

hipError_t launch_persistent(const LatDev& L, const PersistArgs& A, int block, int spt, bool dbuf,
                             hipStream_t st) {
    const size_t lds = persistent_lds_bytes(L, block, dbuf);
    switch (spt) {
        case 1: return launch_p_z<1>(L, A, block, dbuf, lds, st);
        case 2: return launch_p_z<2>(L, A, block, dbuf, lds, st);
        case 3: return launch_p_z<3>(L, A, block, dbuf, lds, st);
        case 4: return launch_p_z<4>(L, A, block, dbuf, lds, st);
        case 6: return launch_p_z<6>(L, A, block, dbuf, lds, st);
        default: return hipErrorInvalidValue;
    }
}

bool persistent_spt_supported(int spt) {
    return spt == 1 || spt == 2 || spt == 3 || spt == 4 || spt == 6;
}

int persistent_block_limit(int spt) {
    switch (spt) {
        case 1: return persistent_max_block<1>();
        case 2: return persistent_max_block<2>();
        case 3: return persistent_max_block<3>();
        case 4: return persistent_max_block<4>();
        case 6: return persistent_max_block<6>();
        default: return 0;
    }
}

hipError_t launch_stage(const LatDev& L, const StageArgs& A, hipStream_t st) {
    const int nblk = (L.NP + 255) / 256;
    if (L.zmax <= 3) return launch_s_uni<3>(L, A, nblk, st);
    if (L.zmax <= 6) return launch_s_uni<6>(L, A, nblk, st);
    if (L.zmax <= 12) return launch_s_uni<12>(L, A, nblk, st);
    return hipErrorInvalidValue;
}

int stage_blocks(const LatDev& L) { return (L.NP + 255) / 256; }

hipError_t launch_observe(const LatDev& L, const double* y, double* partials, int B,
                          hipStream_t st) {
    const int nblk = (L.NP + 255) / 256;
    if (L.zmax <= 3) return launch_o_uni<3>(L, y, partials, nblk, B, st);
    if (L.zmax <= 6) return launch_o_uni<6>(L, y, partials, nblk, B, st);
    if (L.zmax <= 12) return launch_o_uni<12>(L, y, partials, nblk, B, st);
    return hipErrorInvalidValue;
}

hipError_t launch_reduce_obs(const double* partials, int nblk, int B, int ri, double inv_n,
                             double* mags, double* energy, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_obs, dim3(B), dim3(64), 0, st, partials, nblk, B, ri, inv_n, mags,
                       energy);
    return hipGetLastError();
}

hipError_t launch_eval(const LatDev& L, const double* y, double* H, double* dsdt, double f1,
                       double f2, int drive_on, hipStream_t st) {
    if (L.zmax <= 3) return launch_e_uni<3>(L, y, H, dsdt, f1, f2, drive_on, st);
    if (L.zmax <= 6) return launch_e_uni<6>(L, y, H, dsdt, f1, f2, drive_on, st);
    if (L.zmax <= 12) return launch_e_uni<12>(L, y, H, dsdt, f1, f2, drive_on, st);
    return hipErrorInvalidValue;
}

hipError_t launch_perm_to_soa(const LatDev& L, const double* aos, size_t aos_stride, double* soa,
                              int B, hipStream_t st) {
    hipLaunchKernelGGL(k_perm_to_soa, dim3((L.NP + 255) / 256, B), dim3(256), 0, st, L, aos,
                       aos_stride, soa);
    return hipGetLastError();
}

hipError_t launch_perm_to_aos(const LatDev& L, const double* soa, double* aos, int B,
                              hipStream_t st) {
    hipLaunchKernelGGL(k_perm_to_aos, dim3((L.NP + 255) / 256, B), dim3(256), 0, st, L, soa, aos);
    return hipGetLastError();
}

hipError_t launch_normalize(double* s, int N, double len, hipStream_t st) {
    hipLaunchKernelGGL(k_normalize, dim3((N + 255) / 256), dim3(256), 0, st, s, N, len);
    return hipGetLastError();
}

hipError_t launch_drive_table(const double* tgrid, int n_steps, double dt, int B,
                              const double* t_pulse, const double* amp, double width, double freq,
                              double* fac, hipStream_t st) {
    const size_t total = size_t(n_steps) * 3 * B * 2;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_drive_table, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, tgrid,
                       n_steps, dt, B, t_pulse, amp, width, freq, fac);
    return hipGetLastError();
}

}  // namespace csl
