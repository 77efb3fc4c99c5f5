// csl_kernels.hip — gfx950 kernels of the Landau–Lifshitz RK4 path.
//
//   (the persistent batched kernel lives in csl_persistent.hip)
//   k_stage        grid-wide fused RK4 stage: local field + H x S + stage
//                  update + accumulator update for one stage of every site,
//                  one launch per stage (the "plain" path and large
//                  lattices).  Replaces LLG_flat_kernel + update_arrays_kernel
//                  + rk4_final_update_kernel (lattice_gpu.cu:73-117,
//                  gpu_common_helpers.cu:9-52).
//   k_eval         local field / dS/dt / energy terms of one state.
//   k_reduce_obs   block partials -> one record (M, E).
//   k_perm_*       reference AoS order <-> device SoA order.
#include "csl_device.hpp"

namespace csl {

// grid-wide fused stage kernel (plain path, any lattice size, B in grid.y)

// Workgroups are handed to the 8 XCDs round-robin in launch order; when the
// grid size is a multiple of 8, XCD x is given the contiguous logical range
// [x T/8, (x+1) T/8) instead, so the workgroups of one trajectory share one
// L2 and its neighbour gathers hit there.
__device__ __forceinline__ void xcd_block(int& bx, int& by) {
    const int nx = gridDim.x;
    const int T = nx * gridDim.y;
    bx = blockIdx.x;
    by = blockIdx.y;
    if ((T & 7) == 0) {
        const int h = bx + by * nx;
        const int logical = (h & 7) * (T >> 3) + (h >> 3);
        bx = logical % nx;
        by = logical / nx;
    }
}

template <int ZMAX, bool UNI>
__global__ void __launch_bounds__(256) k_stage(LatDev L, StageArgs A) {
    __shared__ double red[4 * 10];
    const int NP = L.NP;
    int bx, b;
    xcd_block(bx, b);
    const int tid = threadIdx.x;
    const int p = bx * blockDim.x + tid;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & (kWave - 1);
    const int pw = bx * blockDim.x + wave * kWave;
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
        if (drive_on && (f1 != 0.0 || f2 != 0.0)) sub_drive(L, atom, f1, f2, H);
        double k[3];
        ll_su2(H, S, L.alpha, k);
        // stage 0 reads its input from x itself, so S is x there; stage 3
        // needs only the accumulator
        if (s == 0) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const size_t q = size_t(d) * NP + p;
                acc[q] = S[d] + bdt * k[d];
                nxt[q] = S[d] + cdt * k[d];
            }
        } else if (s < 3) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const size_t q = size_t(d) * NP + p;
                const double xd = x[q];
                acc[q] = acc[q] + bdt * k[d];
                nxt[q] = xd + cdt * k[d];
            }
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const size_t q = size_t(d) * NP + p;
                x[q] = acc[q] + bdt * k[d];
            }
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
            A.partials[(size_t(b) * gridDim.x + bx) * 10 + tid] = v;
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
__global__ void __launch_bounds__(256) k_reduce_obs(const double* __restrict__ partials, int nblk,
                                                    int B, int ri, double inv_n,
                                                    double* __restrict__ mags,
                                                    double* __restrict__ energy) {
    // one workgroup per trajectory: strided partial sums per thread, wave64
    // butterfly sums, then the 4 wave totals in a fixed order
    __shared__ double red[10][4];
    const int b = blockIdx.x;
    const int t = threadIdx.x;
    double v[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) v[j] = 0.0;
    for (int k = t; k < nblk; k += 256) {
        const double* q = partials + (size_t(b) * nblk + k) * 10;
#pragma unroll
        for (int j = 0; j < 10; ++j) v[j] += q[j];
    }
#pragma unroll
    for (int j = 0; j < 10; ++j) v[j] = wave_sum(v[j]);
    if ((t & (kWave - 1)) == 0)
#pragma unroll
        for (int j = 0; j < 10; ++j) red[j][t >> 6] = v[j];
    __syncthreads();
    if (t < 10) {
        const double tot = (red[t][0] + red[t][1]) + (red[t][2] + red[t][3]);
        if (t < 9) {
            if (mags) mags[(size_t(ri) * B + b) * 9 + t] = tot * inv_n;
        } else if (energy) {
            energy[size_t(ri) * B + b] = tot * inv_n;
        }
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

// stage input / result of the generic fixed-step engine, element-wise (the
// reference's update_arrays / rk_stage_update / final-update kernels of one
// stage folded into one pass, gpu_common_helpers.cu:9-120).  out may alias x:
// each element is read before it is written by the same thread.
__global__ void __launch_bounds__(256) k_rk_combine(double* out, const double* x, const double* __restrict__ K,
                                                    size_t n, RkCombine c) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = c.g * x[i];
    for (int j = 0; j < c.m; ++j) v = fma(c.w[j], K[size_t(c.idx[j]) * n + i], v);
    out[i] = v;
}

__global__ void __launch_bounds__(256) k_fill_u32(unsigned* __restrict__ p, unsigned v, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// launchers

namespace {

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

namespace {
hipError_t persistent_dispatch(const LatDev& L, const PersistArgs& A, int block, int spt, int flags,
                               hipStream_t st, int* occ) {
    const bool dbuf = flags & kPersistDoubleBuffer;
    const size_t lds = persistent_lds_bytes(L, block, dbuf);
    switch (spt) {
        case 1: return launch_persistent_spt<1>(L, A, block, flags, lds, st, occ);
        case 2: return launch_persistent_spt<2>(L, A, block, flags, lds, st, occ);
        case 3: return launch_persistent_spt<3>(L, A, block, flags, lds, st, occ);
        case 4: return launch_persistent_spt<4>(L, A, block, flags, lds, st, occ);
        case 6: return launch_persistent_spt<6>(L, A, block, flags, lds, st, occ);
        default: return hipErrorInvalidValue;
    }
}
}  // namespace

hipError_t launch_persistent(const LatDev& L, const PersistArgs& A, int block, int spt, int flags,
                             hipStream_t st) {
    return persistent_dispatch(L, A, block, spt, flags, st, nullptr);
}

int persistent_occupancy(const LatDev& L, int block, int spt, int flags) {
    int occ = 0;
    PersistArgs A{};
    if (persistent_dispatch(L, A, block, spt, flags, nullptr, &occ) != hipSuccess) return 0;
    return occ;
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
    hipLaunchKernelGGL(k_reduce_obs, dim3(B), dim3(256), 0, st, partials, nblk, B, ri, inv_n, mags,
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

hipError_t launch_fill_u32(unsigned* p, unsigned v, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_rk_combine(double* out, const double* x, const double* K, size_t n, const RkCombine& c,
                             hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (c.m < 0 || c.m > 13) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_rk_combine, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, out, x, K, n, c);
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
