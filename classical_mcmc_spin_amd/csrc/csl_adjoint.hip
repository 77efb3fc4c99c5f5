// csl_adjoint.hip — discrete adjoint of the RK4 Landau–Lifshitz integration:
// gradient of  L_b = sum_n sum_j w_j (m_bj(t_n) - m*_j(t_n))^2  (m = the
// 9 observer magnetisations [M_anti, M_local, M_global] at every step) with
// respect to P Hamiltonian parameters theta that enter linearly through the
// bond matrices and fields:  dJ/dtheta_p, dB/dtheta_p are given tables.
//
// Kernels (the host runs them in this order):
//   k_adj_fwd    one workgroup per trajectory: forward RK4 with the
//                LDS-resident state, records m_n, and either the stage
//                inputs y_1..y_4 of every step (Y, when they fit in memory)
//                or checkpoints x_n every C steps;
//   per segment of steps, last segment first:
//     k_adj_fwd  (checkpoint mode) re-integrates the segment from its
//                checkpoint, storing its stage inputs in Y (the same
//                arithmetic, so the same bits as the forward pass);
//     k_adj_bwd  one workgroup per trajectory: the transpose of each step
//                from the segment's end to its start, reading y_s from Y
//                (the next stage's state is prefetched into VGPRs while the
//                current one is transposed) and storing the cotangent u of
//                every site's field at every stage in U; lambda is handed
//                between segments through global memory;
//     k_adj_par  all CUs: the parameter terms sum_i u_i . dH_i/dtheta_p of
//                every stage of the segment (independent of each other, so
//                they leave the sequential sweep);
//   k_adj_final  loss and gradient sums in a fixed order.
// The values handed between kernels are exact copies, so every checkpoint
// spacing gives the same bits.
//
// Transpose of one step x' = x + dt/6 k1 + dt/3 k2 + dt/3 k3 + dt/6 k4
// (k_s = f(y_s), y1 = x, y2 = x + dt/2 k1, y3 = x + dt/2 k2, y4 = x + dt k3),
// given lambda' = dL/dx':
//   c4 = dt/6 lambda'             mu4 = J_f(y4)^T c4
//   c3 = dt/3 lambda' + dt mu4    mu3 = J_f(y3)^T c3
//   c2 = dt/3 lambda' + dt/2 mu3  mu2 = J_f(y2)^T c2
//   c1 = dt/6 lambda' + dt/2 mu2  mu1 = J_f(y1)^T c1
//   lambda = lambda' + mu1 + mu2 + mu3 + mu4 (+ dl_n/dx_n)
//   dL/dtheta += sum_s c_s . df/dtheta (y_s)
// For f_i = H_i x S_i - (alpha/|S_i|) S_i x (H_i x S_i) and a cotangent v_i:
//   w   = v - g (v x S),  g = alpha/|S|,  c = H x S
//   u_i = S x w                                   (cotangent of H_i)
//   direct_i = w x H - g (c x v) + (alpha/|S|^3) (v . (S x c)) S
//   (J_f^T v)_i = direct_i + (2A_i)^T u_i + sum_k J_ik u_p(k)
// The last gather uses the lattice's own bond tables: every bond i->j
// carries J and its reverse j->i carries J^T (lattice.h:790-803), so
// sum_{i->j} J_ij^T u_i = sum_{j->i} J_ji u_i.
//   dH_i/dtheta_p = -dB_ip + sum_k dJ_ikp S_p(k)  ->  dL/dtheta_p += u_i . dH_i/dtheta_p
#include <type_traits>

#include "csl_device.hpp"
#include "csl_lds.hpp"

namespace csl {

namespace {

// One LDS buffer of 3 NP doubles: the (x, y) pairs of all sites, then the z
// plane, so a site is one 16-byte and one 8-byte access (NP is a multiple
// of 64, so every buffer of the kernels below starts 16-byte aligned).
struct Lds3 {
    double* a;  // [(x,y)[NP] | z[NP]]
    int NP;
    __device__ __forceinline__ void load(int q, double S[3]) const {
        const double2 v = reinterpret_cast<const double2*>(a)[q];
        S[0] = v.x;
        S[1] = v.y;
        S[2] = a[2 * NP + q];
    }
    __device__ __forceinline__ void store(int q, const double S[3]) const {
        reinterpret_cast<double2*>(a)[q] = make_double2(S[0], S[1]);
        a[2 * NP + q] = S[2];
    }
};

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

constexpr int kAdjMaxP = 8;

template <int SPT>
constexpr int adjoint_max_block() {
    return SPT == 1 ? 1024 : 640;
}

// Site ownership shared by the forward and backward kernels: site p = tid +
// m*T of the device order; its sublattice is wave-uniform.
template <int SPT, int ZMAX>
struct AdjSites {
    int nb[SPT][ZMAX], cnt[SPT], atm[SPT];
    bool act[SPT];
    __device__ __forceinline__ void init(const LatDev& L, int tid, int T, int wave) {
        const int NP = L.NP;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = tid + m * T;
            const int pw = wave * kWave + m * T;
            atm[m] = uniform(pw < NP ? pw / L.npad : 0);
            act[m] = active_site(L, p);
            const int c = uniform(pw < NP ? as_const(L.cnt_atom)[atm[m]] : 0);
            cnt[m] = c;
#pragma unroll
            for (int k = 0; k < ZMAX; ++k) nb[m][k] = (act[m] && k < c) ? L.col[size_t(k) * NP + p] : 0;
        }
    }
};

// One stored stage input in Y (3 NP doubles, 16-byte aligned): the (x, y)
// pairs of all sites, then the z plane -- one 16-byte and one 8-byte access
// per site for the forward's stores and the sweeps' loads.
__device__ __forceinline__ void ys_put(double* y, int NP, int p, const double S[3]) {
    reinterpret_cast<double2*>(y)[p] = make_double2(S[0], S[1]);
    y[2 * NP + p] = S[2];
}
__device__ __forceinline__ void ys_get(const double* y, int NP, int p, double S[3]) {
    const double2 v = reinterpret_cast<const double2*>(y)[p];
    S[0] = v.x;
    S[1] = v.y;
    S[2] = y[2 * NP + p];
}

// The 9 magnetisation sums [M_anti, M_local, M_global] of the state x of the
// thread's sites, summed over the wave into red[wave][9] (DPP, the total in
// lane 63); record_put sums the waves in order after a barrier.  Shared by the
// forward and the records kernel, so both produce the same bits.
template <int SPT, int ZMAX>
__device__ __forceinline__ void adj_record_sum(const LatDev& L, const AdjSites<SPT, ZMAX>& st,
                                               const double (&x)[SPT][3], int wave, int lane, double* red) {
    double part[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) part[j] = 0.0;
#pragma unroll
    for (int m = 0; m < SPT; ++m)
        if (st.act[m]) {
            cptr<double> R = as_const(L.frames) + st.atm[m] * 9;
            const double sg = as_const(L.afm)[st.atm[m]];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double g = R[d * 3] * x[m][0] + R[d * 3 + 1] * x[m][1] + R[d * 3 + 2] * x[m][2];
                part[d] += sg * g;
                part[3 + d] += x[m][d];
                part[6 + d] += g;
            }
        }
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const double v = wave_sum_lane63(part[j]);
        if (lane == kWave - 1) red[wave * 9 + j] = v;
    }
}

__device__ __forceinline__ void adj_record_put(const AdjArgs& A, int b, int ri, int tid, int nwaves,
                                               const double* red, double inv_n) {
    if (tid < 9) {
        double v = 0.0;
        for (int w = 0; w < nwaves; ++w) v += red[w * 9 + tid];
        A.rec[(size_t(ri) * A.B + b) * 9 + tid] = v * inv_n;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// forward: records, checkpoints, stage inputs

// FO: per-sublattice field without an onsite matrix, zmax bonds at every
// site (translation-invariant tables) -- the runtime branches of the field
// and the bond counts disappear at compile time
template <int SPT, int ZMAX, bool FO>
__global__ void __launch_bounds__(adjoint_max_block<SPT>()) k_adj_fwd(LatDev Lk, AdjArgs A) {
    LatDev L0 = Lk;
    L0.mats += size_t(blockIdx.x) * A.mats_stride;
    L0.field += size_t(blockIdx.x) * A.field_stride;
    const LatDev& L = L0;
    extern __shared__ double smem[];
    const int NP = L.NP;
    const Lds3 bufA{smem, NP}, bufB{smem + 3 * NP, NP};
    double* red = smem + 6 * NP;  // [nwaves][9]
    const int b = blockIdx.x;
    const int T = blockDim.x;
    const int tid = threadIdx.x;
    const int wave = uniform(tid >> 6);
    const int lane = tid & (kWave - 1);
    const int nwaves = T >> 6;
    const double dt = A.dt;
    const double alpha = L.alpha;
    const double inv_n = 1.0 / double(L.N);
    const bool drive = A.fac != nullptr;
    const size_t SZ = size_t(3) * NP;
    AdjSites<SPT, ZMAX> st;
    st.init(L, tid, T, wave);

    auto fac = [&](int n, int sg, int pl) {
        return drive ? A.fac[((size_t(n) * 3 + sg) * A.B + b) * 2 + pl] : 0.0;
    };
    auto rhs = [&](const Lds3& y, int m, const double S[3], double f1, double f2, double k[3]) {
        const int p = tid + m * T;
        double H[3];
        if (FO) {
            cptr<double> Bu = as_const(L.field) + st.atm[m] * 3;
            H[0] = -Bu[0];
            H[1] = -Bu[1];
            H[2] = -Bu[2];
        } else {
            field_onsite(L, st.atm[m], p, S, H);
        }
#pragma unroll
        for (int q = 0; q < ZMAX; ++q)
            if (FO || q < st.cnt[m]) {
                double Sp[3];
                y.load(st.nb[m][q], Sp);
                add_bond_fused(as_const(L.mats) + (st.atm[m] * L.zmax + q) * 9, Sp, H);
            }
        if (drive) sub_drive(L, st.atm[m], f1, f2, H);
        ll_su2_rs(H, S, alpha, k);
    };

    double x[SPT][3];
    // wave sums of the 9 magnetisations of x -> red; the sums over the waves
    // are taken by record_put after the next barrier
    auto record_sum = [&]() { adj_record_sum<SPT, ZMAX>(L, st, x, wave, lane, red); };
    auto record_put = [&](int ri) { adj_record_put(A, b, ri, tid, nwaves, red, inv_n); };
    auto store_own = [&](double* dst, const double (&v)[SPT][3]) {
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = fresh(tid + m * T);
            if (p < NP) {
                dst[p] = v[m][0];
                dst[NP + p] = v[m][1];
                dst[2 * NP + p] = v[m][2];
            }
        }
    };

    // initial state: a checkpoint (SoA, device order) or the initial spins
    if (A.ckpt_src >= 0) {
        const double* src = A.ckpt + (size_t(b) * A.n_ckpt + A.ckpt_src) * SZ;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = tid + m * T;
            x[m][0] = x[m][1] = x[m][2] = 0.0;
            if (p < NP) {
                x[m][0] = src[p];
                x[m][1] = src[NP + p];
                x[m][2] = src[2 * NP + p];
                bufA.store(p, x[m]);
            }
        }
    } else {
        const double* init = A.init + size_t(b) * A.init_stride;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = tid + m * T;
            x[m][0] = x[m][1] = x[m][2] = 0.0;
            if (st.act[m]) {
                const int cell = p - st.atm[m] * L.npad;
                const size_t i = size_t(cell) * L.n_atoms + st.atm[m];
                x[m][0] = init[3 * i];
                x[m][1] = init[3 * i + 1];
                x[m][2] = init[3 * i + 2];
            }
            if (p < NP) bufA.store(p, x[m]);
        }
    }
    __syncthreads();

    double* Yb = A.store_stages ? A.Y + size_t(b) * A.seg_len * 4 * SZ : nullptr;
    const double b16 = (1.0 / 6.0) * dt, b13 = (1.0 / 3.0) * dt;
    double acc[SPT][3];
    // drive factors of the step's three stage times, read one step ahead
    double fcur[6], fnext[6];
    auto load_fac = [&](int n, double* f) {
#pragma unroll
        for (int q = 0; q < 6; ++q) f[q] = (n < A.n1) ? fac(n, q >> 1, q & 1) : 0.0;
    };
    load_fac(A.n0, fnext);
    // one RK4 stage of step n (compile-time stage index): the stage input of
    // every owned site is read from LDS (and stored to Y), the next one
    // written to the other buffer
    auto stage = [&](auto s_, int n, bool rec) {
        constexpr int s = decltype(s_)::value;
        const Lds3& cur = (s & 1) ? bufB : bufA;
        const Lds3& nxt = (s & 1) ? bufA : bufB;
        constexpr int sg = s == 0 ? 0 : (s == 3 ? 2 : 1);
        const double f1 = fcur[2 * sg], f2 = fcur[2 * sg + 1];
        const double cdt = (s == 2) ? 1.0 * dt : 0.5 * dt;
        const double bdt = (s == 0 || s == 3) ? b16 : b13;
        double* ys = Yb ? Yb + (size_t(n - A.y0) * 4 + s) * SZ : nullptr;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = fresh(tid + m * T);
            double S[3] = {0.0, 0.0, 0.0};
            if (p < NP) cur.load(p, S);
            if (ys && p < NP) ys_put(ys, NP, p, S);
            if (st.act[m]) {
                double k[3], y[3];
                rhs(cur, m, S, f1, f2, k);
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    if (s == 0) acc[m][d] = x[m][d] + bdt * k[d];
                    else if (s < 3) acc[m][d] = acc[m][d] + bdt * k[d];
                    else x[m][d] = acc[m][d] + bdt * k[d];
                    y[d] = (s < 3) ? x[m][d] + cdt * k[d] : x[m][d];
                }
                nxt.store(p, y);
            }
        }
        __syncthreads();
        if (s == 0 && rec) record_put(n);
    };
#pragma unroll 1
    for (int n = A.n0; n < A.n1; ++n) {
        if (A.ckpt_interval && n % A.ckpt_interval == 0)
            store_own(A.ckpt + (size_t(b) * A.n_ckpt + n / A.ckpt_interval) * SZ, x);
#pragma unroll
        for (int q = 0; q < 6; ++q) fcur[q] = fnext[q];
        load_fac(n + 1, fnext);
        const bool rec = A.record == 1;  // 2: the final record only (the rest from Y by k_adj_rec)
        if (rec) record_sum();
        stage(std::integral_constant<int, 0>{}, n, rec);
        stage(std::integral_constant<int, 1>{}, n, rec);
        stage(std::integral_constant<int, 2>{}, n, rec);
        stage(std::integral_constant<int, 3>{}, n, rec);
    }
    if (A.record && A.n1 == A.n_steps) {
        record_sum();
        __syncthreads();
        record_put(A.n_steps);
    }
}

// Records m_n of the steps n0 .. n1-1 from the stored stage inputs (the
// stage-0 input of step n is x_n), one workgroup per (step, trajectory) with
// the forward's thread layout and summation order: the forward then leaves
// the per-step reductions (9 wave sums and a barrier-bound write every step,
// on its sequential path) to this kernel, which runs on the whole GPU.
template <int SPT, int ZMAX>
__global__ void __launch_bounds__(adjoint_max_block<SPT>()) k_adj_rec(LatDev L, AdjArgs A) {
    __shared__ double red[(adjoint_max_block<SPT>() / kWave) * 9];
    const int n = A.n0 + int(blockIdx.x);
    const int b = blockIdx.y;
    const int T = blockDim.x;
    const int tid = threadIdx.x;
    const int wave = uniform(tid >> 6);
    const int lane = tid & (kWave - 1);
    const int nwaves = T >> 6;
    const int NP = L.NP;
    const size_t SZ = size_t(3) * NP;
    AdjSites<SPT, ZMAX> st;
    st.init(L, tid, T, wave);
    const double* y = A.Y + (size_t(b) * A.seg_len * 4 + size_t(n - A.y0) * 4) * SZ;
    double x[SPT][3];
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        const int p = tid + m * T;
        x[m][0] = x[m][1] = x[m][2] = 0.0;
        if (st.act[m]) ys_get(y, NP, p, x[m]);
    }
    adj_record_sum<SPT, ZMAX>(L, st, x, wave, lane, red);
    __syncthreads();
    adj_record_put(A, b, n, tid, nwaves, red, 1.0 / double(L.N));
}

// ---------------------------------------------------------------------------
// backward: transposed steps of one segment
//
// With A(s) the first half of a VJP at y_s (field, u_s -> LDS, direct term)
// and B(s) the second (gather of the partners' u_s), the transposed step is
// A(3) B(3) A(2) B(2) A(1) B(1) A(0) B(0).  B(s+1) needs only the site's own
// cotangent and LDS values written before the last barrier, so every phase
// runs B of one stage and A of the next with ONE barrier: two u buffers and
// two y buffers alternate (y of the next phase is fetched from Y into VGPRs
// at the start of a phase and stored into the free buffer at its end).  The
// loss terms of step n (records against the target) are read by 9 threads
// two phases ahead of their use and passed through LDS (two parities).

template <int SPT, int ZMAX, bool FO>
__global__ void __launch_bounds__(adjoint_max_block<SPT>()) k_adj_bwd(LatDev Lk, AdjArgs A) {
    LatDev L0 = Lk;
    L0.mats += size_t(blockIdx.x) * A.mats_stride;
    L0.field += size_t(blockIdx.x) * A.field_stride;
    const LatDev& L = L0;
    extern __shared__ double smem[];
    const int NP = L.NP;
    // two stage-input and two cotangent buffers, chosen by parity with a
    // select (an array of them indexed at run time would live in scratch)
    auto bufY = [&](int k) { return Lds3{smem + (k & 1) * 3 * NP, NP}; };
    auto bufU = [&](int k) { return Lds3{smem + (6 + (k & 1) * 3) * NP, NP}; };
    double* red = smem + 12 * NP;  // [2 parities][18]: dl_j, then w_j d_j^2
    const int b = blockIdx.x;
    const int T = blockDim.x;
    const int tid = threadIdx.x;
    const int wave = uniform(tid >> 6);
    const double dt = A.dt;
    const double alpha = L.alpha;
    const double inv_n = 1.0 / double(L.N);
    const bool drive = A.fac != nullptr;
    const size_t SZ = size_t(3) * NP;
    AdjSites<SPT, ZMAX> st;
    st.init(L, tid, T, wave);
    auto fac = [&](int n, int sg, int pl) {
        return drive ? A.fac[((size_t(n) * 3 + sg) * A.B + b) * 2 + pl] : 0.0;
    };
    const double* Yb = A.Y + size_t(b) * A.seg_len * 4 * SZ;
    double* Ub = A.U + size_t(b) * A.seg_len * 4 * SZ;
    double* lamg = A.lam + size_t(b) * SZ;
    const double b16 = (1.0 / 6.0) * dt, b13 = (1.0 / 3.0) * dt;

    double lam[SPT][3], acc[SPT][3], c[SPT][3], dir[SPT][3];

    // loss terms of record n -> red[n & 1] (threads j < 9); read after two barriers
    auto obs_load = [&](int n) {
        if (tid < 9) {
            const double d = A.rec[(size_t(n) * A.B + b) * 9 + tid] - A.target[size_t(n) * 9 + tid];
            red[(n & 1) * 18 + tid] = 2.0 * A.w[tid] * d * inv_n;
            red[(n & 1) * 18 + 9 + tid] = A.w[tid] * d * d;
        }
    };
    // lambda += d l_n / d x_n; thread 0 keeps the loss term of record n
    auto obs_apply = [&](int n) {
        const double* r = red + (n & 1) * 18;
        if (tid == 0) {
            double l = 0.0;
            for (int j = 0; j < 9; ++j) l += r[9 + j];
            A.loss_n[size_t(b) * (A.n_steps + 1) + n] = l;
        }
        double dl[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) dl[j] = r[j];
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            cptr<double> R = as_const(L.frames) + st.atm[m] * 9;
            const double sg = as_const(L.afm)[st.atm[m]];
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                double v = dl[3 + e];
#pragma unroll
                for (int d = 0; d < 3; ++d) v += R[d * 3 + e] * (sg * dl[d] + dl[6 + d]);
                lam[m][e] += st.act[m] ? v : 0.0;
            }
        }
    };
    // A: field at the state in yb, u -> ub and U (slot ug), direct term -> dir
    auto vjp_a = [&](const Lds3& yb, const Lds3& ub, double f1, double f2, double* ug) {
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = fresh(tid + m * T);
            double u[3] = {0.0, 0.0, 0.0};
            dir[m][0] = dir[m][1] = dir[m][2] = 0.0;
            if (st.act[m]) {
                double S[3];
                yb.load(p, S);
                const double* v = c[m];
                double H[3];
                if (FO) {
                    cptr<double> Bu = as_const(L.field) + st.atm[m] * 3;
                    H[0] = -Bu[0];
                    H[1] = -Bu[1];
                    H[2] = -Bu[2];
                } else {
                    field_onsite(L, st.atm[m], p, S, H);
                }
#pragma unroll
                for (int q = 0; q < ZMAX; ++q)
                    if (FO || q < st.cnt[m]) {
                        double Sp[3];
                        yb.load(fresh(st.nb[m][q]), Sp);
                        add_bond_fused(as_const(L.mats) + (st.atm[m] * L.zmax + q) * 9, Sp, H);
                    }
                if (drive) sub_drive(L, st.atm[m], f1, f2, H);
                double cc[3], w[3];
                cross3(H, S, cc);
                if (alpha > 0.0) {
                    const double S2 = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
                    const double inv = rsqrt_nr(S2);
                    const double g = alpha * inv;
                    double vxS[3], cxv[3], Sxc[3];
                    cross3(v, S, vxS);
                    cross3(cc, v, cxv);
                    cross3(S, cc, Sxc);
                    for (int d = 0; d < 3; ++d) w[d] = v[d] - g * vxS[d];
                    double wxH[3];
                    cross3(w, H, wxH);
                    const double k3 = alpha * inv * inv * inv * (v[0] * Sxc[0] + v[1] * Sxc[1] + v[2] * Sxc[2]);
                    for (int d = 0; d < 3; ++d) dir[m][d] = wxH[d] - g * cxv[d] + k3 * S[d];
                } else {
                    for (int d = 0; d < 3; ++d) w[d] = v[d];
                    cross3(w, H, dir[m]);
                }
                cross3(S, w, u);
            }
            if (p < NP) {
                ub.store(p, u);
                ys_put(ug, NP, p, u);
            }
        }
    };
    // B: mu = direct + (2A)^T u + sum_k J_k u_p(k), then the next cotangent
    // (K = 3, 2, 1: the stage just finished; K = 0 ends the step)
    auto vjp_b = [&](const Lds3& ub, int K) {
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = fresh(tid + m * T);
            double mu[3] = {0.0, 0.0, 0.0};
            if (st.act[m]) {
                double u[3];
                ub.load(p, u);
                double r[3] = {dir[m][0], dir[m][1], dir[m][2]};
                if (!FO && L.has_onsite) {
                    cptr<double> A2 = as_const(L.onsite) + st.atm[m] * 9;
                    for (int e = 0; e < 3; ++e)
                        for (int d = 0; d < 3; ++d) r[e] += A2[d * 3 + e] * u[d];
                }
#pragma unroll
                for (int q = 0; q < ZMAX; ++q)
                    if (FO || q < st.cnt[m]) {
                        double up[3];
                        ub.load(fresh(st.nb[m][q]), up);
                        add_bond_fused(as_const(L.mats) + (st.atm[m] * L.zmax + q) * 9, up, r);
                    }
                mu[0] = r[0];
                mu[1] = r[1];
                mu[2] = r[2];
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (K == 3) {
                    acc[m][d] = lam[m][d] + mu[d];
                    c[m][d] = b13 * lam[m][d] + dt * mu[d];
                } else if (K == 2) {
                    acc[m][d] += mu[d];
                    c[m][d] = b13 * lam[m][d] + (0.5 * dt) * mu[d];
                } else if (K == 1) {
                    acc[m][d] += mu[d];
                    c[m][d] = b16 * lam[m][d] + (0.5 * dt) * mu[d];
                } else {
                    lam[m][d] = acc[m][d] + mu[d];
                }
            }
        }
    };
    double pre[SPT][3];
    auto fetch = [&](int n, int s) {
        const double* src = Yb + (size_t(n - A.y0) * 4 + s) * SZ;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = fresh(tid + m * T);
            pre[m][0] = pre[m][1] = pre[m][2] = 0.0;
            if (p < NP) ys_get(src, NP, p, pre[m]);
        }
    };
    auto put = [&](const Lds3& dst) {
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = fresh(tid + m * T);
            if (p < NP) dst.store(p, pre[m]);
        }
    };
    auto set_c = [&](double w) {
#pragma unroll
        for (int m = 0; m < SPT; ++m)
            for (int d = 0; d < 3; ++d) c[m][d] = w * lam[m][d];
    };
    auto uslot = [&](int n, int s) { return Ub + (size_t(n - A.y0) * 4 + s) * SZ; };

    // lambda at the end of the segment
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        const int p = tid + m * T;
        lam[m][0] = lam[m][1] = lam[m][2] = 0.0;
        if (!A.first && p < NP) {
            lam[m][0] = lamg[p];
            lam[m][1] = lamg[NP + p];
            lam[m][2] = lamg[2 * NP + p];
        }
    }
    // Phases of step n (one barrier each): B(3)+A(2), B(2)+A(1), B(1)+A(0),
    // B(0)+A(3) of step n-1.  y of the next phase is fetched from Y at the
    // start of a phase and stored into the free LDS buffer at its end; the
    // records of step n are read in its first phase and their loss terms
    // stored to LDS in the third (used in the fourth).
    const int nl = A.n1 - 1;
    double orec = 0.0, otgt = 0.0;
    if (A.first) obs_load(A.n_steps);
    fetch(nl, 3);
    put(bufY(0));
    __syncthreads();
    if (A.first) obs_apply(A.n_steps);
    // A(nl, 3)
    set_c(b16);
    fetch(nl, 2);
    vjp_a(bufY(0), bufU(0), fac(nl, 2, 0), fac(nl, 2, 1), uslot(nl, 3));
    put(bufY(1));
    __syncthreads();
#pragma unroll 1
    for (int n = nl; n >= A.n0; --n) {
        // B(3) + A(2) at y_3
        fetch(n, 1);
        if (tid < 9) {
            orec = A.rec[(size_t(n) * A.B + b) * 9 + tid];
            otgt = A.target[size_t(n) * 9 + tid];
        }
        vjp_b(bufU(0), 3);
        vjp_a(bufY(1), bufU(1), fac(n, 1, 0), fac(n, 1, 1), uslot(n, 2));
        put(bufY(0));
        __syncthreads();
        // B(2) + A(1) at y_2
        fetch(n, 0);
        vjp_b(bufU(1), 2);
        vjp_a(bufY(0), bufU(0), fac(n, 1, 0), fac(n, 1, 1), uslot(n, 1));
        put(bufY(1));
        if (tid < 9) {
            const double d = orec - otgt;
            red[(n & 1) * 18 + tid] = 2.0 * A.w[tid] * d * inv_n;
            red[(n & 1) * 18 + 9 + tid] = A.w[tid] * d * d;
        }
        __syncthreads();
        // B(1) + A(0) at y_1 = x_n
        if (n > A.n0) fetch(n - 1, 3);
        vjp_b(bufU(0), 1);
        vjp_a(bufY(1), bufU(1), fac(n, 0, 0), fac(n, 0, 1), uslot(n, 0));
        if (n > A.n0) put(bufY(0));
        __syncthreads();
        // B(0) ends step n: lambda_n = lambda' + mu's + dl_n; then A(n-1, 3)
        vjp_b(bufU(1), 0);
        obs_apply(n);
        if (n > A.n0) {
            set_c(b16);
            fetch(n - 1, 2);
            vjp_a(bufY(0), bufU(0), fac(n - 1, 2, 0), fac(n - 1, 2, 1), uslot(n - 1, 3));
            put(bufY(1));
        }
        __syncthreads();
    }

    // hand lambda to the next (earlier) segment, or the initial-state gradient
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        const int p = tid + m * T;
        if (p < NP) {
            lamg[p] = lam[m][0];
            lamg[NP + p] = lam[m][1];
            lamg[2 * NP + p] = lam[m][2];
        }
    }
    if (A.n0 == 0 && A.grad_init) {
        double* gi = A.grad_init + size_t(b) * L.N * 3;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = tid + m * T;
            if (st.act[m]) {
                const int cell = p - st.atm[m] * L.npad;
                const size_t i = size_t(cell) * L.n_atoms + st.atm[m];
                gi[3 * i] = lam[m][0];
                gi[3 * i + 1] = lam[m][1];
                gi[3 * i + 2] = lam[m][2];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// parameter-gradient terms: one stage slot of one trajectory at a time,
// u . (sum_k dJ_kp S_p(k) - dB_p) summed over the sites (partners of y_s read
// from Y, L2-resident)

template <int ZMAX>
__global__ void __launch_bounds__(256) k_adj_par(LatDev L, AdjArgs A) {
    __shared__ double red[4][kAdjMaxP];
    const int b = blockIdx.y;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid >> 6;
    const int NP = L.NP;
    const size_t SZ = size_t(3) * NP;
    const int P = A.P;
    const int items = (A.n1 - A.n0) * 4;
    const double* Yb = A.Y + size_t(b) * A.seg_len * 4 * SZ;
    const double* Ub = A.U + size_t(b) * A.seg_len * 4 * SZ;
    const size_t off = size_t(A.n0 - A.y0) * 4;  // slot of step n0 in Y and U
    for (int it = blockIdx.x; it < items; it += gridDim.x) {
        const double* y = Yb + (off + it) * SZ;
        const double* u = Ub + (off + it) * SZ;
        double g[kAdjMaxP];
#pragma unroll
        for (int q = 0; q < kAdjMaxP; ++q) g[q] = 0.0;
        for (int p0 = 0; p0 < NP; p0 += 256) {
            const int p = p0 + tid;
            const int pw = p0 + wave * kWave;
            const int atom = uniform(pw < NP ? pw / L.npad : 0);
            if (!active_site(L, p)) continue;
            const int cnt = uniform(as_const(L.cnt_atom)[atom]);
            double uu[3];
            ys_get(u, NP, p, uu);
            double Sp[ZMAX][3];
#pragma unroll
            for (int z = 0; z < ZMAX; ++z)
                if (z < cnt) {
                    const int q = L.col[size_t(z) * NP + p];
                    ys_get(y, NP, q, Sp[z]);
                }
#pragma unroll
            for (int q = 0; q < kAdjMaxP; ++q) {
                if (q < P) {
                    cptr<double> dB = as_const(A.dfield) + (size_t(q) * L.n_atoms + atom) * 3;
                    double dH[3] = {-dB[0], -dB[1], -dB[2]};
#pragma unroll
                    for (int z = 0; z < ZMAX; ++z)
                        if (z < cnt)
                            add_bond_fused(as_const(A.dmats) + ((size_t(q) * L.n_atoms + atom) * L.zmax + z) * 9,
                                           Sp[z], dH);
                    g[q] += uu[0] * dH[0] + uu[1] * dH[1] + uu[2] * dH[2];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kAdjMaxP; ++q) {
            if (q < P) {
                const double v = wave_sum(g[q]);
                if (lane == 0) red[wave][q] = v;
            }
        }
        __syncthreads();
        if (tid < P) {
            const double v = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
            const int n = A.n0 + it / 4, s = it % 4;
            A.part[((size_t(b) * A.n_steps + n) * 4 + s) * P + tid] = v;
        }
        __syncthreads();
    }
}

// loss and gradient of every trajectory: each of the 256 threads sums a
// strided share of the records / stage terms, then a tree over the threads
// (a fixed order, so the same bits for every segment layout)
__global__ void __launch_bounds__(256) k_adj_final(AdjArgs A) {
    __shared__ double buf[256];
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    auto tree = [&](double v) {
        buf[tid] = v;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (tid < h) buf[tid] += buf[tid + h];
            __syncthreads();
        }
        const double r = buf[0];
        __syncthreads();
        return r;
    };
    double l = 0.0;
    for (int n = tid; n <= A.n_steps; n += 256) l += A.loss_n[size_t(b) * (A.n_steps + 1) + n];
    l = tree(l);
    if (tid == 0) A.loss[b] = l;
    const size_t items = size_t(A.n_steps) * 4;
    for (int q = 0; q < A.P; ++q) {
        double g = 0.0;
        for (size_t it = tid; it < items; it += 256) g += A.part[(size_t(b) * items + it) * A.P + q];
        g = tree(g);
        if (tid == 0) A.grad[size_t(b) * A.P + q] = g;
    }
}

namespace {

// Block size for spt sites per thread: the multiple of 64 between the
// smallest that covers NP and the kernels' bound whose busiest SIMD holds the
// fewest wave-sites (a workgroup's waves are dealt to the 4 SIMDs in turn; a
// wave-site is one wave's 64 lanes of one of its spt sites, counted when any
// lane is a site), the smaller block on ties.  NP = 1152, spt 2: 576 threads
// load the SIMDs 6, 4, 4, 4 wave-sites, 640 threads 5, 5, 4, 4.
int adj_block(const LatDev& L, int spt) {
    const int lo = ((L.NP + spt - 1) / spt + kWave - 1) / kWave * kWave;
    const int hi = spt == 1 ? adjoint_max_block<1>() : adjoint_max_block<2>();
    int best = lo, best_load = 1 << 30;
    for (int T = lo; T <= hi; T += kWave) {
        int simd[4] = {0, 0, 0, 0};
        for (int w = 0; w < T / kWave; ++w)
            for (int m = 0; m < spt; ++m)
                if (w * kWave + m * T < L.NP) ++simd[w % 4];
        const int load = std::max(std::max(simd[0], simd[1]), std::max(simd[2], simd[3]));
        if (load < best_load) {
            best_load = load;
            best = T;
        }
    }
    return best;
}

template <int SPT, int ZMAX>
hipError_t launch_adj2(const LatDev& L, const AdjArgs& A, bool fwd, hipStream_t st) {
    const int block = adj_block(L, SPT);
    const size_t lds = adjoint_lds_bytes(L, block);
    const bool fo = L.site_uniform && !L.has_onsite && L.full_coord && L.zmax == ZMAX;
    auto kern = fwd ? (fo ? k_adj_fwd<SPT, ZMAX, true> : k_adj_fwd<SPT, ZMAX, false>)
                    : (fo ? k_adj_bwd<SPT, ZMAX, true> : k_adj_bwd<SPT, ZMAX, false>);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(A.B), dim3(block), lds, st, L, A);
    return hipGetLastError();
}

template <int SPT>
hipError_t launch_adj_z(const LatDev& L, const AdjArgs& A, bool fwd, hipStream_t st) {
    if (L.zmax <= 3) return launch_adj2<SPT, 3>(L, A, fwd, st);
    if (L.zmax <= 6) return launch_adj2<SPT, 6>(L, A, fwd, st);
    return hipErrorInvalidValue;
}

hipError_t launch_adj_spt(const LatDev& L, const AdjArgs& A, bool fwd, hipStream_t st) {
    const int spt = adjoint_sites_per_thread(L);
    if (spt == 0 || A.P > kAdjMaxP) return hipErrorInvalidValue;
    switch (spt) {
        case 1: return launch_adj_z<1>(L, A, fwd, st);
        case 2: return launch_adj_z<2>(L, A, fwd, st);
        case 4: return launch_adj_z<4>(L, A, fwd, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace

size_t adjoint_lds_bytes(const LatDev& L, int block) {
    return (size_t(12) * L.NP + size_t(block / kWave) * 9 + 36) * sizeof(double);
}

int adjoint_sites_per_thread(const LatDev& L) {
    if (L.NP > 1664) return 0;  // 4 LDS state planes of 3*NP doubles must fit 160 KB
    if (L.NP <= adjoint_max_block<1>()) return 1;
    if ((L.NP + 1) / 2 <= adjoint_max_block<2>()) return 2;
    return 4;
}

hipError_t launch_adjoint_forward(const LatDev& L, const AdjArgs& A, hipStream_t st) {
    return launch_adj_spt(L, A, true, st);
}

hipError_t launch_adjoint_records(const LatDev& L, const AdjArgs& A, hipStream_t st) {
    const int spt = adjoint_sites_per_thread(L);
    if (A.n1 <= A.n0) return hipSuccess;
    const int block = adj_block(L, spt > 0 ? spt : 1);
    const dim3 grid(unsigned(A.n1 - A.n0), unsigned(A.B));
    if (spt == 1 && L.zmax <= 3) hipLaunchKernelGGL((k_adj_rec<1, 3>), grid, dim3(block), 0, st, L, A);
    else if (spt == 1 && L.zmax <= 6) hipLaunchKernelGGL((k_adj_rec<1, 6>), grid, dim3(block), 0, st, L, A);
    else if (spt == 2 && L.zmax <= 3) hipLaunchKernelGGL((k_adj_rec<2, 3>), grid, dim3(block), 0, st, L, A);
    else if (spt == 2 && L.zmax <= 6) hipLaunchKernelGGL((k_adj_rec<2, 6>), grid, dim3(block), 0, st, L, A);
    else if (spt == 4 && L.zmax <= 3) hipLaunchKernelGGL((k_adj_rec<4, 3>), grid, dim3(block), 0, st, L, A);
    else if (spt == 4 && L.zmax <= 6) hipLaunchKernelGGL((k_adj_rec<4, 6>), grid, dim3(block), 0, st, L, A);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_adjoint_backward(const LatDev& L, const AdjArgs& A, hipStream_t st) {
    return launch_adj_spt(L, A, false, st);
}

hipError_t launch_adjoint_params(const LatDev& L, const AdjArgs& A, int blocks_per_traj, hipStream_t st) {
    if (A.P == 0) return hipSuccess;
    if (A.P > kAdjMaxP) return hipErrorInvalidValue;
    const int items = (A.n1 - A.n0) * 4;
    const unsigned gx = unsigned(std::max(1, std::min(items, blocks_per_traj)));
    if (L.zmax <= 3) hipLaunchKernelGGL(k_adj_par<3>, dim3(gx, A.B), dim3(256), 0, st, L, A);
    else if (L.zmax <= 6) hipLaunchKernelGGL(k_adj_par<6>, dim3(gx, A.B), dim3(256), 0, st, L, A);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_adjoint_final(const AdjArgs& A, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_final, dim3(A.B), dim3(256), 0, st, A);
    return hipGetLastError();
}

}  // namespace csl
