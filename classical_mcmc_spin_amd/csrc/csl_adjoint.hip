// csl_adjoint.hip — discrete adjoint of the RK4 Landau–Lifshitz integration:
// gradient of  L_b = sum_n sum_j w_j (m_bj(t_n) - m*_j(t_n))^2  (m = the
// 9 observer magnetisations [M_anti, M_local, M_global] at every step) with
// respect to P Hamiltonian parameters theta that enter linearly through the
// bond matrices and fields:  dJ/dtheta_p, dB/dtheta_p are given tables.
//
// One workgroup per trajectory (parameter start).  In one launch it
//   1. integrates forward (the persistent kernel's LDS-resident RK4),
//      recording m_n and storing the state every C steps (checkpoints);
//   2. walks the checkpoint segments backwards: re-integrates the segment
//      from its checkpoint, keeping the C states in global memory, then runs
//      the exact transpose of each RK4 step from its end to its start.
// With C ~ sqrt(n_steps) the memory is O(sqrt(n_steps)) states per
// trajectory for 2x the forward work plus the adjoint steps.
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
#include "csl_device.hpp"

namespace csl {

namespace {

struct Lds3 {
    double* a;  // [3][NP]
    int NP;
    __device__ __forceinline__ void load(int q, double S[3]) const {
        S[0] = a[q];
        S[1] = a[NP + q];
        S[2] = a[2 * NP + q];
    }
    __device__ __forceinline__ void store(int q, const double S[3]) const {
        a[q] = S[0];
        a[NP + q] = S[1];
        a[2 * NP + q] = S[2];
    }
};

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace

constexpr int kAdjMaxP = 8;

template <int SPT>
constexpr int adjoint_max_block() {
    return SPT == 1 ? 1024 : 640;
}

template <int SPT, int ZMAX>
__global__ void __launch_bounds__(adjoint_max_block<SPT>()) k_adjoint(LatDev Lk, AdjArgs A) {
    LatDev L0 = Lk;
    L0.mats += size_t(blockIdx.x) * A.mats_stride;
    L0.field += size_t(blockIdx.x) * A.field_stride;
    const LatDev& L = L0;
    extern __shared__ double smem[];
    const int NP = L.NP;
    const Lds3 bufA{smem, NP}, bufB{smem + 3 * NP, NP}, bufU{smem + 6 * NP, NP};
    double* red = smem + 9 * NP;  // [nwaves][kAdjMaxP + 10]
    constexpr int RW = kAdjMaxP + 10;

    const int b = blockIdx.x;
    const int T = blockDim.x;
    const int tid = threadIdx.x;
    const int wave = uniform(tid >> 6);
    const int lane = tid & (kWave - 1);
    const int nwaves = T >> 6;
    const int P = A.P;
    const double dt = A.dt;
    const double alpha = L.alpha;
    const double inv_n = 1.0 / double(L.N);
    const bool drive = A.fac != nullptr;

    int nb[SPT][ZMAX], cnt[SPT], atm[SPT];
    bool act[SPT];
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

    // drive factor of pulse pl at stage time st (0: t, 1: t+dt/2, 2: t+dt) of step n
    auto fac = [&](int n, int st, int pl) {
        return drive ? A.fac[((size_t(n) * 3 + st) * A.B + b) * 2 + pl] : 0.0;
    };
    // dS/dt of site (m) at state S gathered from buffer y
    auto rhs = [&](const Lds3& y, int m, const double S[3], double f1, double f2, double k[3]) {
        const int p = tid + m * T;
        double H[3];
        field_onsite(L, atm[m], p, S, H);
#pragma unroll
        for (int q = 0; q < ZMAX; ++q)
            if (q < cnt[m]) {
                double Sp[3];
                y.load(nb[m][q], Sp);
                add_bond_fused(as_const(L.mats) + (atm[m] * L.zmax + q) * 9, Sp, H);
            }
        if (drive) sub_drive(L, atm[m], f1, f2, H);
        ll_su2(H, S, alpha, k);
    };

    double x[SPT][3];
    // ---- records: wave sums of the 9 magnetisations of the owned sites
    auto record = [&](int ri) {
        double part[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) part[j] = 0.0;
#pragma unroll
        for (int m = 0; m < SPT; ++m)
            if (act[m]) {
                cptr<double> R = as_const(L.frames) + atm[m] * 9;
                const double sg = as_const(L.afm)[atm[m]];
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
            const double v = wave_sum(part[j]);
            if (lane == 0) red[wave * RW + j] = v;
        }
        __syncthreads();
        if (tid < 9) {
            double v = 0.0;
            for (int w = 0; w < nwaves; ++w) v += red[w * RW + tid];
            A.rec[(size_t(ri) * A.B + b) * 9 + tid] = v * inv_n;
        }
        __syncthreads();
    };

    // ---- one forward RK4 step from x (state also in bufA); leaves x_{n+1} in x and bufA
    auto fwd_step = [&](int n) {
        double acc[SPT][3];
        const double b16 = (1.0 / 6.0) * dt, b13 = (1.0 / 3.0) * dt;
#pragma unroll 1
        for (int s = 0; s < 4; ++s) {
            const Lds3& cur = (s & 1) ? bufB : bufA;
            const Lds3& nxt = (s & 1) ? bufA : bufB;
            const int st = s == 0 ? 0 : (s == 3 ? 2 : 1);
            const double f1 = fac(n, st, 0), f2 = fac(n, st, 1);
            const double cdt = (s == 2) ? 1.0 * dt : 0.5 * dt;
            const double bdt = (s == 0 || s == 3) ? b16 : b13;
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int p = tid + m * T;
                if (act[m]) {
                    double S[3], k[3], y[3];
                    cur.load(p, S);
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
        }
    };

    auto store_state = [&](double* dst) {
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = tid + m * T;
            if (p < NP) {
                dst[p] = x[m][0];
                dst[NP + p] = x[m][1];
                dst[2 * NP + p] = x[m][2];
            }
        }
    };
    auto load_state = [&](const double* src) {
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = tid + m * T;
            if (p < NP) {
                x[m][0] = src[p];
                x[m][1] = src[NP + p];
                x[m][2] = src[2 * NP + p];
                bufA.store(p, x[m]);
            }
        }
        __syncthreads();
    };

    const size_t SZ = size_t(3) * NP;
    double* ckpt = A.ckpt + size_t(b) * A.n_ckpt * SZ;
    double* seg = A.seg + size_t(b) * A.ckpt_interval * SZ;

    // ---- forward pass
    {
        const double* init = A.init + size_t(b) * A.init_stride;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = tid + m * T;
            x[m][0] = x[m][1] = x[m][2] = 0.0;
            if (act[m]) {
                const int cell = p - atm[m] * L.npad;
                const size_t i = size_t(cell) * L.n_atoms + atm[m];
                x[m][0] = init[3 * i];
                x[m][1] = init[3 * i + 1];
                x[m][2] = init[3 * i + 2];
            }
            if (p < NP) bufA.store(p, x[m]);
        }
        __syncthreads();
    }
#pragma unroll 1
    for (int n = 0; n < A.n_steps; ++n) {
        if (n % A.ckpt_interval == 0) store_state(ckpt + size_t(n / A.ckpt_interval) * SZ);
        record(n);
        fwd_step(n);
    }
    record(A.n_steps);

    // ---- backward pass
    double lam[SPT][3];
    double gp[kAdjMaxP];
#pragma unroll
    for (int q = 0; q < kAdjMaxP; ++q) gp[q] = 0.0;
    double loss = 0.0;
    // lambda += d l_n / d x_n; l_n = sum_j w_j (m_j - m*_j)^2.  Thread j < 9
    // reads back the record component it wrote in the forward pass and
    // shares dl_j through LDS.
    auto obs_grad = [&](int n) {
        if (tid < 9) {
            const double d = A.rec[(size_t(n) * A.B + b) * 9 + tid] - A.target[size_t(n) * 9 + tid];
            red[tid] = 2.0 * A.w[tid] * d * inv_n;
            loss += A.w[tid] * d * d;
        }
        __syncthreads();
        double dl[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) dl[j] = red[j];
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            cptr<double> R = as_const(L.frames) + atm[m] * 9;
            const double sg = as_const(L.afm)[atm[m]];
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                double r = dl[3 + e];
#pragma unroll
                for (int d = 0; d < 3; ++d) r += R[d * 3 + e] * (sg * dl[d] + dl[6 + d]);
                lam[m][e] += act[m] ? r : 0.0;
            }
        }
    };
    // phase A of a VJP at state y (in LDS buffer yb) with cotangent c:
    // writes u to bufU, returns the direct term, accumulates parameter grads
    auto vjp_a = [&](const Lds3& yb, const double (&c)[SPT][3], double f1, double f2, double (&dir)[SPT][3]) {
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = tid + m * T;
            double u[3] = {0.0, 0.0, 0.0};
            dir[m][0] = dir[m][1] = dir[m][2] = 0.0;
            if (act[m]) {
                double S[3];
                yb.load(p, S);
                const double* v = c[m];
                double H[3];
                field_onsite(L, atm[m], p, S, H);
                double Sp[ZMAX][3];
#pragma unroll
                for (int q = 0; q < ZMAX; ++q)
                    if (q < cnt[m]) {
                        yb.load(nb[m][q], Sp[q]);
                        add_bond_fused(as_const(L.mats) + (atm[m] * L.zmax + q) * 9, Sp[q], H);
                    }
                if (drive) sub_drive(L, atm[m], f1, f2, H);
                double cc[3], w[3];
                cross3(H, S, cc);
                if (alpha > 0.0) {
                    const double S2 = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
                    const double inv = 1.0 / sqrt(S2);
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
                // parameter gradients: u . (sum_k dJ_kp S_p(k) - dB_p)
#pragma unroll
                for (int q = 0; q < kAdjMaxP; ++q) {
                    if (q < P) {
                    cptr<double> dB = as_const(A.dfield) + (size_t(q) * L.n_atoms + atm[m]) * 3;
                    double dH[3] = {-dB[0], -dB[1], -dB[2]};
#pragma unroll
                    for (int z = 0; z < ZMAX; ++z)
                        if (z < cnt[m])
                            add_bond_fused(as_const(A.dmats) + ((size_t(q) * L.n_atoms + atm[m]) * L.zmax + z) * 9, Sp[z],
                                     dH);
                    gp[q] += u[0] * dH[0] + u[1] * dH[1] + u[2] * dH[2];
                    }
                }
            }
            if (p < NP) bufU.store(p, u);
        }
        __syncthreads();
    };
    // phase B: mu = direct + (2A)^T u + sum_k J_k u_p(k)
    auto vjp_b = [&](const double (&dir)[SPT][3], double (&mu)[SPT][3]) {
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = tid + m * T;
            mu[m][0] = mu[m][1] = mu[m][2] = 0.0;
            if (act[m]) {
                double u[3];
                bufU.load(p, u);
                double r[3] = {dir[m][0], dir[m][1], dir[m][2]};
                if (L.has_onsite) {
                    cptr<double> A2 = as_const(L.onsite) + atm[m] * 9;
                    for (int e = 0; e < 3; ++e)
                        for (int d = 0; d < 3; ++d) r[e] += A2[d * 3 + e] * u[d];
                }
#pragma unroll
                for (int q = 0; q < ZMAX; ++q)
                    if (q < cnt[m]) {
                        double up[3];
                        bufU.load(nb[m][q], up);
                        add_bond_fused(as_const(L.mats) + (atm[m] * L.zmax + q) * 9, up, r);
                    }
                mu[m][0] = r[0];
                mu[m][1] = r[1];
                mu[m][2] = r[2];
            }
        }
    };
    auto store_buf = [&](const Lds3& dst, const double (&v)[SPT][3]) {
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = tid + m * T;
            if (p < NP) dst.store(p, v[m]);
        }
    };

#pragma unroll
    for (int m = 0; m < SPT; ++m) lam[m][0] = lam[m][1] = lam[m][2] = 0.0;
    obs_grad(A.n_steps);

    const int n_seg = (A.n_steps + A.ckpt_interval - 1) / A.ckpt_interval;
#pragma unroll 1
    for (int cseg = n_seg - 1; cseg >= 0; --cseg) {
        const int n0 = cseg * A.ckpt_interval;
        const int n1 = min(n0 + A.ckpt_interval, A.n_steps);
        load_state(ckpt + size_t(cseg) * SZ);
#pragma unroll 1
        for (int n = n0; n < n1; ++n) {
            store_state(seg + size_t(n - n0) * SZ);
            if (n + 1 < n1) fwd_step(n);
        }
        __syncthreads();
#pragma unroll 1
        for (int n = n1 - 1; n >= n0; --n) {
            load_state(seg + size_t(n - n0) * SZ);   // x_n -> x, bufA
            // recompute the stage inputs y2, y3, y4 (bufB, bufA, bufB)
            double y2[SPT][3], yt[SPT][3];
            {
                const double f1 = fac(n, 0, 0), f2 = fac(n, 0, 1);
#pragma unroll
                for (int m = 0; m < SPT; ++m) {
                    double k[3] = {0.0, 0.0, 0.0};
                    if (act[m]) rhs(bufA, m, x[m], f1, f2, k);
                    for (int d = 0; d < 3; ++d) y2[m][d] = x[m][d] + (0.5 * dt) * k[d];
                }
                store_buf(bufB, y2);
                __syncthreads();
            }
            {
                const double f1 = fac(n, 1, 0), f2 = fac(n, 1, 1);
#pragma unroll
                for (int m = 0; m < SPT; ++m) {
                    double k[3] = {0.0, 0.0, 0.0};
                    if (act[m]) rhs(bufB, m, y2[m], f1, f2, k);
                    for (int d = 0; d < 3; ++d) yt[m][d] = x[m][d] + (0.5 * dt) * k[d];
                }
                store_buf(bufA, yt);
                __syncthreads();
            }
            {
                const double f1 = fac(n, 1, 0), f2 = fac(n, 1, 1);
#pragma unroll
                for (int m = 0; m < SPT; ++m) {
                    double k[3] = {0.0, 0.0, 0.0};
                    if (act[m]) {
                        double y3[3];
                        bufA.load(tid + m * T, y3);
                        rhs(bufA, m, y3, f1, f2, k);
                    }
                    for (int d = 0; d < 3; ++d) yt[m][d] = x[m][d] + (1.0 * dt) * k[d];
                }
                store_buf(bufB, yt);
                __syncthreads();
            }
            double c[SPT][3], dir[SPT][3], mu[SPT][3], lnew[SPT][3];
            const double b16 = (1.0 / 6.0) * dt, b13 = (1.0 / 3.0) * dt;
            // stage 4 at y4 (bufB)
#pragma unroll
            for (int m = 0; m < SPT; ++m)
                for (int d = 0; d < 3; ++d) c[m][d] = b16 * lam[m][d];
            vjp_a(bufB, c, fac(n, 2, 0), fac(n, 2, 1), dir);
            vjp_b(dir, mu);
#pragma unroll
            for (int m = 0; m < SPT; ++m)
                for (int d = 0; d < 3; ++d) {
                    lnew[m][d] = lam[m][d] + mu[m][d];
                    c[m][d] = b13 * lam[m][d] + dt * mu[m][d];
                }
            __syncthreads();
            // stage 3 at y3 (bufA)
            vjp_a(bufA, c, fac(n, 1, 0), fac(n, 1, 1), dir);
            vjp_b(dir, mu);
            store_buf(bufB, y2);
#pragma unroll
            for (int m = 0; m < SPT; ++m)
                for (int d = 0; d < 3; ++d) {
                    lnew[m][d] += mu[m][d];
                    c[m][d] = b13 * lam[m][d] + (0.5 * dt) * mu[m][d];
                }
            __syncthreads();
            // stage 2 at y2 (bufB)
            vjp_a(bufB, c, fac(n, 1, 0), fac(n, 1, 1), dir);
            vjp_b(dir, mu);
            store_buf(bufA, x);
#pragma unroll
            for (int m = 0; m < SPT; ++m)
                for (int d = 0; d < 3; ++d) {
                    lnew[m][d] += mu[m][d];
                    c[m][d] = b16 * lam[m][d] + (0.5 * dt) * mu[m][d];
                }
            __syncthreads();
            // stage 1 at x_n (bufA)
            vjp_a(bufA, c, fac(n, 0, 0), fac(n, 0, 1), dir);
            vjp_b(dir, mu);
#pragma unroll
            for (int m = 0; m < SPT; ++m)
                for (int d = 0; d < 3; ++d) lam[m][d] = lnew[m][d] + mu[m][d];
            __syncthreads();
            obs_grad(n);
        }
    }

    // ---- outputs: loss and the block sums of the parameter gradients
#pragma unroll
    for (int q = 0; q < kAdjMaxP; ++q) {
        const double v = wave_sum(gp[q]);
        if (lane == 0) red[wave * RW + q] = v;
    }
    __syncthreads();
    if (tid < P) {
        double v = 0.0;
        for (int w = 0; w < nwaves; ++w) v += red[w * RW + tid];
        A.grad[size_t(b) * P + tid] = v;
    }
    // loss partials live in threads 0..8 (wave 0)
    loss = wave_sum(loss);
    if (tid == 0) A.loss[b] = loss;
    if (A.grad_init) {
        double* gi = A.grad_init + size_t(b) * L.N * 3;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int p = tid + m * T;
            if (act[m]) {
                const int cell = p - atm[m] * L.npad;
                const size_t i = size_t(cell) * L.n_atoms + atm[m];
                gi[3 * i] = lam[m][0];
                gi[3 * i + 1] = lam[m][1];
                gi[3 * i + 2] = lam[m][2];
            }
        }
    }
}

namespace {

template <int SPT, int ZMAX>
hipError_t launch_adj(const LatDev& L, const AdjArgs& A, int block, size_t lds, hipStream_t st) {
    auto kern = k_adjoint<SPT, ZMAX>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(A.B), dim3(block), lds, st, L, A);
    return hipGetLastError();
}

template <int SPT>
hipError_t launch_adj_z(const LatDev& L, const AdjArgs& A, int block, size_t lds, hipStream_t st) {
    if (L.zmax <= 3) return launch_adj<SPT, 3>(L, A, block, lds, st);
    if (L.zmax <= 6) return launch_adj<SPT, 6>(L, A, block, lds, st);
    return hipErrorInvalidValue;
}

}  // namespace

size_t adjoint_lds_bytes(const LatDev& L, int block) {
    return (size_t(9) * L.NP + size_t(block / kWave) * (kAdjMaxP + 10)) * sizeof(double);
}

int adjoint_sites_per_thread(const LatDev& L) {
    if (L.NP > 2048) return 0;  // 9 LDS planes of 3*NP doubles must fit 160 KB
    if (L.NP <= adjoint_max_block<1>()) return 1;
    if ((L.NP + 1) / 2 <= adjoint_max_block<2>()) return 2;
    return 4;
}

hipError_t launch_adjoint(const LatDev& L, const AdjArgs& A, hipStream_t st) {
    const int spt = adjoint_sites_per_thread(L);
    if (spt == 0 || A.P > kAdjMaxP) return hipErrorInvalidValue;
    const int block = ((L.NP + spt - 1) / spt + kWave - 1) / kWave * kWave;
    const size_t lds = adjoint_lds_bytes(L, block);
    switch (spt) {
        case 1: return launch_adj_z<1>(L, A, block, lds, st);
        case 2: return launch_adj_z<2>(L, A, block, lds, st);
        case 4: return launch_adj_z<4>(L, A, block, lds, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace csl
