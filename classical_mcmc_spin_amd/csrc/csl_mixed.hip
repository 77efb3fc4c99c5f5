// csl_mixed.hip — mixed SU(2) + SU(3) Landau–Lifshitz RK4 path: device
// layout, gfx950 stage kernels and the C ABI of include/csl_mixed.h.
//
// Device layout: both species are renumbered sublattice-major like the
// SU(2) path (p = atom * npad + cell, npad = n_cells rounded up to 64), so a
// wavefront holds 64 sites of one sublattice of one species and the
// per-(sublattice, slot) bond matrices of translation-invariant tables are
// wave-uniform (scalar loads through the constant address space).  One
// trajectory's state is [3][NP2] SU(2) planes followed by [8][NP3] SU(3)
// planes, FP64.
//
// The state of the reference workload (2048 SU(2) + 2048 SU(3) sites,
// 180 KB) exceeds one workgroup's LDS, so the path runs grid-wide stage
// kernels: one launch per RK4 stage covers the SU(2) blocks and the SU(3)
// blocks of every trajectory (B in the grid's second dimension), with the
// drive envelopes evaluated once per workgroup and the M(t)/E(t) records
// reduced per workgroup and then per trajectory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/csl_mixed.h"
#include "csl_host.hpp"
#include "csl_internal.hpp"
#include "csl_kernels.hpp"
#include "csl_mixed_device.hpp"
#include "csl_tableau.hpp"

namespace csl {

struct MixStageArgs {
    int stage;                 // 0..3
    int nblk2;                 // workgroups of the SU(2) part
    double* x;                 // [B][SZ] state at step start
    double* acc;               // RK4 accumulator
    const double* ycur;        // stage input (x itself in stage 0)
    double* ynext;             // next stage input
    double t, dt;              // stage time, step
    const double* tp;          // [B][2] SU(2) pulse centres or null (no drive)
    const double* tp3;         // [B][2] SU(3) pulse centres
    const double* scale;       // [B][2] pulse scales or null (1)
    int rec;                   // stage 0 of a recorded step: write partials
    double* partials;          // [B][nblk][34]
    // tangent-linear run (k_mixed_stage<true>): z = d state / d theta for a
    // common scale theta of both Fe-Tm bond kinds (e23, e32), at theta = 1
    double* zx;                // [B][SZ] tangent at step start
    double* zacc;              // tangent RK4 accumulator
    const double* zcur;        // tangent stage input
    double* znext;             // tangent next stage input
    double* tpartials;         // [B][nblk][34] record partials of z (entry 33 stays 0)
};

__device__ __forceinline__ size_t mix_stride(const MixDev& M) {
    return size_t(3) * M.NP2 + size_t(8) * M.NP3;
}

// Loads and stores of values other workgroups exchange within one launch
// (the cooperative kernel): COH = agent scope, performed at the device's
// coherence point (the writers' L2 may sit on another XCD); otherwise plain.
template <bool COH>
__device__ __forceinline__ double ld_x(const double* p) {
    if constexpr (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}
template <bool COH>
__device__ __forceinline__ void st_x(double* p, double v) {
    if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

// y_p += sum_k J_k x_partner(k), J m = DA x DB, accumulated straight into y
// (one FMA chain per component)
template <int DA, int DB, bool COH = false>
__device__ __forceinline__ void gather_bonds(const Ell& E, const double* __restrict__ src, int NPsrc,
                                             int NPpart, int atom, int p, double* H) {
    const int c = E.uniform ? E.cnt_atom[atom] : E.cnt[p];
    for (int k = 0; k < c; ++k) {
        const int q = E.col[size_t(k) * NPsrc + p];
        double s[DB];
#pragma unroll
        for (int e = 0; e < DB; ++e) s[e] = ld_x<COH>(src + size_t(e) * NPpart + q);
        if (E.uniform) {
            const cptr<double> J = as_const(E.mats) + (size_t(atom) * E.zmax + k) * (DA * DB);
#pragma unroll
            for (int a = 0; a < DA; ++a)
#pragma unroll
                for (int e = 0; e < DB; ++e) H[a] = fma(J[a * DB + e], s[e], H[a]);
        } else {
            const double* J = E.mats + (size_t(k) * NPsrc + p) * (DA * DB);
#pragma unroll
            for (int a = 0; a < DA; ++a)
#pragma unroll
                for (int e = 0; e < DB; ++e) H[a] = fma(J[a * DB + e], s[e], H[a]);
        }
    }
}

// -B + 2A S of one site (tables hold 2A)
template <int D>
__device__ __forceinline__ void field_onsite(const double* __restrict__ field, const double* __restrict__ onsite,
                                             int site_uniform, int has_onsite, int atom, int p,
                                             const double* S, double* H) {
    if (site_uniform) {
        const cptr<double> f = as_const(field) + size_t(atom) * D;
#pragma unroll
        for (int a = 0; a < D; ++a) H[a] = -f[a];
        if (has_onsite) {
            const cptr<double> A = as_const(onsite) + size_t(atom) * D * D;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                double r = 0.0;
#pragma unroll
                for (int e = 0; e < D; ++e) r += A[a * D + e] * S[e];
                H[a] += r;
            }
        }
    } else {
        const double* f = field + size_t(p) * D;
#pragma unroll
        for (int a = 0; a < D; ++a) H[a] = -f[a];
        if (has_onsite) {
            const double* A = onsite + size_t(p) * D * D;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                double r = 0.0;
#pragma unroll
                for (int e = 0; e < D; ++e) r += A[a * D + e] * S[e];
                H[a] += r;
            }
        }
    }
}

// 2A dS of one site (the linear part of field_onsite; tables hold 2A)
template <int D>
__device__ __forceinline__ void onsite_lin(const double* __restrict__ onsite, int site_uniform, int has_onsite,
                                           int atom, int p, const double* dS, double* dH) {
    if (!has_onsite) return;
    if (site_uniform) {
        const cptr<double> A = as_const(onsite) + size_t(atom) * D * D;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            double r = 0.0;
#pragma unroll
            for (int e = 0; e < D; ++e) r += A[a * D + e] * dS[e];
            dH[a] += r;
        }
    } else {
        const double* A = onsite + size_t(p) * D * D;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            double r = 0.0;
#pragma unroll
            for (int e = 0; e < D; ++e) r += A[a * D + e] * dS[e];
            dH[a] += r;
        }
    }
}

// SU(2) local field without drive at device site p of one trajectory y
template <bool COH = false>
__device__ __forceinline__ void field_su2(const MixDev& M, const double* __restrict__ y, int atom, int p,
                                          const double* S, double* H) {
    field_onsite<3>(M.field2, M.onsite2, M.site_uniform2, M.has_onsite2, atom, p, S, H);
    gather_bonds<3, 3, COH>(M.e22, y, M.NP2, M.NP2, atom, p, H);
    gather_bonds<3, 8, COH>(M.e23, y + size_t(3) * M.NP2, M.NP2, M.NP3, atom, p, H);
}

template <bool COH = false>
__device__ __forceinline__ void field_su3(const MixDev& M, const double* __restrict__ y, int atom, int p,
                                          const double* S, double* H) {
    field_onsite<8>(M.field3, M.onsite3, M.site_uniform3, M.has_onsite3, atom, p, S, H);
    gather_bonds<8, 8, COH>(M.e33, y + size_t(3) * M.NP2, M.NP3, M.NP3, atom, p, H);
    gather_bonds<8, 3, COH>(M.e32, y, M.NP3, M.NP2, atom, p, H);
}

// Workgroup -> (block, trajectory) with the XCD-contiguous mapping of k_stage.
__device__ __forceinline__ void mix_block(int& bx, int& by) {
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

// RK4 update of one state component (stage s) from its derivative kd
__device__ __forceinline__ void rk_update(int s, size_t q, double Sd, double kd, double bdt, double cdt,
                                          double* x, double* acc, double* nxt) {
    if (s == 0) {
        acc[q] = Sd + bdt * kd;
        nxt[q] = Sd + cdt * kd;
    } else if (s < 3) {
        const double xd = x[q];
        acc[q] = acc[q] + bdt * kd;
        nxt[q] = xd + cdt * kd;
    } else {
        x[q] = acc[q] + bdt * kd;
    }
}

// One RK4 stage of every SU(2) and SU(3) site of every trajectory.  With TL
// the same launch also advances the tangent z = d state / d theta (theta the
// scale of both Fe-Tm bond kinds): dk = J_f(y) z + d f / d theta, with
//   SU(2)  dH = 2A dS + sum J22 dS_p + sum J23 dl_p + sum J23 l_p,
//          dk = dH x S + H x dS (+ the linearised Gilbert term)
//   SU(3)  dH = 2A dl + sum J33 dl_p + sum J32 dS_p + sum J32 S_p,
//          dk = f(dH, l) + f(H, dl) - Gamma dl
// and the same RK4 coefficients, i.e. the exact derivative of the discrete
// scheme (what a discrete adjoint computes for one parameter).
//
// mixed_stage_wg is the work of one workgroup (block bx of trajectory b);
// with COH the values other workgroups read (next stage values, the state
// written in stage 3, record partials) are stored and gathered at agent
// scope, for the cooperative kernel that runs all stages in one launch.
template <bool TL, bool COH>
__device__ __forceinline__ void mixed_stage_wg(const MixDev& M, const MixStageArgs& A, int bx, int b, double* fac,
                                               double (*red)[kRec]) {
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = uniform(tid >> 6);
    const bool is3 = bx >= A.nblk2;                       // uniform per workgroup
    const size_t off = size_t(b) * mix_stride(M);
    const double* cur = A.ycur + off;
    double* nxt = A.ynext + off;
    double* x = A.x + off;
    double* acc = A.acc + off;
    if (A.tp) {  // drive envelopes of this trajectory at the stage time
        if (tid < 4) {
            const int sp = tid >> 1, k = tid & 1;
            const double amp = sp ? M.amp3 : M.amp2;
            double f = 0.0;
            if (amp != 0.0) {
                f = envelope(amp, A.t, (sp ? A.tp3 : A.tp)[size_t(b) * 2 + k], sp ? M.width3 : M.width2,
                             sp ? M.freq3 : M.freq2);
                if (A.scale) f *= A.scale[size_t(b) * 2 + k];
            }
            fac[tid] = f;
        }
        __syncthreads();
    }
    const int s = A.stage;
    const double dt = A.dt;
    const double cdt = (s == 2) ? 1.0 * dt : 0.5 * dt;
    const double bdt = (s == 0 || s == 3) ? (1.0 / 6.0) * dt : (1.0 / 3.0) * dt;
    double part[kRec], tpart[kRec];
#pragma unroll
    for (int j = 0; j < kRec; ++j) part[j] = tpart[j] = 0.0;
    const double* zc = TL ? A.zcur + off : nullptr;
    if (!is3) {
        const int p = bx * 256 + tid;
        const int pw = uniform(bx * 256 + wave * kWave);
        const int atom = uniform(pw < M.NP2 ? pw / M.npad : 0);
        if (p < M.NP2 && (p - atom * M.npad) < M.n_cells) {
            const int NP = M.NP2;
            const double S[3] = {cur[p], cur[NP + p], cur[2 * NP + p]};
            double H[3];
            field_su2<COH>(M, cur, atom, p, S, H);
            if (A.rec) {
                const cptr<double> R = as_const(M.frames2) + size_t(atom) * 9;
                const double sg = M.afm2[atom];
                const double B0 = M.site_uniform2 ? M.field2[atom * 3 + 0] : M.field2[size_t(p) * 3 + 0];
                const double B1 = M.site_uniform2 ? M.field2[atom * 3 + 1] : M.field2[size_t(p) * 3 + 1];
                const double B2 = M.site_uniform2 ? M.field2[atom * 3 + 2] : M.field2[size_t(p) * 3 + 2];
#pragma unroll
                for (int mu = 0; mu < 3; ++mu) {
                    const double g = R[mu * 3] * S[0] + R[mu * 3 + 1] * S[1] + R[mu * 3 + 2] * S[2];
                    part[mu] = sg * g;
                    part[3 + mu] = S[mu];
                    part[6 + mu] = g;
                }
                part[33] = 0.5 * (S[0] * (H[0] - B0) + S[1] * (H[1] - B1) + S[2] * (H[2] - B2));
            }
            if (A.tp) {
                const double f1 = fac[0], f2 = fac[1];
                if (f1 != 0.0 || f2 != 0.0) {
                    const cptr<double> d0 = as_const(M.fdrive2) + size_t(atom) * 3;
                    const cptr<double> d1 = as_const(M.fdrive2) + size_t(M.A2 + atom) * 3;
#pragma unroll
                    for (int a = 0; a < 3; ++a) H[a] -= d0[a] * f1 + d1[a] * f2;
                }
            }
            double k[3] = {H[1] * S[2] - H[2] * S[1], H[2] * S[0] - H[0] * S[2], H[0] * S[1] - H[1] * S[0]};
            if (M.alpha != 0.0) {  // mixed_lattice_md.cpp:570-578
                const double S2 = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
                const double SH = S[0] * H[0] + S[1] * H[1] + S[2] * H[2];
                const double g = (S2 > 0.0) ? M.alpha / sqrt(S2) : 0.0;
#pragma unroll
                for (int a = 0; a < 3; ++a) k[a] += g * (S[a] * SH - H[a] * S2);
            }
            if (TL) {
                const double dS[3] = {zc[p], zc[NP + p], zc[2 * NP + p]};
                double dH[3] = {0.0, 0.0, 0.0};
                onsite_lin<3>(M.onsite2, M.site_uniform2, M.has_onsite2, atom, p, dS, dH);
                gather_bonds<3, 3>(M.e22, zc, M.NP2, M.NP2, atom, p, dH);
                gather_bonds<3, 8>(M.e23, zc + size_t(3) * M.NP2, M.NP2, M.NP3, atom, p, dH);
                gather_bonds<3, 8>(M.e23, cur + size_t(3) * M.NP2, M.NP2, M.NP3, atom, p, dH);  // d H / d theta
                double dk[3] = {dH[1] * S[2] - dH[2] * S[1] + (H[1] * dS[2] - H[2] * dS[1]),
                                dH[2] * S[0] - dH[0] * S[2] + (H[2] * dS[0] - H[0] * dS[2]),
                                dH[0] * S[1] - dH[1] * S[0] + (H[0] * dS[1] - H[1] * dS[0])};
                if (M.alpha != 0.0) {
                    const double S2 = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
                    if (S2 > 0.0) {
                        const double SH = S[0] * H[0] + S[1] * H[1] + S[2] * H[2];
                        const double g = M.alpha / sqrt(S2);
                        const double SdS = S[0] * dS[0] + S[1] * dS[1] + S[2] * dS[2];
                        const double dSH = (dS[0] * H[0] + dS[1] * H[1] + dS[2] * H[2]) +
                                           (S[0] * dH[0] + S[1] * dH[1] + S[2] * dH[2]);
                        const double dg = -g * SdS / S2;
#pragma unroll
                        for (int a = 0; a < 3; ++a)
                            dk[a] += dg * (S[a] * SH - H[a] * S2) +
                                     g * (dS[a] * SH + S[a] * dSH - dH[a] * S2 - H[a] * (2.0 * SdS));
                    }
                }
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    rk_update(s, size_t(d) * NP + p, dS[d], dk[d], bdt, cdt, A.zx + off, A.zacc + off,
                              A.znext + off);
                if (A.rec) {
                    const cptr<double> R = as_const(M.frames2) + size_t(atom) * 9;
                    const double sg = M.afm2[atom];
#pragma unroll
                    for (int mu = 0; mu < 3; ++mu) {
                        const double g = R[mu * 3] * dS[0] + R[mu * 3 + 1] * dS[1] + R[mu * 3 + 2] * dS[2];
                        tpart[mu] = sg * g;
                        tpart[3 + mu] = dS[mu];
                        tpart[6 + mu] = g;
                    }
                }
            }
            if (s == 0) {
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const size_t q = size_t(d) * NP + p;
                    acc[q] = S[d] + bdt * k[d];
                    st_x<COH>(nxt + q, S[d] + cdt * k[d]);
                }
            } else if (s < 3) {
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const size_t q = size_t(d) * NP + p;
                    const double xd = x[q];
                    acc[q] = acc[q] + bdt * k[d];
                    st_x<COH>(nxt + q, xd + cdt * k[d]);
                }
            } else {
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const size_t q = size_t(d) * NP + p;
                    st_x<COH>(x + q, acc[q] + bdt * k[d]);
                }
            }
        }
    } else {
        const int p = (bx - A.nblk2) * 256 + tid;
        const int pw = uniform((bx - A.nblk2) * 256 + wave * kWave);
        const int atom = uniform(pw < M.NP3 ? pw / M.npad : 0);
        const int cell = p - atom * M.npad;
        if (p < M.NP3 && cell < M.n_cells) {
            const int NP = M.NP3;
            const size_t o3 = size_t(3) * M.NP2;
            double S[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) S[a] = cur[o3 + size_t(a) * NP + p];
            double H[8];
            field_su3<COH>(M, cur, atom, p, S, H);
            if (A.rec) {
                const cptr<double> R = as_const(M.frames3) + size_t(atom) * 64;
                const double sg = ((cell * M.A3 + atom) & 1) ? -1.0 : 1.0;
                double e = 0.0;
#pragma unroll
                for (int mu = 0; mu < 8; ++mu) {
                    double g = 0.0;
#pragma unroll
                    for (int nu = 0; nu < 8; ++nu) g += R[mu * 8 + nu] * S[nu];
                    part[9 + mu] = sg * S[mu];
                    part[17 + mu] = S[mu];
                    part[25 + mu] = g;
                    const double Bm = M.site_uniform3 ? M.field3[atom * 8 + mu] : M.field3[size_t(p) * 8 + mu];
                    e += S[mu] * (H[mu] - Bm);
                }
                part[33] = 0.5 * e;
            }
            if (A.tp) {
                const double f1 = fac[2], f2 = fac[3];
                if (f1 != 0.0 || f2 != 0.0) {
                    const cptr<double> d0 = as_const(M.fdrive3) + size_t(atom) * 8;
                    const cptr<double> d1 = as_const(M.fdrive3) + size_t(M.A3 + atom) * 8;
#pragma unroll
                    for (int a = 0; a < 8; ++a) H[a] -= d0[a] * f1 + d1[a] * f2;
                }
            }
            double k[8];
            su3_cross(H, S, k);
            if (M.has_bloch) {  // mixed_lattice_md.cpp:596-605
#pragma unroll
                for (int a = 0; a < 8; ++a)
                    if (M.gamma[a] != 0.0) k[a] -= M.gamma[a] * (S[a] - M.eq3[size_t(a) * NP + p]);
            }
            if (TL) {
                double dl[8];
#pragma unroll
                for (int a = 0; a < 8; ++a) dl[a] = zc[o3 + size_t(a) * NP + p];
                double dH[8];
#pragma unroll
                for (int a = 0; a < 8; ++a) dH[a] = 0.0;
                onsite_lin<8>(M.onsite3, M.site_uniform3, M.has_onsite3, atom, p, dl, dH);
                gather_bonds<8, 8>(M.e33, zc + o3, M.NP3, M.NP3, atom, p, dH);
                gather_bonds<8, 3>(M.e32, zc, M.NP3, M.NP2, atom, p, dH);
                gather_bonds<8, 3>(M.e32, cur, M.NP3, M.NP2, atom, p, dH);  // d H / d theta
                double dk[8], t2[8];
                su3_cross(dH, S, dk);
                su3_cross(H, dl, t2);
#pragma unroll
                for (int a = 0; a < 8; ++a) dk[a] += t2[a];
                if (M.has_bloch) {
#pragma unroll
                    for (int a = 0; a < 8; ++a) dk[a] -= M.gamma[a] * dl[a];
                }
#pragma unroll
                for (int d = 0; d < 8; ++d)
                    rk_update(s, o3 + size_t(d) * NP + p, dl[d], dk[d], bdt, cdt, A.zx + off, A.zacc + off,
                              A.znext + off);
                if (A.rec) {
                    const cptr<double> R = as_const(M.frames3) + size_t(atom) * 64;
                    const double sg = ((cell * M.A3 + atom) & 1) ? -1.0 : 1.0;
#pragma unroll
                    for (int mu = 0; mu < 8; ++mu) {
                        double g = 0.0;
#pragma unroll
                        for (int nu = 0; nu < 8; ++nu) g += R[mu * 8 + nu] * dl[nu];
                        tpart[9 + mu] = sg * dl[mu];
                        tpart[17 + mu] = dl[mu];
                        tpart[25 + mu] = g;
                    }
                }
            }
            if (s == 0) {
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    const size_t q = o3 + size_t(d) * NP + p;
                    acc[q] = S[d] + bdt * k[d];
                    st_x<COH>(nxt + q, S[d] + cdt * k[d]);
                }
            } else if (s < 3) {
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    const size_t q = o3 + size_t(d) * NP + p;
                    const double xd = x[q];
                    acc[q] = acc[q] + bdt * k[d];
                    st_x<COH>(nxt + q, xd + cdt * k[d]);
                }
            } else {
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    const size_t q = o3 + size_t(d) * NP + p;
                    st_x<COH>(x + q, acc[q] + bdt * k[d]);
                }
            }
        }
    }
    if (A.rec) {
        // SU(2) workgroups own record entries 0..8 and 33, SU(3) ones 9..33
        const int j0 = is3 ? 9 : 0, j1 = is3 ? 33 : 9;
#pragma unroll
        for (int j = 0; j < kRec; ++j) {
            if ((j >= j0 && j < j1) || j == 33) {
                const double v = wave_sum(part[j]);
                if (lane == 0) red[wave][j] = v;
            }
        }
        __syncthreads();
        if (tid < kRec) {
            const bool mine = (tid >= j0 && tid < j1) || tid == 33;
            const double v = mine ? (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]) : 0.0;
            st_x<COH>(A.partials + (size_t(b) * gridDim.x + bx) * kRec + tid, v);
        }
        if (TL) {
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kRec; ++j) {
                if (j >= j0 && j < j1) {
                    const double v = wave_sum(tpart[j]);
                    if (lane == 0) red[wave][j] = v;
                }
            }
            __syncthreads();
            if (tid < kRec) {
                const bool mine = tid >= j0 && tid < j1;
                const double v = mine ? (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]) : 0.0;
                A.tpartials[(size_t(b) * gridDim.x + bx) * kRec + tid] = v;
            }
        }
    }
}


template <bool TL>
__global__ void __launch_bounds__(256) k_mixed_stage(MixDev M, MixStageArgs A) {
    __shared__ double fac[4];
    __shared__ double red[4][kRec];
    int bx, b;
    mix_block(bx, b);
    mixed_stage_wg<TL, false>(M, A, bx, b, fac, red);
}

// Cooperative form of the stage kernels for a small batch of a large
// lattice (csl_mixed_set_persistent mode -2): the nblk x B workgroups of one stage
// launch, all co-resident, run every stage of every step in ONE launch with
// a grid-wide barrier between stages instead of one launch per stage.  The
// values other workgroups gather are written and read at agent scope; a
// barrier is one agent-scope counter (every wave's stores complete, then one
// increment per workgroup, then a bounded wait for the stage's count).  The
// first workgroup of a trajectory reduces its record partials after the
// barrier that follows a recorded stage 0.
struct MixCoopArgs {
    MixStageArgs A;        // x, acc, y1 (A.ynext), y2 (A.ycur), pulses, partials, dt, nblk2
    double* aux;           // scratch outputs of the final record's evaluation
    const double* tgrid;   // [n_steps + 1]
    int n_steps, save_interval, n_rec;
    int rec_stride;        // row stride of recs (trajectories)
    double* recs;          // [n_rec][rec_stride][34], this batch's first trajectory at column 0
    double inv2, inv3, invE;
    unsigned* ctr;         // barrier counter, zero at launch
    unsigned* status;      // set when a barrier wait gives up
};

constexpr int kCoopSpin = 1 << 22;

__global__ void __launch_bounds__(256) k_mixed_coop(MixDev M, MixCoopArgs C) {
    __shared__ double fac[4];
    __shared__ double red[4][kRec];
    __shared__ int ok;
    int bx, b;
    mix_block(bx, b);
    const int tid = threadIdx.x;
    const int nblk = int(gridDim.x);
    const unsigned nwg = gridDim.x * gridDim.y;
    unsigned epoch = 0;
    double* const y1 = C.A.ynext;
    double* const y2 = const_cast<double*>(C.A.ycur);
    auto barrier = [&]() -> bool {
        __builtin_amdgcn_s_waitcnt(0);  // this thread's agent-scope stores are performed
        __syncthreads();
        ++epoch;
        if (tid == 0) {
            __hip_atomic_fetch_add(C.ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned want = epoch * nwg;
            int good = 0;
            for (int it = 0; it < kCoopSpin; ++it) {
                if (__hip_atomic_load(C.ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= want) {
                    good = 1;
                    break;
                }
                if ((it & 255) == 255 && __hip_atomic_load(C.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    break;
                __builtin_amdgcn_s_sleep(1);
            }
            if (!good) __hip_atomic_store(C.status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = good;
        }
        __syncthreads();
        return ok != 0;
    };
    auto reduce = [&](int ri) {
        if (bx == 0 && tid < kRec) {
            double tot = 0.0;
            for (int k = 0; k < nblk; ++k) tot += ld_x<true>(C.A.partials + (size_t(b) * nblk + k) * kRec + tid);
            const double sc = tid < 9 ? C.inv2 : (tid < 33 ? C.inv3 : C.invE);
            C.recs[(size_t(ri) * C.rec_stride + b) * kRec + tid] = tot * sc;
        }
    };
    const double dt = C.A.dt;
    for (int n = 0; n < C.n_steps; ++n) {
        const bool rec = (n % C.save_interval) == 0;
        const double t = C.tgrid[n];
        // stage times t, t + dt/2, t + dt/2, t + dt; inputs x, y1, y2, y1
        for (int st = 0; st < 4; ++st) {
            MixStageArgs A = C.A;
            A.stage = st;
            A.ycur = st == 0 ? A.x : (st == 2 ? y2 : y1);
            A.ynext = st == 0 ? y1 : (st == 1 ? y2 : (st == 2 ? y1 : y2));
            A.t = t + (st == 0 ? 0.0 : (st == 3 ? 1.0 * dt : 0.5 * dt));
            A.rec = (st == 0 && rec) ? 1 : 0;
            mixed_stage_wg<false, true>(M, A, bx, b, fac, red);
            if (!barrier()) return;
            if (A.rec) reduce(n / C.save_interval);
        }
    }
    // final record: a stage-0 evaluation of the final state into scratch
    MixStageArgs A = C.A;
    A.stage = 0;
    A.acc = C.aux;
    A.ycur = A.x;
    A.ynext = C.aux;
    A.t = C.tgrid[C.n_steps];
    A.rec = 1;
    mixed_stage_wg<false, true>(M, A, bx, b, fac, red);
    if (!barrier()) return;
    reduce(C.n_rec - 1);
}

// Loss sum_t sum_j w_j (rec_j(t) - target_j(t))^2 of each trajectory and its
// derivative 2 sum_t sum_j w_j (rec_j - target_j) trec_j from the records and
// the tangent records ([n_rec][B][34]); target [n_rec][34] shared or
// [n_rec][B][34].  One workgroup per trajectory.
struct Weights34 {
    double w[kRec];
};
__global__ void __launch_bounds__(256) k_mixed_loss_grad(const double* __restrict__ recs,
                                                         const double* __restrict__ trecs,
                                                         const double* __restrict__ target, int per_replica,
                                                         int n_rec, int B, Weights34 W, double* __restrict__ loss,
                                                         double* __restrict__ grad) {
    __shared__ double red[2][4];
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    double l = 0.0, g = 0.0;
    for (long i = tid; i < long(n_rec) * kRec; i += 256) {
        const int r = int(i / kRec), j = int(i % kRec);
        const double w = W.w[j];
        if (w == 0.0) continue;
        const size_t q = (size_t(r) * B + b) * kRec + j;
        const double tg = per_replica ? target[q] : target[size_t(r) * kRec + j];
        const double d = recs[q] - tg;
        l += w * d * d;
        g += 2.0 * w * d * trecs[q];
    }
    l = wave_sum(l);
    g = wave_sum(g);
    if ((tid & (kWave - 1)) == 0) {
        red[0][tid >> 6] = l;
        red[1][tid >> 6] = g;
    }
    __syncthreads();
    if (tid == 0) {
        loss[b] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        grad[b] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}

// One workgroup per trajectory: partials -> record ri
__global__ void __launch_bounds__(256) k_mixed_reduce(const double* __restrict__ partials, int nblk, int B,
                                                      int ri, double inv2, double inv3, double invE,
                                                      double* __restrict__ recs, int rec_stride) {
    __shared__ double red[kRec][4];
    const int b = blockIdx.x;
    const int t = threadIdx.x;
    double v[kRec];
#pragma unroll
    for (int j = 0; j < kRec; ++j) v[j] = 0.0;
    for (int k = t; k < nblk; k += 256) {
        const double* q = partials + (size_t(b) * nblk + k) * kRec;
#pragma unroll
        for (int j = 0; j < kRec; ++j) v[j] += q[j];
    }
#pragma unroll
    for (int j = 0; j < kRec; ++j) v[j] = wave_sum(v[j]);
    if ((t & (kWave - 1)) == 0)
#pragma unroll
        for (int j = 0; j < kRec; ++j) red[j][t >> 6] = v[j];
    __syncthreads();
    if (t < kRec) {
        const double tot = (red[t][0] + red[t][1]) + (red[t][2] + red[t][3]);
        const double sc = t < 9 ? inv2 : (t < 33 ? inv3 : invE);
        recs[(size_t(ri) * rec_stride + b) * kRec + t] = tot * sc;
    }
}

// Local fields (no drive) and dS/dt at time t with drive factors f[4] of one state
__global__ void __launch_bounds__(256) k_mixed_eval(MixDev M, const double* __restrict__ y, double f21,
                                                    double f22, double f31, double f32,
                                                    double* __restrict__ dsdt) {
    const int nblk2 = (M.NP2 + 255) / 256;
    const int bx = blockIdx.x;
    const int tid = threadIdx.x;
    const int wave = uniform(tid >> 6);
    if (bx < nblk2) {
        const int p = bx * 256 + tid;
        const int pw = uniform(bx * 256 + wave * kWave);
        const int atom = uniform(pw < M.NP2 ? pw / M.npad : 0);
        if (p >= M.NP2 || p - atom * M.npad >= M.n_cells) return;
        const int NP = M.NP2;
        const double S[3] = {y[p], y[NP + p], y[2 * NP + p]};
        double H[3];
        field_su2(M, y, atom, p, S, H);
        if (f21 != 0.0 || f22 != 0.0)
            for (int a = 0; a < 3; ++a)
                H[a] -= M.fdrive2[atom * 3 + a] * f21 + M.fdrive2[(M.A2 + atom) * 3 + a] * f22;
        double k[3] = {H[1] * S[2] - H[2] * S[1], H[2] * S[0] - H[0] * S[2], H[0] * S[1] - H[1] * S[0]};
        if (M.alpha != 0.0) {
            const double S2 = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
            const double SH = S[0] * H[0] + S[1] * H[1] + S[2] * H[2];
            const double g = (S2 > 0.0) ? M.alpha / sqrt(S2) : 0.0;
            for (int a = 0; a < 3; ++a) k[a] += g * (S[a] * SH - H[a] * S2);
        }
        for (int d = 0; d < 3; ++d) dsdt[size_t(d) * NP + p] = k[d];
    } else {
        const int p = (bx - nblk2) * 256 + tid;
        const int pw = uniform((bx - nblk2) * 256 + wave * kWave);
        const int atom = uniform(pw < M.NP3 ? pw / M.npad : 0);
        if (p >= M.NP3 || p - atom * M.npad >= M.n_cells) return;
        const int NP = M.NP3;
        const size_t o3 = size_t(3) * M.NP2;
        double S[8], H[8], k[8];
        for (int a = 0; a < 8; ++a) S[a] = y[o3 + size_t(a) * NP + p];
        field_su3(M, y, atom, p, S, H);
        if (f31 != 0.0 || f32 != 0.0)
            for (int a = 0; a < 8; ++a)
                H[a] -= M.fdrive3[atom * 8 + a] * f31 + M.fdrive3[(M.A3 + atom) * 8 + a] * f32;
        su3_cross(H, S, k);
        if (M.has_bloch)
            for (int a = 0; a < 8; ++a)
                if (M.gamma[a] != 0.0) k[a] -= M.gamma[a] * (S[a] - M.eq3[size_t(a) * NP + p]);
        for (int d = 0; d < 8; ++d) dsdt[o3 + size_t(d) * NP + p] = k[d];
    }
}

// flat reference state(s) [N2*3 | N3*8] <-> device SoA, B trajectories
__global__ void k_mixed_perm(MixDev M, const double* __restrict__ src, size_t src_stride, double* __restrict__ dst,
                             size_t dst_stride, int to_soa) {
    const int b = blockIdx.y;
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;   // flat component index
    const long n2 = long(M.N2) * 3, n = n2 + long(M.N3) * 8;
    if (i >= n) return;
    size_t q;
    if (i < n2) {
        const int site = int(i / 3), d = int(i % 3);
        const int p = (site % M.A2) * M.npad + site / M.A2;
        q = size_t(d) * M.NP2 + p;
    } else {
        const long j = i - n2;
        const int site = int(j / 8), d = int(j % 8);
        const int p = (site % M.A3) * M.npad + site / M.A3;
        q = size_t(3) * M.NP2 + size_t(d) * M.NP3 + p;
    }
    if (to_soa) dst[size_t(b) * dst_stride + q] = src[size_t(b) * src_stride + size_t(i)];
    else dst[size_t(b) * dst_stride + size_t(i)] = src[size_t(b) * src_stride + q];
}

__global__ void k_mixed_normalize(MixDev M, double* __restrict__ y, double len2, double len3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M.NP2) {
        const double a = y[i], b = y[M.NP2 + i], c = y[2 * M.NP2 + i];
        const double n = sqrt(a * a + b * b + c * c);
        if (n > 0.0) {
            const double s = len2 / n;
            y[i] = a * s;
            y[M.NP2 + i] = b * s;
            y[2 * M.NP2 + i] = c * s;
        }
    }
    if (i < M.NP3) {
        double* z = y + size_t(3) * M.NP2;
        double n2 = 0.0;
        for (int a = 0; a < 8; ++a) n2 += z[size_t(a) * M.NP3 + i] * z[size_t(a) * M.NP3 + i];
        if (n2 > 0.0) {
            const double s = len3 / sqrt(n2);
            for (int a = 0; a < 8; ++a) z[size_t(a) * M.NP3 + i] *= s;
        }
    }
}

// ---------------------------------------------------------------------------
// Dormand-Prince 5(4) on the mixed lattice (MixedLattice::integrate_ode_system
// with "dopri5", mixed_lattice.h:3960-4010).  The state exceeds one
// workgroup's LDS, so every attempt is a sequence of grid-wide launches over
// all trajectories: stages 2..6 (k_s of the stage input, then the next stage
// input of the owned site), stage 7 (k7 of the candidate, the per-workgroup
// error maximum and record partials) and one decision workgroup per
// trajectory (k_mixdp_decide), which accepts or rejects, publishes the
// observer records and sets up the trajectory's next attempt or derivative
// evaluation.  Each trajectory keeps its own step-size sequence; its x and
// k1 live in one of two buffers (par), so accepting a step is a flip.

// Dormand-Prince tableau (Dormand & Prince 1980)
namespace dp {
constexpr double c2 = 1.0 / 5.0, c3 = 3.0 / 10.0, c4 = 4.0 / 5.0, c5 = 8.0 / 9.0;
constexpr double a21 = 1.0 / 5.0;
constexpr double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
constexpr double a41 = 44.0 / 45.0, a42 = -56.0 / 15.0, a43 = 32.0 / 9.0;
constexpr double a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0, a54 = -212.0 / 729.0;
constexpr double a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0, a64 = 49.0 / 176.0,
                 a65 = -5103.0 / 18656.0;
constexpr double b1 = 35.0 / 384.0, b3 = 500.0 / 1113.0, b4 = 125.0 / 192.0, b5 = -2187.0 / 6784.0,
                 b6 = 11.0 / 84.0;
constexpr double e1 = 35.0 / 384.0 - 5179.0 / 57600.0, e3 = 500.0 / 1113.0 - 7571.0 / 16695.0,
                 e4 = 125.0 / 192.0 - 393.0 / 640.0, e5 = -2187.0 / 6784.0 + 92097.0 / 339200.0,
                 e6 = 11.0 / 84.0 - 187.0 / 2100.0, e7 = -1.0 / 40.0;
constexpr double eps = 2.220446049250313e-16;  // numeric_limits<double>::epsilon()
}  // namespace dp

__device__ __forceinline__ double wave_max_f64(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ bool dp_less(double t1, double t2) { return t2 - t1 > dp::eps; }
__device__ __forceinline__ bool dp_less_eq(double t1, double t2) { return t1 - t2 <= dp::eps; }
__device__ __forceinline__ double dp_max(double a, double b) { return a < b ? b : a; }
__device__ __forceinline__ double dp_grid_time(double start, long long k, double dt) {
#pragma clang fp contract(off)
    return start + double(k) * dt;
}

// Per-trajectory controller state, device resident.
struct MixDpCtl {
    int mode;        // 0 finished, 1 attempt, 2 derivative evaluation k1 = f(x, te)
    int par;         // x and k1 are X[par], K1[par]; the candidate goes to X[1 - par]
    int obs;         // this iteration's state is observed (attempt: if accepted)
    int eobs;        // the evaluation is an observer call
    int ends;        // an accepted attempt ends the run / interval
    int fresh;       // control 2: at the start of interval kk
    int refresh;     // the attempt first takes k1 = f(x, te) again (stage 1)
    int fails, status, sg, gi;
    long long kk, obs_calls, saved, n_acc, n_rej, attempts;
    double t, h, t7, tobs, te, tend, s0, s1, tcur;
};

struct MixDpArgs {
    int stage;                    // 1..7 (1: k1 refresh before an attempt)
    int nblk2;
    double *X0, *X1, *K10, *K11;  // x and k1, two buffers each
    double* K[5];                 // k2..k6
    double *Y0, *Y1;              // stage inputs y3..y6
    const double* tp;             // [B][2] pulse centres (both species)
    const double* scale;          // [B][2]
    double abs_tol, rel_tol;
    int control;
    const MixDpCtl* ctl;
    double* errp;                 // [B][nblk]
    double* partials;             // [B][nblk][34]
};

// Stage input of trajectory b: a[i], or a[i] + c b[i] (y2 = x + h a21 k1
// formed on the fly from x and k1 of the owned site and of the partners).
struct DpSrc {
    const double* a;
    const double* k;
    double c;
    __device__ __forceinline__ double operator()(size_t i) const { return k ? a[i] + c * k[i] : a[i]; }
};

template <int DA, int DB>
__device__ __forceinline__ void gather_src(const Ell& E, const DpSrc& src, size_t base, int NPsrc, int NPpart,
                                           int atom, int p, double* H) {
    const int c = E.uniform ? E.cnt_atom[atom] : E.cnt[p];
    for (int k = 0; k < c; ++k) {
        const int q = E.col[size_t(k) * NPsrc + p];
        double s[DB];
#pragma unroll
        for (int e = 0; e < DB; ++e) s[e] = src(base + size_t(e) * NPpart + q);
        const double* J = E.uniform ? E.mats + (size_t(atom) * E.zmax + k) * (DA * DB)
                                    : E.mats + (size_t(k) * NPsrc + p) * (DA * DB);
#pragma unroll
        for (int a = 0; a < DA; ++a)
#pragma unroll
            for (int e = 0; e < DB; ++e) H[a] = fma(J[a * DB + e], s[e], H[a]);
    }
}

__global__ void __launch_bounds__(256) k_mixdp_stage(MixDev M, MixDpArgs A) {
    __shared__ double fac[4];
    __shared__ double red[4][kRec];
    __shared__ double emx[4];
    int bx, b;
    mix_block(bx, b);
    const MixDpCtl& C = A.ctl[b];
    const int mode = C.mode;
    const int S = A.stage;
    // uniform per workgroup: finished trajectories, evaluations outside
    // stage 7, stage 1 without a refresh
    if (mode == 0 || (mode == 2 && S != 7) || (S == 1 && (mode != 1 || C.refresh == 0))) return;
    const bool eval = mode == 2 || S == 1;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = uniform(tid >> 6);
    const bool is3 = bx >= A.nblk2;
    const int par = C.par;
    const double h = C.h;
    const bool rec = C.obs != 0 && S == 7;
    const size_t off = size_t(b) * mix_stride(M);
    const double* X = (par ? A.X1 : A.X0) + off;
    double* XN = (par ? A.X0 : A.X1) + off;
    double* K1 = (par ? A.K11 : A.K10) + off;
    double* K7 = (par ? A.K10 : A.K11) + off;
    double ts;
    switch (S) {
        case 1: ts = C.te; break;
        case 2: ts = C.t + dp::c2 * h; break;
        case 3: ts = C.t + dp::c3 * h; break;
        case 4: ts = C.t + dp::c4 * h; break;
        case 5: ts = C.t + dp::c5 * h; break;
        case 6: ts = C.t + h; break;
        default: ts = eval ? C.te : C.t7; break;
    }
    if (tid < 4) {  // drive envelopes of this trajectory at the stage time
        const int sp = tid >> 1, k = tid & 1;
        const double amp = sp ? M.amp3 : M.amp2;
        double f = 0.0;
        if (amp != 0.0) {
            f = envelope(amp, ts, A.tp[size_t(b) * 2 + k], sp ? M.width3 : M.width2, sp ? M.freq3 : M.freq2);
            f *= A.scale[size_t(b) * 2 + k];
        }
        fac[tid] = f;
    }
    __syncthreads();
    DpSrc src;
    src.k = nullptr;
    src.c = 0.0;
    if (S == 2) {
        src.a = X;
        src.k = K1;
        src.c = h * dp::a21;
    } else if (S >= 3 && S <= 6) {
        src.a = ((S & 1) ? A.Y1 : A.Y0) + off;
    } else {
        src.a = eval ? X : XN;
    }
    double* ynext = S <= 5 ? (((S + 1) & 1) ? A.Y1 : A.Y0) + off : XN;
    const double* Kb[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) Kb[j] = A.K[j] + off;
    double part[kRec];
#pragma unroll
    for (int j = 0; j < kRec; ++j) part[j] = 0.0;
    double emax = 0.0;
    const double ah = fabs(h);
    // one derivative component kd of plane index q: store it and form the
    // next stage input / candidate, or (stage 7) the error ratio
    auto finish = [&](size_t q, double kd) {
        if (S == 1 || S == 7) {
            if (eval) {
                K1[q] = kd;
            } else {
                K7[q] = kd;
                const double k1 = K1[q];
                const double xerr = h * (dp::e1 * k1 + dp::e3 * Kb[1][q] + dp::e4 * Kb[2][q] + dp::e5 * Kb[3][q] +
                                         dp::e6 * Kb[4][q] + dp::e7 * kd);
                const double r = fabs(xerr) / (A.abs_tol + A.rel_tol * (fabs(X[q]) + ah * fabs(k1)));
                emax = fmax(emax, r);
            }
            return;
        }
        const double x = X[q], k1 = K1[q];
        double y;
        if (S == 2) {
            y = x + h * (dp::a31 * k1 + dp::a32 * kd);
        } else if (S == 3) {
            y = x + h * (dp::a41 * k1 + dp::a42 * Kb[0][q] + dp::a43 * kd);
        } else if (S == 4) {
            y = x + h * (dp::a51 * k1 + dp::a52 * Kb[0][q] + dp::a53 * Kb[1][q] + dp::a54 * kd);
        } else if (S == 5) {
            y = x + h * (dp::a61 * k1 + dp::a62 * Kb[0][q] + dp::a63 * Kb[1][q] + dp::a64 * Kb[2][q] + dp::a65 * kd);
        } else {
            y = x + h * (dp::b1 * k1 + dp::b3 * Kb[1][q] + dp::b4 * Kb[2][q] + dp::b5 * Kb[3][q] + dp::b6 * kd);
        }
        A.K[S - 2][off + q] = kd;
        ynext[q] = y;
    };
    const size_t o3 = size_t(3) * M.NP2;
    if (!is3) {
        const int p = bx * 256 + tid;
        const int pw = uniform(bx * 256 + wave * kWave);
        const int atom = uniform(pw < M.NP2 ? pw / M.npad : 0);
        if (p < M.NP2 && (p - atom * M.npad) < M.n_cells) {
            const int NP = M.NP2;
            const double S3[3] = {src(p), src(size_t(NP) + p), src(size_t(2) * NP + p)};
            double H[3];
            field_onsite<3>(M.field2, M.onsite2, M.site_uniform2, M.has_onsite2, atom, p, S3, H);
            gather_src<3, 3>(M.e22, src, 0, M.NP2, M.NP2, atom, p, H);
            gather_src<3, 8>(M.e23, src, o3, M.NP2, M.NP3, atom, p, H);
            if (rec) {
                const cptr<double> R = as_const(M.frames2) + size_t(atom) * 9;
                const double sg = M.afm2[atom];
                const double B0 = M.site_uniform2 ? M.field2[atom * 3 + 0] : M.field2[size_t(p) * 3 + 0];
                const double B1 = M.site_uniform2 ? M.field2[atom * 3 + 1] : M.field2[size_t(p) * 3 + 1];
                const double B2 = M.site_uniform2 ? M.field2[atom * 3 + 2] : M.field2[size_t(p) * 3 + 2];
#pragma unroll
                for (int mu = 0; mu < 3; ++mu) {
                    const double g = R[mu * 3] * S3[0] + R[mu * 3 + 1] * S3[1] + R[mu * 3 + 2] * S3[2];
                    part[mu] = sg * g;
                    part[3 + mu] = S3[mu];
                    part[6 + mu] = g;
                }
                part[33] = 0.5 * (S3[0] * (H[0] - B0) + S3[1] * (H[1] - B1) + S3[2] * (H[2] - B2));
            }
            const double f1 = fac[0], f2 = fac[1];
            if (f1 != 0.0 || f2 != 0.0) {
                const cptr<double> d0 = as_const(M.fdrive2) + size_t(atom) * 3;
                const cptr<double> d1 = as_const(M.fdrive2) + size_t(M.A2 + atom) * 3;
#pragma unroll
                for (int a = 0; a < 3; ++a) H[a] -= d0[a] * f1 + d1[a] * f2;
            }
            double k[3] = {H[1] * S3[2] - H[2] * S3[1], H[2] * S3[0] - H[0] * S3[2], H[0] * S3[1] - H[1] * S3[0]};
            if (M.alpha != 0.0) {  // mixed_lattice_md.cpp:570-578
                const double S2 = S3[0] * S3[0] + S3[1] * S3[1] + S3[2] * S3[2];
                const double SH = S3[0] * H[0] + S3[1] * H[1] + S3[2] * H[2];
                const double g = (S2 > 0.0) ? M.alpha / sqrt(S2) : 0.0;
#pragma unroll
                for (int a = 0; a < 3; ++a) k[a] += g * (S3[a] * SH - H[a] * S2);
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) finish(size_t(d) * NP + p, k[d]);
        }
    } else {
        const int p = (bx - A.nblk2) * 256 + tid;
        const int pw = uniform((bx - A.nblk2) * 256 + wave * kWave);
        const int atom = uniform(pw < M.NP3 ? pw / M.npad : 0);
        const int cell = p - atom * M.npad;
        if (p < M.NP3 && cell < M.n_cells) {
            const int NP = M.NP3;
            double S8[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) S8[a] = src(o3 + size_t(a) * NP + p);
            double H[8];
            field_onsite<8>(M.field3, M.onsite3, M.site_uniform3, M.has_onsite3, atom, p, S8, H);
            gather_src<8, 8>(M.e33, src, o3, M.NP3, M.NP3, atom, p, H);
            gather_src<8, 3>(M.e32, src, 0, M.NP3, M.NP2, atom, p, H);
            if (rec) {
                const cptr<double> R = as_const(M.frames3) + size_t(atom) * 64;
                const double sg = ((cell * M.A3 + atom) & 1) ? -1.0 : 1.0;
                double e = 0.0;
#pragma unroll
                for (int mu = 0; mu < 8; ++mu) {
                    double g = 0.0;
#pragma unroll
                    for (int nu = 0; nu < 8; ++nu) g += R[mu * 8 + nu] * S8[nu];
                    part[9 + mu] = sg * S8[mu];
                    part[17 + mu] = S8[mu];
                    part[25 + mu] = g;
                    const double Bm = M.site_uniform3 ? M.field3[atom * 8 + mu] : M.field3[size_t(p) * 8 + mu];
                    e += S8[mu] * (H[mu] - Bm);
                }
                part[33] = 0.5 * e;
            }
            const double f1 = fac[2], f2 = fac[3];
            if (f1 != 0.0 || f2 != 0.0) {
                const cptr<double> d0 = as_const(M.fdrive3) + size_t(atom) * 8;
                const cptr<double> d1 = as_const(M.fdrive3) + size_t(M.A3 + atom) * 8;
#pragma unroll
                for (int a = 0; a < 8; ++a) H[a] -= d0[a] * f1 + d1[a] * f2;
            }
            double k[8];
            su3_cross(H, S8, k);
            if (M.has_bloch) {  // mixed_lattice_md.cpp:596-605
#pragma unroll
                for (int a = 0; a < 8; ++a)
                    if (M.gamma[a] != 0.0) k[a] -= M.gamma[a] * (S8[a] - M.eq3[size_t(a) * NP + p]);
            }
#pragma unroll
            for (int d = 0; d < 8; ++d) finish(o3 + size_t(d) * NP + p, k[d]);
        }
    }
    if (S != 7) return;  // uniform
    const size_t nblk = gridDim.x;
    if (!eval) {
        const double e = wave_max_f64(emax);
        if (lane == 0) emx[wave] = e;
    }
    if (rec) {
        const int j0 = is3 ? 9 : 0, j1 = is3 ? 33 : 9;
#pragma unroll
        for (int j = 0; j < kRec; ++j) {
            if ((j >= j0 && j < j1) || j == 33) {
                const double v = wave_sum(part[j]);
                if (lane == 0) red[wave][j] = v;
            }
        }
    }
    __syncthreads();
    if (!eval && tid == 0) A.errp[size_t(b) * nblk + bx] = fmax(fmax(emx[0], emx[1]), fmax(emx[2], emx[3]));
    if (rec && tid < kRec) {
        const int j0 = is3 ? 9 : 0, j1 = is3 ? 33 : 9;
        const bool mine = (tid >= j0 && tid < j1) || tid == 33;
        const double v = mine ? (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]) : 0.0;
        A.partials[(size_t(b) * nblk + bx) * kRec + tid] = v;
    }
}

struct MixDpDecideArgs {
    int nblk, control, save_interval, n_grid, max_seg, max_rec;
    long long max_attempts;
    double T0, T1, dt, inv2, inv3, invE;
    const double* tgrid;      // control 0
    const double* segs;       // control 2: [B][max_seg][2] or null (one segment [T0, T1])
    const int* n_segs;
    const double* tp;         // [B][2]
    const double* scale;      // [B][2]
    double amp2, amp3;
    MixDpCtl* ctl;
    const double* errp;
    const double* partials;
    double* recs;             // [B][max_rec][34]
    double* times;            // [B][max_rec]
};

// Decision of one trajectory after an iteration, then the set-up of its next
// attempt or evaluation (the loop of k_dopri5 split at the launches).
__global__ void __launch_bounds__(256) k_mixdp_decide(MixDpDecideArgs D) {
    __shared__ double red[kRec][4];
    __shared__ double emx[4];
    __shared__ long long pub_ri;
    __shared__ double pub_t;
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid >> 6;
    MixDpCtl* cp = D.ctl + b;
    const int mode = cp->mode;
    if (mode == 0) return;  // uniform
    const bool attempt = mode == 1;
    const bool ob = cp->obs != 0;
    // workgroup maximum of the error ratios, sums of the record partials
    double e = 0.0;
    if (attempt && D.control != 0)
        for (int k = tid; k < D.nblk; k += 256) e = fmax(e, D.errp[size_t(b) * D.nblk + k]);
    e = wave_max_f64(e);
    if (lane == 0) emx[wave] = e;
    if (ob) {
        double v[kRec];
#pragma unroll
        for (int j = 0; j < kRec; ++j) v[j] = 0.0;
        for (int k = tid; k < D.nblk; k += 256) {
            const double* q = D.partials + (size_t(b) * D.nblk + k) * kRec;
#pragma unroll
            for (int j = 0; j < kRec; ++j) v[j] += q[j];
        }
#pragma unroll
        for (int j = 0; j < kRec; ++j) {
            v[j] = wave_sum(v[j]);
            if (lane == 0) red[j][wave] = v[j];
        }
    }
    __syncthreads();
    if (tid == 0) {
        MixDpCtl c = *cp;
        const double err = fmax(fmax(emx[0], emx[1]), fmax(emx[2], emx[3]));
        long long ri = -1;
        double rt = 0.0;
        bool done = false;
        const double dt = D.dt;
        const double* sg = D.segs ? D.segs + size_t(b) * D.max_seg * 2 : nullptr;
        const int nseg = D.control == 2 ? (D.segs ? D.n_segs[b] : 1) : 0;
        const bool drive_on = (D.amp2 != 0.0 || D.amp3 != 0.0) &&
                              (D.scale[size_t(b) * 2] != 0.0 || D.scale[size_t(b) * 2 + 1] != 0.0);
        // the iteration that just ran
        if (!attempt) {
            if (c.eobs) {
                if (ob) {
                    ri = c.saved++;
                    rt = c.te;
                }
                ++c.obs_calls;
            }
        } else if (D.control != 0 && !(err <= 1.0)) {
            c.h = c.h * dp_max(0.9 * pow(err, -1.0 / 3.0), 0.2);
            ++c.n_rej;
            if (++c.fails >= 500) {
                c.status = 1;
                done = true;
            }
        } else {
            c.fails = 0;
            c.par ^= 1;
            if (D.control != 0) {
                c.t = c.t + c.h;
                if (err < 0.5) c.h = c.h * (0.9 * pow(dp_max(1.0 / 3125.0, err), -1.0 / 5.0));
            }
            ++c.n_acc;
            if (c.ends) {
                if (c.obs) {
                    ri = c.saved++;
                    rt = c.tobs;
                }
                ++c.obs_calls;
                if (D.control == 0) ++c.gi;
                if (D.control == 2) {
                    ++c.kk;
                    c.tcur = c.tobs;
                    c.fresh = 1;
                }
            }
        }
        // the next attempt or evaluation
        c.mode = 0;
        c.refresh = 0;
        while (!done) {
            if (D.control == 0) {
                if (c.gi + 1 >= D.n_grid) break;
                c.t = D.tgrid[c.gi];
                c.h = dt;
                c.t7 = D.tgrid[c.gi + 1];
                c.tobs = c.t7;
                c.ends = 1;
            } else {
                if (D.control == 1) {
                    if (!dp_less(c.t, D.T1)) break;
                    c.tend = D.T1;
                } else if (c.fresh) {
                    if (!dp_less_eq(c.tcur + dt, c.s1)) {
                        // segment done: the next integrate_const call
                        if (++c.sg >= nseg) break;
                        c.s0 = sg[size_t(c.sg) * 2];
                        c.s1 = sg[size_t(c.sg) * 2 + 1];
                        c.kk = 0;
                        c.t = c.tcur = c.te = c.s0;
                        c.h = dt;
                        c.mode = 2;
                        c.eobs = 1;
                        break;
                    }
                    c.tend = c.tcur + dt;
                    if (!dp_less(c.tcur, c.tend)) {
                        ++c.kk;
                        c.t = c.tcur = c.te = dp_grid_time(c.s0, c.kk, dt);
                        c.mode = 2;
                        c.eobs = 1;
                        break;
                    }
                    c.fresh = 0;
                    c.fails = 0;
                    // a new stepper copy takes the derivative at tcur; it
                    // differs from the carried one only if the accepted
                    // time is not tcur itself and the drive depends on t
                    c.refresh = (drive_on && c.t != c.tcur) ? 1 : 0;
                    c.t = c.te = c.tcur;
                }
                if (c.fails == 0 && dp_less(c.tend, c.t + c.h)) c.h = c.tend - c.t;
                if (!(c.h > 0.0) || !(c.t + c.h > c.t)) {
                    c.status = 3;
                    break;
                }
                c.t7 = c.t + c.h;
                c.tobs = D.control == 1 ? c.t7 : dp_grid_time(c.s0, c.kk + 1, dt);
                c.ends = (D.control == 1 || !dp_less(c.t + c.h, c.tend)) ? 1 : 0;
            }
            if (c.attempts >= D.max_attempts) {
                c.status = 2;
                break;
            }
            ++c.attempts;
            c.mode = 1;
            break;
        }
        if (c.mode == 1) c.obs = (c.ends && (c.obs_calls % D.save_interval) == 0) ? 1 : 0;
        if (c.mode == 2) c.obs = (c.eobs && (c.obs_calls % D.save_interval) == 0) ? 1 : 0;
        *cp = c;
        pub_ri = (ri >= 0 && ri < D.max_rec) ? ri : -1;
        pub_t = rt;
    }
    __syncthreads();
    const long long ri = pub_ri;
    if (ri >= 0 && tid < kRec) {
        const double tot = (red[tid][0] + red[tid][1]) + (red[tid][2] + red[tid][3]);
        const double sc = tid < 9 ? D.inv2 : (tid < 33 ? D.inv3 : D.invE);
        const size_t r = size_t(b) * D.max_rec + size_t(ri);
        D.recs[r * kRec + tid] = tot * sc;
        if (tid == 0) D.times[r] = pub_t;
    }
}

// X0 <- X1 for the trajectories whose final state is in X1
__global__ void k_mixdp_final(const MixDpCtl* __restrict__ ctl, const double* __restrict__ X1,
                              double* __restrict__ X0, size_t sz) {
    const int b = blockIdx.y;
    if (ctl[b].par == 0) return;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < sz) X0[size_t(b) * sz + i] = X1[size_t(b) * sz + i];
}

}  // namespace csl

using namespace cslh;

namespace {

// Host form of one ELL bond kind.
struct EllHost {
    int zmax = 0;
    bool uniform = true;
    std::vector<int> col, cnt, cnt_atom;
    std::vector<double> mats;
};

// CSR (reference per-site order) -> ELL in the sublattice-major layout.  For
// translation-invariant tables the bonds of each site are put in its
// sublattice's majority slot order (sites on wrapped cells list the same
// bonds in another order) and one matrix per (sublattice, slot) is kept.
EllHost build_ell(size_t N, size_t A, int npad, const size_t* off, const size_t* par, const double* J, int m,
                  size_t N_partner, size_t A_partner) {
    EllHost E;
    const int NP = int(A) * npad;
    for (size_t i = 0; i < N; ++i) {
        if (off[i + 1] < off[i]) fail(CSL_ERR_ARG, "bond offsets must be non-decreasing");
        E.zmax = std::max<int>(E.zmax, int(off[i + 1] - off[i]));
    }
    if (E.zmax > 64) fail(CSL_ERR_UNSUPPORTED, "more than 64 bonds of one kind per site");
    for (size_t k = 0; k < off[N]; ++k) {
        if (par[k] >= N_partner) fail(CSL_ERR_ARG, "bond partner index out of range");
        for (int q = 0; q < m; ++q)
            if (!std::isfinite(J[k * m + q])) fail(CSL_ERR_NONFINITE, "non-finite bond matrix");
    }
    auto pdev = [&](size_t s) { return int(s % A_partner) * npad + int(s / A_partner); };
    auto sdev = [&](size_t s) { return int(s % A) * npad + int(s / A); };
    std::vector<size_t> slot_of(off[N]);
    for (size_t i = 0; i < N; ++i)
        for (size_t k = off[i]; k < off[i + 1]; ++k) slot_of[k] = k - off[i];
    std::vector<size_t> canon(A, 0);
    auto same = [&](size_t a, size_t b) {
        const size_t ca = off[a + 1] - off[a];
        return ca == off[b + 1] - off[b] && std::memcmp(J + off[a] * m, J + off[b] * m, ca * m * sizeof(double)) == 0;
    };
    bool uni = true;
    for (size_t at = 0; at < A && uni; ++at) {
        size_t cand = at, votes = 0;
        for (size_t i = at; i < N; i += A) {
            if (votes == 0) { cand = i; votes = 1; }
            else if (same(i, cand)) ++votes;
            else --votes;
        }
        canon[at] = cand;
        const size_t c = off[cand + 1] - off[cand];
        for (size_t i = at; i < N && uni; i += A) {
            if (same(i, cand)) continue;
            if (off[i + 1] - off[i] != c) { uni = false; break; }
            std::vector<char> used(c, 0);
            for (size_t q = 0; q < c && uni; ++q) {
                bool found = false;
                for (size_t r = 0; r < c; ++r)
                    if (!used[r] && std::memcmp(J + (off[i] + r) * m, J + (off[cand] + q) * m, m * sizeof(double)) == 0) {
                        used[r] = 1;
                        slot_of[off[i] + r] = q;
                        found = true;
                        break;
                    }
                if (!found) uni = false;
            }
        }
    }
    if (!uni)
        for (size_t i = 0; i < N; ++i)
            for (size_t k = off[i]; k < off[i + 1]; ++k) slot_of[k] = k - off[i];
    E.uniform = uni;
    const int z = std::max(E.zmax, 1);
    E.col.assign(size_t(z) * NP, 0);
    E.cnt.assign(NP, 0);
    E.cnt_atom.assign(A, 0);
    E.mats.assign(uni ? A * size_t(z) * m : size_t(z) * NP * m, 0.0);
    for (size_t i = 0; i < N; ++i) {
        const int p = sdev(i);
        const int c = int(off[i + 1] - off[i]);
        E.cnt[p] = c;
        for (int k = 0; k < c; ++k) {
            const size_t q = slot_of[off[i] + k];
            E.col[q * NP + p] = pdev(par[off[i] + k]);
            double* dst = uni ? &E.mats[((i % A) * z + q) * m] : &E.mats[(q * NP + p) * m];
            std::memcpy(dst, J + (off[i] + k) * m, m * sizeof(double));
        }
    }
    for (size_t a = 0; a < A; ++a) E.cnt_atom[a] = int(off[canon[a] + 1] - off[canon[a]]);
    return E;
}

// For translation-invariant SU(2)->SU(3) and SU(3)->SU(2) tables: the
// offset in e23.mats of the matrix whose transpose is the SU(3)->SU(2)
// matrix of each (SU(3) sublattice, slot) — the reverse bond of the same
// pair, with the same slot for every cell and a bit-identical transposed
// matrix — as [A3][8]; empty when any slot has no such twin or a slot
// exceeds 8.
std::vector<int> transpose_offsets(const EllHost& e23, const EllHost& e32, size_t A2, size_t A3, int n_cells,
                                   int npad) {
    const int z23 = std::max(e23.zmax, 1), z32 = std::max(e32.zmax, 1);
    if (!e23.uniform || !e32.uniform || e32.zmax > 8 || e23.zmax == 0 || e32.zmax == 0) return {};
    const size_t NP2 = A2 * size_t(npad), NP3 = A3 * size_t(npad);
    std::vector<int> off(A3 * 8, 0);
    for (size_t ta = 0; ta < A3; ++ta) {
        const int c = e32.cnt_atom[ta];
        for (int tk = 0; tk < c; ++tk) {
            const double* J32 = &e32.mats[(ta * z32 + tk) * 24];
            int found = -1;
            for (int cell = 0; cell < n_cells; ++cell) {
                const size_t p = ta * size_t(npad) + size_t(cell);
                const int q = e32.col[size_t(tk) * NP3 + p];
                const size_t fa = size_t(q) / size_t(npad);
                if (fa >= A2) return {};
                int here = -1;
                for (int fk = 0; fk < e23.cnt_atom[fa] && here < 0; ++fk) {
                    if (e23.col[size_t(fk) * NP2 + size_t(q)] != int(p)) continue;
                    const double* J23 = &e23.mats[(fa * z23 + fk) * 24];
                    bool same = true;
                    for (int d = 0; d < 8 && same; ++d)
                        for (int e = 0; e < 3 && same; ++e) same = J32[d * 3 + e] == J23[e * 8 + d];
                    if (same) here = int((fa * z23 + fk) * 24);
                }
                if (here < 0 || (found >= 0 && here != found)) return {};
                found = here;
            }
            off[ta * 8 + tk] = found;
        }
    }
    return off;
}

struct EllDev {
    DBuf<int> col, cnt, cnt_atom;
    DBuf<double> mats;
    void upload(const EllHost& E, csl::Ell& d, hipStream_t s) {
        col.upload(E.col.data(), E.col.size(), s);
        cnt.upload(E.cnt.data(), E.cnt.size(), s);
        cnt_atom.upload(E.cnt_atom.data(), E.cnt_atom.size(), s);
        mats.upload(E.mats.data(), E.mats.size(), s);
        d.col = col.p;
        d.cnt = cnt.p;
        d.cnt_atom = cnt_atom.p;
        d.mats = mats.p;
        d.zmax = std::max(E.zmax, 1);
        d.uniform = E.uniform ? 1 : 0;
    }
};

}  // namespace

struct csl_mixed {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t N2 = 0, N3 = 0, A2 = 0, A3 = 0, n_cells = 0;
    csl::MixDev M{};
    EllDev e22, e23, e33, e32;
    DBuf<double> field2, onsite2, field3, onsite3, frames2, afm2, frames3, fdrive2, fdrive3, eq3;
    DBuf<double> state;                      // SoA [SZ]
    bool has_state = false;
    DBuf<double> X, ACC, Y1, Y2, partials, recs, tp, tp3, scale, init, flat, aux;
    DBuf<double> KS;                         // stage derivatives of the generic fixed-step engine [S][SZ]
    DBuf<double> ZX, ZACC, ZY1, ZY2, tpartials, trecs, zaux, target, loss, grad;  // tangent-linear runs
    // Dormand-Prince runs: x / k1 pairs, k2..k6, stage inputs, controller state
    DBuf<double> dpX0, dpX1, dpK10, dpK11, dpK2, dpK3, dpK4, dpK5, dpK6, dpY0, dpY1, dp_errp, dp_recs, dp_times,
        dp_segs, dp_grid;
    DBuf<int> dp_nseg;
    DBuf<csl::MixDpCtl> dp_ctl;
    double h_tp2[2] = {0.0, 0.0}, h_tp3[2] = {0.0, 0.0};  // single-trajectory pulse centres
    double last_ms = 0.0;
    size_t last_launches = 0;
    // persistent batched kernel: 0 when the lattice qualifies, -1 stage
    // kernels, n > 0 at most n co-resident workgroups
    int persist_mode = 0;
    int last_persist = 0, last_persist_grid = 0;
    size_t last_split = 0;  // trajectories the persistent kernel ran in two pieces
    double last_persist_ms = 0.0;  // the persistent launch alone
    DBuf<double> pG, ptgrid;
    DBuf<int> j32off;  // see MixDev::j32off
    DBuf<csl::MixDev> pM;  // device copy of M for the persistent kernel
    DBuf<unsigned> coop_ctl;  // cooperative kernel: barrier counter, status
    DBuf<int> psched;         // work split of the persistent kernel
    DBuf<unsigned> pflags;    // its seam flags [grid] and status
    int zmax[4] = {0, 0, 0, 0};  // bonds per site of kinds 22, 23, 33, 32 (0: none)
    // batched stage-kernel launch sequence replayed from a hipGraph
    int stage_graph = 0;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<unsigned char> graph_key;
    ~csl_mixed() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (stream) (void)hipStreamDestroy(stream);
    }
    size_t sz() const { return size_t(3) * M.NP2 + size_t(8) * M.NP3; }
    size_t nflat() const { return N2 * 3 + N3 * 8; }
};

namespace {

csl_mixed* check_mixed(csl_mixed* h) {
    if (!h) fail(CSL_ERR_NULL, "null mixed-lattice handle");
    hip_check(hipSetDevice(h->device), "hipSetDevice");
    return h;
}

// per-site field / onsite -> per-sublattice when translation invariant; 2A
template <int D>
void site_tables(size_t N, size_t A, int npad, const double* field, const double* onsite, DBuf<double>& f_dev,
                 DBuf<double>& o_dev, int& site_uniform, int& has_onsite, hipStream_t s) {
    check_finite(field, N * D, "field");
    check_finite(onsite, N * D * D, "onsite");
    bool uni = true;
    for (size_t i = 0; i < N && uni; ++i) {
        const size_t r = i % A;
        if (std::memcmp(field + i * D, field + r * D, D * sizeof(double)) ||
            std::memcmp(onsite + i * D * D, onsite + r * D * D, D * D * sizeof(double)))
            uni = false;
    }
    std::vector<double> f, o;
    if (uni) {
        f.assign(field, field + A * D);
        o.assign(onsite, onsite + A * D * D);
    } else {
        const size_t NP = A * npad;
        f.assign(NP * D, 0.0);
        o.assign(NP * D * D, 0.0);
        for (size_t i = 0; i < N; ++i) {
            const size_t p = (i % A) * npad + i / A;
            std::memcpy(&f[p * D], field + i * D, D * sizeof(double));
            std::memcpy(&o[p * D * D], onsite + i * D * D, D * D * sizeof(double));
        }
    }
    bool any = false;
    for (double& v : o) {
        any |= (v != 0.0);
        v *= 2.0;
    }
    site_uniform = uni ? 1 : 0;
    has_onsite = any ? 1 : 0;
    f_dev.upload(f.data(), f.size(), s);
    o_dev.upload(o.data(), o.size(), s);
}

int blocks_of(int n) { return (n + 255) / 256; }

void to_soa(csl_mixed* h, const double* d_flat, size_t src_stride, double* d_soa, size_t B) {
    const long n = long(h->nflat());
    dim3 g(unsigned((n + 255) / 256), unsigned(B));
    hipLaunchKernelGGL(csl::k_mixed_perm, g, dim3(256), 0, h->stream, h->M, d_flat, src_stride, d_soa, h->sz(), 1);
    hip_check(hipGetLastError(), "k_mixed_perm");
}

void to_flat(csl_mixed* h, const double* d_soa, double* d_flat, size_t B) {
    const long n = long(h->nflat());
    dim3 g(unsigned((n + 255) / 256), unsigned(B));
    hipLaunchKernelGGL(csl::k_mixed_perm, g, dim3(256), 0, h->stream, h->M, d_soa, h->sz(), d_flat, h->nflat(), 0);
    hip_check(hipGetLastError(), "k_mixed_perm");
}

// Integrate B trajectories whose SoA states are in h->X; records into h->recs
// ([n_rec][B][34]); optional state snapshots (flat, B = 1) into d_states.
// With tl the tangent z = d state / d theta (h->ZX, zero at the start) is
// advanced in the same launches and its records go to h->trecs.
// X0 / recs_out / rec_stride: a sub-batch whose states start at X0 and whose
// records go to recs_out with row stride rec_stride (defaults: h->X, h->recs, B).
void run_mixed(csl_mixed* h, size_t B, const double* d_tp, const double* d_tp3, const double* d_scale,
               const std::vector<double>& grid, double dt, size_t si, double* d_states, bool tl = false,
               bool use_graph = false, double* X0 = nullptr, double* recs_out = nullptr, size_t rec_stride = 0) {
    const size_t n_steps = grid.size() - 1;
    const size_t n_rec = records_for(n_steps, si);
    hipStream_t s = h->stream;
    if (!recs_out) h->recs.ensure(n_rec * B * CSL_MIXED_RECORD);
    double* const Xp = X0 ? X0 : h->X.p;
    double* const Rp = recs_out ? recs_out : h->recs.p;
    const int RS = int(rec_stride ? rec_stride : B);
    const size_t plane = B * h->sz();
    h->ACC.ensure(plane);
    h->Y1.ensure(plane);
    h->Y2.ensure(plane);
    const int nblk2 = blocks_of(h->M.NP2), nblk3 = blocks_of(h->M.NP3);
    const int nblk = nblk2 + nblk3;
    h->partials.ensure(B * size_t(nblk) * CSL_MIXED_RECORD);
    if (tl) {
        h->ZACC.ensure(plane);
        h->ZY1.ensure(plane);
        h->ZY2.ensure(plane);
        h->tpartials.ensure(B * size_t(nblk) * CSL_MIXED_RECORD);
        h->trecs.ensure(n_rec * B * CSL_MIXED_RECORD);
    }
    const bool drive = d_tp && (h->M.amp2 != 0.0 || h->M.amp3 != 0.0);
    const double inv2 = 1.0 / double(h->N2), inv3 = 1.0 / double(h->N3), invE = 1.0 / double(h->N2 + h->N3);
    h->aux.ensure(plane);
    if (tl) h->zaux.ensure(plane);
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    size_t launches = 0;
    const dim3 grid_dim{unsigned(nblk), unsigned(B), 1u};
    auto stage = [&](int st, double t, int rec) {
        csl::MixStageArgs A{};
        A.stage = st;
        A.nblk2 = nblk2;
        A.x = Xp;
        A.acc = h->ACC.p;
        A.ycur = st == 0 ? Xp : (st == 2 ? h->Y2.p : h->Y1.p);
        A.ynext = st == 0 ? h->Y1.p : (st == 1 ? h->Y2.p : (st == 2 ? h->Y1.p : h->Y2.p));
        A.t = t;
        A.dt = dt;
        A.tp = drive ? d_tp : nullptr;
        A.tp3 = drive ? d_tp3 : nullptr;
        A.scale = drive ? d_scale : nullptr;
        A.rec = rec;
        A.partials = h->partials.p;
        if (tl) {
            A.zx = h->ZX.p;
            A.zacc = h->ZACC.p;
            A.zcur = st == 0 ? h->ZX.p : (st == 2 ? h->ZY2.p : h->ZY1.p);
            A.znext = st == 0 ? h->ZY1.p : (st == 1 ? h->ZY2.p : (st == 2 ? h->ZY1.p : h->ZY2.p));
            A.tpartials = h->tpartials.p;
            hipLaunchKernelGGL(csl::k_mixed_stage<true>, grid_dim, dim3(256), 0, s, h->M, A);
        } else {
            hipLaunchKernelGGL(csl::k_mixed_stage<false>, grid_dim, dim3(256), 0, s, h->M, A);
        }
        hip_check(hipGetLastError(), "k_mixed_stage");
        ++launches;
    };
    auto reduce = [&](size_t ri) {
        hipLaunchKernelGGL(csl::k_mixed_reduce, dim3(unsigned(B)), dim3(256), 0, s, h->partials.p, nblk, int(B),
                           int(ri), inv2, inv3, invE, Rp, RS);
        hip_check(hipGetLastError(), "k_mixed_reduce");
        if (tl) {
            hipLaunchKernelGGL(csl::k_mixed_reduce, dim3(unsigned(B)), dim3(256), 0, s, h->tpartials.p, nblk,
                               int(B), int(ri), inv2, inv3, invE, h->trecs.p, int(B));
            hip_check(hipGetLastError(), "k_mixed_reduce");
        }
    };
    // The whole launch sequence; with use_graph it is captured once into a
    // hipGraph and replayed while the key (shapes, time grid, device
    // buffers, lattice parameters) is unchanged.
    auto issue = [&](bool timed) {
    if (timed) hip_check(hipEventRecord(e0, s), "event record");
    launches = 0;
    for (size_t n = 0; n < n_steps; ++n) {
        const bool rec = (n % si) == 0;
        const size_t ri = n / si;
        if (rec && d_states) to_flat(h, Xp, d_states + ri * h->nflat(), 1);
        const double t = grid[n];
        // Boost runge_kutta4 stage times t, t + dt/2, t + dt/2, t + dt
        stage(0, t, rec ? 1 : 0);
        stage(1, t + 0.5 * dt, 0);
        stage(2, t + 0.5 * dt, 0);
        stage(3, t + 1.0 * dt, 0);
        if (rec) reduce(ri);
    }
    if (timed) hip_check(hipEventRecord(e1, s), "event record");
    // final record: a stage-0 evaluation whose outputs go to scratch
    {
        csl::MixStageArgs A{};
        A.stage = 0;
        A.nblk2 = nblk2;
        A.x = Xp;
        A.acc = h->aux.p;
        A.ycur = Xp;
        A.ynext = h->aux.p;
        A.t = grid[n_steps];
        A.dt = dt;
        A.rec = 1;
        A.partials = h->partials.p;
        if (tl) {
            A.zx = h->ZX.p;
            A.zacc = h->zaux.p;
            A.zcur = h->ZX.p;
            A.znext = h->zaux.p;
            A.tpartials = h->tpartials.p;
            hipLaunchKernelGGL(csl::k_mixed_stage<true>, grid_dim, dim3(256), 0, s, h->M, A);
        } else {
            hipLaunchKernelGGL(csl::k_mixed_stage<false>, grid_dim, dim3(256), 0, s, h->M, A);
        }
        hip_check(hipGetLastError(), "k_mixed_stage");
        reduce(n_rec - 1);
    }
    if (d_states) to_flat(h, Xp, d_states + (n_rec - 1) * h->nflat(), 1);
    };
    if (!use_graph) {
        issue(true);
    } else {
        std::vector<unsigned char> key;
        auto put = [&](const void* p, size_t nb) {
            const auto* c = static_cast<const unsigned char*>(p);
            key.insert(key.end(), c, c + nb);
        };
        put(&h->M, sizeof(csl::MixDev));
        put(grid.data(), grid.size() * sizeof(double));
        const size_t shape[5] = {B, si, size_t(tl ? 1 : 0), grid.size(), size_t(RS)};
        put(shape, sizeof(shape));
        put(&dt, sizeof(dt));
        const void* ptrs[] = {Xp, h->ACC.p, h->Y1.p, h->Y2.p, h->partials.p, Rp, h->aux.p, d_tp, d_tp3,
                              d_scale, d_states, tl ? h->ZX.p : nullptr, tl ? h->ZACC.p : nullptr,
                              tl ? h->ZY1.p : nullptr, tl ? h->ZY2.p : nullptr, tl ? h->tpartials.p : nullptr,
                              tl ? h->trecs.p : nullptr, tl ? h->zaux.p : nullptr};
        put(ptrs, sizeof(ptrs));
        if (!h->graph_exec || key != h->graph_key) {
            if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
            h->graph_exec = nullptr;
            hipGraph_t g = nullptr;
            hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
            try {
                issue(false);
            } catch (...) {
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hip_check(hipStreamEndCapture(s, &g), "end capture");
            const hipError_t ie = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hip_check(ie, "graph instantiate");
            h->graph_key = key;
        }
        hip_check(hipEventRecord(e0, s), "event record");
        hip_check(hipGraphLaunch(h->graph_exec, s), "graph launch");
        hip_check(hipEventRecord(e1, s), "event record");
        launches = 4 * n_steps;  // stage launches of the replayed sequence
    }
    hip_check(hipStreamSynchronize(s), "mixed integration");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->last_ms = launches ? ms / double(launches) : 0.0;
    h->last_launches = launches;
}

// B trajectories (states at X0, records to recs_out with row stride
// rec_stride) on the cooperative stage kernel: one
// launch for the whole time grid.  False (nothing launched) when its
// workgroups cannot all be resident; the caller checks h->coop_ctl[1] after
// synchronising.
bool run_mixed_coop(csl_mixed* h, size_t B, const double* d_tp, const double* d_tp3, const double* d_scale,
                    const std::vector<double>& grid, double dt, size_t si, double* X0, double* recs_out,
                    size_t rec_stride) {
    const size_t n_steps = grid.size() - 1;
    const size_t n_rec = records_for(n_steps, si);
    hipStream_t s = h->stream;
    const int nblk2 = blocks_of(h->M.NP2), nblk3 = blocks_of(h->M.NP3);
    const int nblk = nblk2 + nblk3;
    int occ = 0, n_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, csl::k_mixed_coop, 256, 0) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    hip_check(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, h->device), "CU count");
    if (size_t(nblk) * B > size_t(occ) * size_t(n_cu)) return false;
    const size_t plane = B * h->sz();
    h->ACC.ensure(plane);
    h->Y1.ensure(plane);
    h->Y2.ensure(plane);
    h->aux.ensure(plane);
    h->partials.ensure(B * size_t(nblk) * CSL_MIXED_RECORD);
    h->coop_ctl.ensure(2);
    hip_check(csl::launch_fill_u32(h->coop_ctl.p, 0u, 2, s), "fill");
    const bool drive = d_tp && (h->M.amp2 != 0.0 || h->M.amp3 != 0.0);
    csl::MixCoopArgs C{};
    C.A.nblk2 = nblk2;
    C.A.x = X0;
    C.A.acc = h->ACC.p;
    C.A.ycur = h->Y2.p;
    C.A.ynext = h->Y1.p;
    C.A.dt = dt;
    C.A.tp = drive ? d_tp : nullptr;
    C.A.tp3 = drive ? d_tp3 : nullptr;
    C.A.scale = drive ? d_scale : nullptr;
    C.A.partials = h->partials.p;
    C.aux = h->aux.p;
    C.tgrid = h->ptgrid.p;
    C.n_steps = int(n_steps);
    C.save_interval = int(si);
    C.n_rec = int(n_rec);
    C.rec_stride = int(rec_stride);
    C.recs = recs_out;
    C.inv2 = 1.0 / double(h->N2);
    C.inv3 = 1.0 / double(h->N3);
    C.invE = 1.0 / double(h->N2 + h->N3);
    C.ctr = h->coop_ctl.p;
    C.status = h->coop_ctl.p + 1;
    hipLaunchKernelGGL(csl::k_mixed_coop, dim3(unsigned(nblk), unsigned(B), 1u), dim3(256), 0, s, h->M, C);
    hip_check(hipGetLastError(), "k_mixed_coop");
    return true;
}

// Launch shape of the persistent kernel for this lattice (block 0 when it
// does not qualify: it needs translation-invariant bond tables of all four
// kinds, per-sublattice fields and onsite matrices, at most 8 bonds per site
// of kinds 22, 23, 32 and 4 of kind 33, and at most 2048 sites of each
// species, so that the stage values fit LDS plus two global components).
csl::MixPersistShape mixed_persist_shape(const csl_mixed* h) {
    const csl::MixDev& M = h->M;
    if (!(M.e22.uniform && M.e23.uniform && M.e33.uniform && M.e32.uniform && M.site_uniform2 && M.site_uniform3))
        return {0, 0, false};
    if (h->zmax[0] > csl::kMixZ22 || h->zmax[1] > csl::kMixZ23 || h->zmax[2] > csl::kMixZ33 ||
        h->zmax[3] > csl::kMixZ32)
        return {0, 0, false};
    return csl::mixed_persist_shape(M.NP2, M.NP3, h->zmax[2] > 0);
}

// The batch on the persistent kernel (states in h->X, records into h->recs):
// one workgroup per trajectory, as many as are co-resident.  With more
// trajectories than workgroups, workgroup g runs the range [g W, (g+1) W) of
// the trajectory-major step order (W = ceil(B n / grid)): the head of the
// trajectory it shares with g + 1 first, then whole trajectories, last the
// rest of the trajectory begun by g - 1, from the state that workgroup left
// in X (a seam flag per workgroup).  False (and nothing launched) when the
// kernel cannot be resident.
bool run_mixed_persist(csl_mixed* h, size_t B, const double* d_tp, const double* d_tp3, const double* d_scale,
                       const std::vector<double>& grid, double dt, size_t si) {
    const csl::MixPersistShape shape = mixed_persist_shape(h);
    if (!shape.block) return false;
    const size_t n_steps = grid.size() - 1;
    const size_t n_rec = records_for(n_steps, si);
    hipStream_t s = h->stream;
    csl::MixPersistArgs A{};
    int occ = 0;
    h->pM.upload(&h->M, 1, s);
    if (csl::launch_mixed_persist(h->M, h->pM.p, A, shape, 0, s, &occ) != hipSuccess || occ < 1) {
        (void)hipGetLastError();
        return false;
    }
    int n_cu = 0;
    hip_check(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, h->device), "CU count");
    size_t co = size_t(occ) * size_t(n_cu);
    if (h->persist_mode > 0) co = std::min(co, size_t(h->persist_mode));
    const size_t grid_wg = std::min(B, co);
    size_t n_split = 0;
    if (B > grid_wg && n_steps > 0) {
        size_t max_p = 1;
        const std::vector<int> sched = split_schedule(B, grid_wg, n_steps, max_p);
        h->psched.upload(sched.data(), sched.size(), s);
        h->pflags.ensure(grid_wg + 1);
        hip_check(csl::launch_fill_u32(h->pflags.p, 0u, grid_wg + 1, s), "fill");
        for (size_t r = 0; r < sched.size(); r += 5) n_split += sched[r + 4] >= 0;
        A.sched = h->psched.p;
        A.max_pieces = int(max_p);
        A.flags = h->pflags.p;
        A.status = h->pflags.p + grid_wg;
    }
    h->recs.ensure(n_rec * B * CSL_MIXED_RECORD);
    h->pG.ensure(grid_wg * 4 * size_t(h->M.NP3));
    h->ptgrid.upload(grid.data(), grid.size(), s);
    A.b0 = 0;
    A.B = int(B);
    A.B_rec = int(B);
    A.n_steps = int(n_steps);
    A.save_interval = int(si);
    A.n_rec = int(n_rec);
    A.dt = dt;
    A.tgrid = h->ptgrid.p;
    A.X = h->X.p;
    A.G = h->pG.p;
    const bool drive = d_tp && (h->M.amp2 != 0.0 || h->M.amp3 != 0.0);
    A.tp = drive ? d_tp : nullptr;
    A.tp3 = drive ? d_tp3 : nullptr;
    A.scale = drive ? d_scale : nullptr;
    A.recs = h->recs.p;
    A.inv2 = 1.0 / double(h->N2);
    A.inv3 = 1.0 / double(h->N3);
    A.invE = 1.0 / double(h->N2 + h->N3);
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, s), "event record");
    hip_check(csl::launch_mixed_persist(h->M, h->pM.p, A, shape, int(grid_wg), s, nullptr), "k_mixed_persist");
    hip_check(hipEventRecord(e1, s), "event record");
    hip_check(hipStreamSynchronize(s), "mixed persistent integration");
    if (A.sched) {
        unsigned st = 0;
        hip_check(hipMemcpy(&st, A.status, sizeof(unsigned), hipMemcpyDeviceToHost), "status");
        if (st) fail(CSL_ERR_HIP, "mixed persistent kernel: a wait for a trajectory's head piece timed out");
    }
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->last_ms = ms;
    h->last_persist_ms = ms;
    h->last_launches = 1;
    h->last_persist = 1;
    h->last_persist_grid = int(grid_wg);
    h->last_split = n_split;
    return true;
}

// The whole batch on the cooperative stage kernel (states in h->X, records
// into h->recs): one launch for the time grid.  False (nothing launched)
// when its workgroups cannot all be resident.
bool run_mixed_coop_batch(csl_mixed* h, size_t B, const double* d_tp, const double* d_tp3, const double* d_scale,
                          const std::vector<double>& grid, double dt, size_t si) {
    const size_t n_rec = records_for(grid.size() - 1, si);
    hipStream_t s = h->stream;
    h->recs.ensure(n_rec * B * CSL_MIXED_RECORD);
    h->ptgrid.upload(grid.data(), grid.size(), s);
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, s), "event record");
    const bool ran = run_mixed_coop(h, B, d_tp, d_tp3, d_scale, grid, dt, si, h->X.p, h->recs.p, B);
    hip_check(hipEventRecord(e1, s), "event record");
    hip_check(hipStreamSynchronize(s), "cooperative integration");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!ran) return false;
    unsigned st = 0;
    hip_check(hipMemcpy(&st, h->coop_ctl.p + 1, sizeof(unsigned), hipMemcpyDeviceToHost), "status");
    if (st) fail(CSL_ERR_HIP, "cooperative stage kernel: a grid barrier wait timed out");
    h->last_ms = ms;
    h->last_persist_ms = ms;
    h->last_launches = 1;
    h->last_persist = 2;
    h->last_persist_grid = int(B) * (blocks_of(h->M.NP2) + blocks_of(h->M.NP3));
    h->last_split = 0;
    return true;
}

void need_state(const csl_mixed* h) {
    if (!h->has_state) fail(CSL_ERR_NOSTATE, "no spin state set");
}

// Method name of the reference GPU stepper -> stage table (unknown names are
// reported on stderr and run ssprk53, integrator.cuh:454-458).
Tableau mixed_method_tableau(const char* method) {
    bool known = true;
    Tableau T = tableau_for(method, &known);
    if (!known)
        std::fprintf(stderr, "Warning: Unknown GPU integration method '%s', using ssprk53\n",
                     method ? method : "");
    return T;
}

// Generic fixed-step engine for the reference GPU stepper's methods other
// than rk4 (integrator.cuh:91-458; the mixed lattice steps through the same
// gpu::ode::step, mixed_lattice_gpu.cu:456-488), one trajectory in h->X:
// per stage one k_mixed_eval launch (derivative of the stage input at
// t + c_s dt with host-evaluated envelopes into K[s]) and one k_rk_combine
// launch (next stage input, or the result in place of x).  Records at every
// save_interval-th step start and after the last step, as run_mixed.
void run_mixed_tableau(csl_mixed* h, const Tableau& T, const std::vector<double>& grid, double dt, size_t si,
                       double* d_states) {
    const size_t n_steps = grid.size() - 1;
    const size_t n_rec = records_for(n_steps, si);
    hipStream_t s = h->stream;
    const size_t n = h->sz();
    const int S = T.S;
    h->KS.ensure(size_t(S) * n);
    h->Y1.ensure(n);
    h->aux.ensure(n);
    hip_check(hipMemsetAsync(h->KS.p, 0, size_t(S) * n * sizeof(double), s), "memset");
    hip_check(hipMemsetAsync(h->Y1.p, 0, n * sizeof(double), s), "memset");
    const int nblk2 = blocks_of(h->M.NP2), nblk3 = blocks_of(h->M.NP3);
    const int nblk = nblk2 + nblk3;
    h->partials.ensure(size_t(nblk) * CSL_MIXED_RECORD);
    h->recs.ensure(n_rec * CSL_MIXED_RECORD);
    const csl::MixDev& M = h->M;
    const double inv2 = 1.0 / double(h->N2), inv3 = 1.0 / double(h->N3), invE = 1.0 / double(h->N2 + h->N3);
    std::vector<csl::RkCombine> rows(static_cast<size_t>(S));
    for (int st = 0; st < S; ++st) {
        const Lin& l = st + 1 < S ? T.in[st + 1] : T.out;
        rows[size_t(st)].g = l.g;
        rows[size_t(st)].m = lin_terms(l, st + 1, dt, rows[size_t(st)].idx, rows[size_t(st)].w);
    }
    auto env = [](double amp, double t, double c, double w, double fr) {
        if (amp == 0.0) return 0.0;
        const double d = t - c;
        const double a = d / (2.0 * w);
        return amp * std::exp(-(a * a)) * std::cos(fr * d);
    };
    // record of the state in X: a stage-0 evaluation whose outputs go to scratch
    auto record = [&](size_t ri) {
        if (d_states) to_flat(h, h->X.p, d_states + ri * h->nflat(), 1);
        csl::MixStageArgs A{};
        A.stage = 0;
        A.nblk2 = nblk2;
        A.x = h->X.p;
        A.acc = h->aux.p;
        A.ycur = h->X.p;
        A.ynext = h->aux.p;
        A.dt = dt;
        A.rec = 1;
        A.partials = h->partials.p;
        hipLaunchKernelGGL(csl::k_mixed_stage<false>, dim3(unsigned(nblk), 1u, 1u), dim3(256), 0, s, M, A);
        hip_check(hipGetLastError(), "k_mixed_stage");
        hipLaunchKernelGGL(csl::k_mixed_reduce, dim3(1u), dim3(256), 0, s, h->partials.p, nblk, 1, int(ri), inv2,
                           inv3, invE, h->recs.p, 1);
        hip_check(hipGetLastError(), "k_mixed_reduce");
    };
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, s), "event record");
    for (size_t k = 0; k < n_steps; ++k) {
        if (k % si == 0) record(k / si);
        const double t = grid[k];
        for (int st = 0; st < S; ++st) {
            const double ts = t + T.c[st] * dt;
            const double f21 = env(M.amp2, ts, h->h_tp2[0], M.width2, M.freq2);
            const double f22 = env(M.amp2, ts, h->h_tp2[1], M.width2, M.freq2);
            const double f31 = env(M.amp3, ts, h->h_tp3[0], M.width3, M.freq3);
            const double f32 = env(M.amp3, ts, h->h_tp3[1], M.width3, M.freq3);
            const double* in = st == 0 ? h->X.p : h->Y1.p;
            hipLaunchKernelGGL(csl::k_mixed_eval, dim3(unsigned(nblk)), dim3(256), 0, s, M, in, f21, f22, f31, f32,
                               h->KS.p + size_t(st) * n);
            hip_check(hipGetLastError(), "k_mixed_eval");
            double* out = st + 1 < S ? h->Y1.p : h->X.p;
            hip_check(csl::launch_rk_combine(out, h->X.p, h->KS.p, n, rows[size_t(st)], s), "k_rk_combine");
        }
    }
    hip_check(hipEventRecord(e1, s), "event record");
    record(n_rec - 1);
    hip_check(hipStreamSynchronize(s), "mixed integration");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->last_launches = n_steps * size_t(S) * 2;
    h->last_ms = h->last_launches ? ms / double(h->last_launches) : 0.0;
}

}  // namespace

extern "C" {

int csl_mixed_create(size_t A2, size_t A3, size_t n_cells, const double* field2, const double* onsite2,
                     const size_t* off22, const size_t* par22, const double* J22, const double* field3,
                     const double* onsite3, const size_t* off33, const size_t* par33, const double* J33,
                     const size_t* off23, const size_t* par23, const double* J23, const size_t* off32,
                     const size_t* par32, const double* J32, csl_mixed** out) {
    return guarded([&] {
        if (!out || !field2 || !onsite2 || !field3 || !onsite3 || !off22 || !off33 || !off23 || !off32)
            fail(CSL_ERR_NULL, "null table pointer");
        *out = nullptr;
        if (A2 == 0 || A3 == 0 || n_cells == 0) fail(CSL_ERR_ARG, "empty lattice");
        const size_t N2 = A2 * n_cells, N3 = A3 * n_cells;
        if (N2 > size_t(1) << 26 || N3 > size_t(1) << 26) fail(CSL_ERR_ARG, "lattice too large");
        if ((off22[N2] && (!par22 || !J22)) || (off33[N3] && (!par33 || !J33)) || (off23[N2] && (!par23 || !J23)) ||
            (off32[N3] && (!par32 || !J32)))
            fail(CSL_ERR_NULL, "null bond table");
        if (off23[N2] != off32[N3]) fail(CSL_ERR_ARG, "SU(2)-SU(3) and SU(3)-SU(2) tables differ in size");
        auto* h = new csl_mixed();
        try {
            hip_check(hipGetDevice(&h->device), "hipGetDevice");
            hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
            hipStream_t s = h->stream;
            h->N2 = N2;
            h->N3 = N3;
            h->A2 = A2;
            h->A3 = A3;
            h->n_cells = n_cells;
            const int npad = int((n_cells + 63) / 64 * 64);
            csl::MixDev& M = h->M;
            M.N2 = int(N2);
            M.N3 = int(N3);
            M.A2 = int(A2);
            M.A3 = int(A3);
            M.n_cells = int(n_cells);
            M.npad = npad;
            M.NP2 = int(A2) * npad;
            M.NP3 = int(A3) * npad;
            const double dummy = 0.0;
            EllHost e22 = build_ell(N2, A2, npad, off22, par22, J22 ? J22 : &dummy, 9, N2, A2);
            EllHost e23 = build_ell(N2, A2, npad, off23, par23, J23 ? J23 : &dummy, 24, N3, A3);
            EllHost e33 = build_ell(N3, A3, npad, off33, par33, J33 ? J33 : &dummy, 64, N3, A3);
            EllHost e32 = build_ell(N3, A3, npad, off32, par32, J32 ? J32 : &dummy, 24, N2, A2);
            h->e22.upload(e22, M.e22, s);
            h->e23.upload(e23, M.e23, s);
            h->e33.upload(e33, M.e33, s);
            h->e32.upload(e32, M.e32, s);
            h->zmax[0] = e22.zmax;
            h->zmax[1] = e23.zmax;
            h->zmax[2] = e33.zmax;
            h->zmax[3] = e32.zmax;
            const std::vector<int> jo = transpose_offsets(e23, e32, A2, A3, int(n_cells), npad);
            M.j32off = nullptr;
            if (!jo.empty()) {
                h->j32off.upload(jo.data(), jo.size(), s);
                M.j32off = h->j32off.p;
            }
            site_tables<3>(N2, A2, npad, field2, onsite2, h->field2, h->onsite2, M.site_uniform2, M.has_onsite2, s);
            site_tables<8>(N3, A3, npad, field3, onsite3, h->field3, h->onsite3, M.site_uniform3, M.has_onsite3, s);
            M.field2 = h->field2.p;
            M.onsite2 = h->onsite2.p;
            M.field3 = h->field3.p;
            M.onsite3 = h->onsite3.p;
            std::vector<double> f2(A2 * 9, 0.0), a2(A2, 1.0), f3(A3 * 64, 0.0);
            for (size_t a = 0; a < A2; ++a)
                for (int d = 0; d < 3; ++d) f2[a * 9 + d * 4] = 1.0;
            for (size_t a = 0; a < A3; ++a)
                for (int d = 0; d < 8; ++d) f3[a * 64 + d * 9] = 1.0;
            h->frames2.upload(f2.data(), f2.size(), s);
            h->afm2.upload(a2.data(), a2.size(), s);
            h->frames3.upload(f3.data(), f3.size(), s);
            std::vector<double> z2(2 * A2 * 3, 0.0), z3(2 * A3 * 8, 0.0);
            h->fdrive2.upload(z2.data(), z2.size(), s);
            h->fdrive3.upload(z3.data(), z3.size(), s);
            M.frames2 = h->frames2.p;
            M.afm2 = h->afm2.p;
            M.frames3 = h->frames3.p;
            M.fdrive2 = h->fdrive2.p;
            M.fdrive3 = h->fdrive3.p;
            M.width2 = M.width3 = 1.0;
            hip_check(hipStreamSynchronize(s), "table upload");
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int csl_mixed_destroy(csl_mixed* h) {
    return guarded([&] {
        if (!h) return;
        (void)hipSetDevice(h->device);
        (void)hipStreamSynchronize(h->stream);
        delete h;
    });
}

int csl_mixed_info(csl_mixed* h, size_t* N2, size_t* N3, int* u22, int* u23, int* u33, int* u32) {
    return guarded([&] {
        check_mixed(h);
        if (N2) *N2 = h->N2;
        if (N3) *N3 = h->N3;
        if (u22) *u22 = h->M.e22.uniform;
        if (u23) *u23 = h->M.e23.uniform;
        if (u33) *u33 = h->M.e33.uniform;
        if (u32) *u32 = h->M.e32.uniform;
    });
}

int csl_mixed_set_frames(csl_mixed* h, const double* frames2, const double* afm2, const double* frames3) {
    return guarded([&] {
        check_mixed(h);
        if (frames2) {
            check_finite(frames2, h->A2 * 9, "frames2");
            h->frames2.upload(frames2, h->A2 * 9, h->stream);
        }
        if (afm2) {
            check_finite(afm2, h->A2, "afm2");
            h->afm2.upload(afm2, h->A2, h->stream);
        }
        if (frames3) {
            check_finite(frames3, h->A3 * 64, "frames3");
            h->frames3.upload(frames3, h->A3 * 64, h->stream);
        }
        hip_check(hipStreamSynchronize(h->stream), "frames");
    });
}

int csl_mixed_set_pulse(csl_mixed* h, int species, const double* drive, double amp, double width, double freq,
                        double t1, double t2) {
    return guarded([&] {
        check_mixed(h);
        if (species != 2 && species != 3) fail(CSL_ERR_ARG, "species must be 2 (SU(2)) or 3 (SU(3))");
        if (!drive) fail(CSL_ERR_NULL, "null drive");
        if (!(width > 0.0) || !std::isfinite(amp) || !std::isfinite(freq) || !std::isfinite(t1) || !std::isfinite(t2))
            fail(CSL_ERR_ARG, "pulse needs finite parameters and width > 0");
        const size_t n = species == 2 ? 2 * h->A2 * 3 : 2 * h->A3 * 8;
        check_finite(drive, n, "drive");
        (species == 2 ? h->fdrive2 : h->fdrive3).upload(drive, n, h->stream);
        if (species == 2) {
            h->M.amp2 = amp;
            h->M.width2 = width;
            h->M.freq2 = freq;
        } else {
            h->M.amp3 = amp;
            h->M.width3 = width;
            h->M.freq3 = freq;
        }
        double* tc = species == 2 ? h->h_tp2 : h->h_tp3;
        tc[0] = t1;
        tc[1] = t2;
        hip_check(hipStreamSynchronize(h->stream), "pulse");
    });
}

int csl_mixed_set_damping(csl_mixed* h, double alpha, const double* gamma8, const double* eq) {
    return guarded([&] {
        check_mixed(h);
        if (!std::isfinite(alpha)) fail(CSL_ERR_ARG, "non-finite alpha");
        h->M.alpha = alpha;
        bool any = false;
        for (int a = 0; a < 8; ++a) {
            h->M.gamma[a] = gamma8 ? gamma8[a] : 0.0;
            if (!std::isfinite(h->M.gamma[a])) fail(CSL_ERR_ARG, "non-finite gamma");
            any |= h->M.gamma[a] != 0.0;
        }
        h->M.has_bloch = any ? 1 : 0;
        std::vector<double> e(size_t(8) * h->M.NP3, 0.0);
        if (eq) {
            check_finite(eq, h->N3 * 8, "eq");
            for (size_t i = 0; i < h->N3; ++i) {
                const size_t p = (i % h->A3) * h->M.npad + i / h->A3;
                for (int a = 0; a < 8; ++a) e[size_t(a) * h->M.NP3 + p] = eq[i * 8 + a];
            }
        }
        h->eq3.upload(e.data(), e.size(), h->stream);
        h->M.eq3 = h->eq3.p;
        hip_check(hipStreamSynchronize(h->stream), "damping");
    });
}

int csl_mixed_set_spins(csl_mixed* h, const double* flat) {
    return guarded([&] {
        check_mixed(h);
        if (!flat) fail(CSL_ERR_NULL, "null spins");
        check_finite(flat, h->nflat(), "spins");
        h->flat.upload(flat, h->nflat(), h->stream);
        h->state.ensure(h->sz());
        hip_check(hipMemsetAsync(h->state.p, 0, h->sz() * sizeof(double), h->stream), "memset");
        to_soa(h, h->flat.p, h->nflat(), h->state.p, 1);
        hip_check(hipStreamSynchronize(h->stream), "set spins");
        h->has_state = true;
    });
}

int csl_mixed_get_spins(csl_mixed* h, double* flat) {
    return guarded([&] {
        check_mixed(h);
        need_state(h);
        if (!flat) fail(CSL_ERR_NULL, "null output");
        h->flat.ensure(h->nflat());
        to_flat(h, h->state.p, h->flat.p, 1);
        copy_back(flat, h->flat.p, h->nflat(), h->stream);
        hip_check(hipStreamSynchronize(h->stream), "get spins");
    });
}

int csl_mixed_record(csl_mixed* h, double* out34) {
    return guarded([&] {
        check_mixed(h);
        need_state(h);
        if (!out34) fail(CSL_ERR_NULL, "null output");
        // a zero-step integration of a copy of the state yields the record
        h->X.ensure(h->sz());
        hip_check(hipMemcpyAsync(h->X.p, h->state.p, h->sz() * sizeof(double), hipMemcpyDeviceToDevice, h->stream),
                  "copy");
        std::vector<double> grid{0.0};
        run_mixed(h, 1, nullptr, nullptr, nullptr, grid, 1.0, 1, nullptr);
        copy_back(out34, h->recs.p, CSL_MIXED_RECORD, h->stream);
        hip_check(hipStreamSynchronize(h->stream), "record");
    });
}

int csl_mixed_energy(csl_mixed* h, double* energy) {
    return guarded([&] {
        if (!energy) fail(CSL_ERR_NULL, "null output");
        double r[CSL_MIXED_RECORD];
        const int rc = csl_mixed_record(h, r);
        if (rc != CSL_OK) fail(rc, g_err);
        *energy = r[33] * double(h->N2 + h->N3);
    });
}

int csl_mixed_rhs(csl_mixed* h, double t, double* dsdt) {
    return guarded([&] {
        check_mixed(h);
        need_state(h);
        if (!dsdt) fail(CSL_ERR_NULL, "null output");
        auto env = [&](double amp, double w, double fr, double c) {
            const double d = t - c;
            const double a = d / (2.0 * w);
            return amp * std::exp(-(a * a)) * std::cos(fr * d);
        };
        const csl::MixDev& M = h->M;
        const double f21 = M.amp2 != 0.0 ? env(M.amp2, M.width2, M.freq2, h->h_tp2[0]) : 0.0;
        const double f22 = M.amp2 != 0.0 ? env(M.amp2, M.width2, M.freq2, h->h_tp2[1]) : 0.0;
        const double f31 = M.amp3 != 0.0 ? env(M.amp3, M.width3, M.freq3, h->h_tp3[0]) : 0.0;
        const double f32 = M.amp3 != 0.0 ? env(M.amp3, M.width3, M.freq3, h->h_tp3[1]) : 0.0;
        h->aux.ensure(h->sz());
        hip_check(hipMemsetAsync(h->aux.p, 0, h->sz() * sizeof(double), h->stream), "memset");
        const int nb = blocks_of(M.NP2) + blocks_of(M.NP3);
        hipLaunchKernelGGL(csl::k_mixed_eval, dim3(nb), dim3(256), 0, h->stream, M, h->state.p, f21, f22, f31, f32,
                           h->aux.p);
        hip_check(hipGetLastError(), "k_mixed_eval");
        h->flat.ensure(h->nflat());
        to_flat(h, h->aux.p, h->flat.p, 1);
        copy_back(dsdt, h->flat.p, h->nflat(), h->stream);
        hip_check(hipStreamSynchronize(h->stream), "rhs");
    });
}

int csl_mixed_normalize(csl_mixed* h, double len2, double len3) {
    return guarded([&] {
        check_mixed(h);
        need_state(h);
        const int n = std::max(h->M.NP2, h->M.NP3);
        hipLaunchKernelGGL(csl::k_mixed_normalize, dim3(blocks_of(n)), dim3(256), 0, h->stream, h->M, h->state.p,
                           len2, len3);
        hip_check(hipGetLastError(), "k_mixed_normalize");
        hip_check(hipStreamSynchronize(h->stream), "normalize");
    });
}

int csl_mixed_integrate(csl_mixed* h, double T0, double T1, double dt, size_t save_interval, const char* method,
                        int tgrid, size_t max_rec, double* times, double* states, double* recs, size_t* n_rec_out) {
    return guarded([&] {
        check_mixed(h);
        need_state(h);
        if (save_interval == 0) fail(CSL_ERR_ARG, "save_interval must be >= 1");
        auto grid = make_grid(T0, T1, dt, tgrid);
        const size_t n_steps = grid.size() - 1;
        const size_t n_rec = records_for(n_steps, save_interval);
        if (n_rec_out) *n_rec_out = n_rec;
        if ((times || states || recs) && max_rec < n_rec) fail(CSL_ERR_CAPACITY, "max_rec is smaller than the number of records");
        hipStream_t s = h->stream;
        h->X.ensure(h->sz());
        hip_check(hipMemcpyAsync(h->X.p, h->state.p, h->sz() * sizeof(double), hipMemcpyDeviceToDevice, s), "copy");
        const double one[2] = {1.0, 1.0};
        h->scale.upload(one, 2, s);
        h->tp.upload(h->h_tp2, 2, s);
        h->tp3.upload(h->h_tp3, 2, s);
        DBuf<double> st;
        if (states) st.ensure(n_rec * h->nflat());
        if (method_is_rk4(method))
            run_mixed(h, 1, h->tp.p, h->tp3.p, h->scale.p, grid, dt, save_interval, states ? st.p : nullptr);
        else
            run_mixed_tableau(h, mixed_method_tableau(method), grid, dt, save_interval, states ? st.p : nullptr);
        hip_check(hipMemcpyAsync(h->state.p, h->X.p, h->sz() * sizeof(double), hipMemcpyDeviceToDevice, s), "copy");
        copy_back(recs, h->recs.p, n_rec * CSL_MIXED_RECORD, s);
        if (states) copy_back(states, st.p, n_rec * h->nflat(), s);
        hip_check(hipStreamSynchronize(s), "integrate");
        if (times)
            for (size_t r = 0; r < n_rec; ++r) times[r] = grid[std::min(r * save_interval, n_steps)];
    });
}

int csl_mixed_step(csl_mixed* h, double t, double dt, const char* method) {
    return guarded([&] {
        check_mixed(h);
        need_state(h);
        if (!(dt > 0.0) || !std::isfinite(t)) fail(CSL_ERR_ARG, "step needs finite t and dt > 0");
        hipStream_t s = h->stream;
        h->X.ensure(h->sz());
        hip_check(hipMemcpyAsync(h->X.p, h->state.p, h->sz() * sizeof(double), hipMemcpyDeviceToDevice, s), "copy");
        const double one[2] = {1.0, 1.0};
        h->scale.upload(one, 2, s);
        h->tp.upload(h->h_tp2, 2, s);
        h->tp3.upload(h->h_tp3, 2, s);
        std::vector<double> grid{t, t + dt};
        if (method_is_rk4(method))
            run_mixed(h, 1, h->tp.p, h->tp3.p, h->scale.p, grid, dt, 1, nullptr);
        else
            run_mixed_tableau(h, mixed_method_tableau(method), grid, dt, 1, nullptr);
        hip_check(hipMemcpyAsync(h->state.p, h->X.p, h->sz() * sizeof(double), hipMemcpyDeviceToDevice, s), "copy");
        hip_check(hipStreamSynchronize(s), "step");
    });
}

int csl_mixed_integrate_batched(csl_mixed* h, size_t B, const double* init, int init_per_replica,
                                const double* t_pulse, const double* pulse_scale, double T0, double T1, double dt,
                                size_t save_interval, const char* method, int tgrid, size_t max_rec, double* times,
                                double* recs, double* final_states, size_t* n_rec_out) {
    return guarded([&] {
        check_mixed(h);
        if (B == 0) fail(CSL_ERR_ARG, "batch size must be >= 1");
        if (B > 65535) fail(CSL_ERR_ARG, "at most 65535 trajectories per call");
        if (!init || !t_pulse) fail(CSL_ERR_NULL, "null init / t_pulse");
        if (!method_is_rk4(method)) fail(CSL_ERR_UNSUPPORTED, "the mixed path implements \"rk4\"");
        if (save_interval == 0) fail(CSL_ERR_ARG, "save_interval must be >= 1");
        auto grid = make_grid(T0, T1, dt, tgrid);
        const size_t n_steps = grid.size() - 1;
        const size_t n_rec = records_for(n_steps, save_interval);
        if (n_rec_out) *n_rec_out = n_rec;
        if ((times || recs) && max_rec < n_rec) fail(CSL_ERR_CAPACITY, "max_rec is smaller than the number of records");
        const size_t nf = h->nflat();
        const size_t n_init = init_per_replica ? B * nf : nf;
        check_finite(init, n_init, "initial state");
        check_finite(t_pulse, 2 * B, "t_pulse");
        hipStream_t s = h->stream;
        h->init.upload(init, n_init, s);
        h->tp.upload(t_pulse, 2 * B, s);
        std::vector<double> sc(2 * B, 1.0);
        if (pulse_scale) {
            check_finite(pulse_scale, 2 * B, "pulse_scale");
            sc.assign(pulse_scale, pulse_scale + 2 * B);
        }
        h->scale.upload(sc.data(), sc.size(), s);
        h->X.ensure(B * h->sz());
        hip_check(hipMemsetAsync(h->X.p, 0, B * h->sz() * sizeof(double), s), "memset");
        to_soa(h, h->init.p, init_per_replica ? nf : 0, h->X.p, B);
        h->last_persist = 0;
        h->last_persist_grid = 0;
        h->last_split = 0;
        h->last_persist_ms = 0.0;
        bool persistent = false;
        if (h->persist_mode >= 0 && !h->stage_graph)
            persistent = run_mixed_persist(h, B, h->tp.p, h->tp.p, h->scale.p, grid, dt, save_interval);
        else if (h->persist_mode == -2 && !h->stage_graph)
            persistent = run_mixed_coop_batch(h, B, h->tp.p, h->tp.p, h->scale.p, grid, dt, save_interval);
        if (!persistent)
            run_mixed(h, B, h->tp.p, h->tp.p, h->scale.p, grid, dt, save_interval, nullptr, false, h->stage_graph != 0);
        copy_back(recs, h->recs.p, n_rec * B * CSL_MIXED_RECORD, s);
        if (final_states) {
            h->flat.ensure(B * nf);
            to_flat(h, h->X.p, h->flat.p, B);
            copy_back(final_states, h->flat.p, B * nf, s);
        }
        hip_check(hipStreamSynchronize(s), "integrate_batched");
        if (times)
            for (size_t r = 0; r < n_rec; ++r) times[r] = grid[std::min(r * save_interval, n_steps)];
    });
}

int csl_mixed_integrate_dopri5(csl_mixed* h, size_t B, const double* init, int init_per_replica,
                               const double* t_pulse, const double* pulse_scale, double T0, double T1, double dt,
                               int control, double abs_tol, double rel_tol, size_t save_interval, int tgrid,
                               const double* segments, const size_t* n_segments, size_t max_segments,
                               size_t max_attempts, size_t max_rec, double* times, double* recs,
                               double* final_states, size_t* n_rec_out, size_t* n_acc_out, size_t* n_rej_out) {
    return guarded([&] {
        check_mixed(h);
        if (B == 0) fail(CSL_ERR_ARG, "batch size must be >= 1");
        if (B > 65535) fail(CSL_ERR_ARG, "at most 65535 trajectories per call");
        if (!init || !t_pulse) fail(CSL_ERR_NULL, "null init / t_pulse");
        if (control < 0 || control > 2) fail(CSL_ERR_ARG, "control must be 0, 1 or 2");
        if (save_interval == 0 || save_interval > size_t(std::numeric_limits<int>::max()))
            fail(CSL_ERR_ARG, "save_interval must be >= 1");
        if (!(dt > 0.0) || !std::isfinite(dt) || !std::isfinite(T0) || !std::isfinite(T1))
            fail(CSL_ERR_ARG, "need finite T0, T1 and dt > 0");
        if (control != 0 && (!(abs_tol >= 0.0) || !(rel_tol >= 0.0) || !(abs_tol + rel_tol > 0.0)))
            fail(CSL_ERR_ARG, "tolerances must be >= 0 and not both 0");
        if (max_rec == 0 || max_rec > size_t(std::numeric_limits<int>::max()))
            fail(CSL_ERR_ARG, "max_rec must be in [1, 2^31)");
        std::vector<double> grid;
        if (control == 0) {
            grid = make_grid(T0, T1, dt, tgrid);
            if (grid.size() > size_t(std::numeric_limits<int>::max())) fail(CSL_ERR_ARG, "too many grid points");
        }
        std::vector<double> sg;
        std::vector<int> nsg;
        if (segments) {
            if (control != 2) fail(CSL_ERR_ARG, "segments apply to control 2");
            if (!n_segments) fail(CSL_ERR_NULL, "null n_segments");
            if (max_segments == 0 || max_segments > 4096) fail(CSL_ERR_ARG, "max_segments must be in [1, 4096]");
            check_finite(segments, B * max_segments * 2, "segments");
            sg.assign(segments, segments + B * max_segments * 2);
            nsg.resize(B);
            for (size_t b = 0; b < B; ++b) {
                if (n_segments[b] == 0 || n_segments[b] > max_segments)
                    fail(CSL_ERR_ARG, "every trajectory needs 1..max_segments segments");
                nsg[b] = int(n_segments[b]);
                for (size_t j = 0; j < n_segments[b]; ++j)
                    if (!(sg[(b * max_segments + j) * 2 + 1] > sg[(b * max_segments + j) * 2]))
                        fail(CSL_ERR_ARG, "a segment must end after it starts");
            }
        }
        const size_t nf = h->nflat();
        const size_t n_init = init_per_replica ? B * nf : nf;
        check_finite(init, n_init, "initial state");
        check_finite(t_pulse, 2 * B, "t_pulse");
        hipStream_t s = h->stream;
        h->init.upload(init, n_init, s);
        h->tp.upload(t_pulse, 2 * B, s);
        std::vector<double> sc(2 * B, 1.0);
        if (pulse_scale) {
            check_finite(pulse_scale, 2 * B, "pulse_scale");
            sc.assign(pulse_scale, pulse_scale + 2 * B);
        }
        h->scale.upload(sc.data(), sc.size(), s);
        if (!grid.empty()) h->dp_grid.upload(grid.data(), grid.size(), s);
        if (!sg.empty()) {
            h->dp_segs.upload(sg.data(), sg.size(), s);
            h->dp_nseg.upload(nsg.data(), nsg.size(), s);
        }
        const size_t plane = B * h->sz();
        for (DBuf<double>* q : {&h->dpX0, &h->dpX1, &h->dpK10, &h->dpK11, &h->dpK2, &h->dpK3, &h->dpK4, &h->dpK5,
                                &h->dpK6, &h->dpY0, &h->dpY1})
            q->ensure(plane);
        // padding entries of the SoA planes stay zero in every buffer
        for (DBuf<double>* q : {&h->dpX0, &h->dpX1, &h->dpY0, &h->dpY1})
            hip_check(hipMemsetAsync(q->p, 0, plane * sizeof(double), s), "memset");
        to_soa(h, h->init.p, init_per_replica ? nf : 0, h->dpX0.p, B);
        const int nblk2 = blocks_of(h->M.NP2), nblk3 = blocks_of(h->M.NP3);
        const int nblk = nblk2 + nblk3;
        h->dp_errp.ensure(B * size_t(nblk));
        h->partials.ensure(B * size_t(nblk) * CSL_MIXED_RECORD);
        h->dp_recs.ensure(B * max_rec * CSL_MIXED_RECORD);
        h->dp_times.ensure(B * max_rec);
        // initial controller state: the derivative at the start with the
        // first observer call
        std::vector<csl::MixDpCtl> ctl(B);
        for (size_t b = 0; b < B; ++b) {
            csl::MixDpCtl& c = ctl[b];
            std::memset(&c, 0, sizeof(c));
            c.s0 = sg.empty() ? T0 : sg[b * max_segments * 2];
            c.s1 = sg.empty() ? T1 : sg[b * max_segments * 2 + 1];
            if (control != 2) c.s0 = T0;
            c.t = c.tcur = c.te = c.s0;
            c.tend = T1;
            c.h = dt;
            c.mode = 2;
            c.eobs = 1;
            c.obs = 1;
            c.fresh = 1;
        }
        h->dp_ctl.upload(ctl.data(), B, s);
        csl::MixDpArgs A{};
        A.nblk2 = nblk2;
        A.X0 = h->dpX0.p;
        A.X1 = h->dpX1.p;
        A.K10 = h->dpK10.p;
        A.K11 = h->dpK11.p;
        A.K[0] = h->dpK2.p;
        A.K[1] = h->dpK3.p;
        A.K[2] = h->dpK4.p;
        A.K[3] = h->dpK5.p;
        A.K[4] = h->dpK6.p;
        A.Y0 = h->dpY0.p;
        A.Y1 = h->dpY1.p;
        A.tp = h->tp.p;
        A.scale = h->scale.p;
        A.abs_tol = abs_tol;
        A.rel_tol = rel_tol;
        A.control = control;
        A.ctl = h->dp_ctl.p;
        A.errp = h->dp_errp.p;
        A.partials = h->partials.p;
        csl::MixDpDecideArgs D{};
        D.nblk = nblk;
        D.control = control;
        D.save_interval = int(save_interval);
        D.n_grid = int(grid.size());
        D.max_seg = int(max_segments);
        D.max_rec = int(max_rec);
        D.max_attempts = max_attempts ? (long long)std::min<size_t>(max_attempts, size_t(1) << 62) : 50000000LL;
        D.T0 = T0;
        D.T1 = T1;
        D.dt = dt;
        D.inv2 = 1.0 / double(h->N2);
        D.inv3 = 1.0 / double(h->N3);
        D.invE = 1.0 / double(h->N2 + h->N3);
        D.tgrid = grid.empty() ? nullptr : h->dp_grid.p;
        D.segs = sg.empty() ? nullptr : h->dp_segs.p;
        D.n_segs = sg.empty() ? nullptr : h->dp_nseg.p;
        D.tp = h->tp.p;
        D.scale = h->scale.p;
        D.amp2 = h->M.amp2;
        D.amp3 = h->M.amp3;
        D.ctl = h->dp_ctl.p;
        D.errp = h->dp_errp.p;
        D.partials = h->partials.p;
        D.recs = h->dp_recs.p;
        D.times = h->dp_times.p;
        const dim3 grid_dim{unsigned(nblk), unsigned(B), 1u};
        // every iteration advances each unfinished trajectory by one attempt
        // or one evaluation; each attempt counts against max_attempts and
        // each evaluation belongs to an observer interval or a segment, so
        // the loop is bounded by the slowest trajectory's work
        const double n_int = std::abs(T1 - T0) / dt;
        const size_t it_cap = size_t(D.max_attempts) + 2 * size_t(std::min(n_int, 1e12)) + 2 * max_segments + 16;
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        hip_check(hipEventRecord(e0, s), "event record");
        std::vector<csl::MixDpCtl> back(B);
        size_t it = 0, launches = 0;
        bool running = true;
        while (running) {
            const size_t batch = it < 64 ? 8 : 64;  // iterations between host checks
            for (size_t k = 0; k < batch; ++k) {
                for (int S = 1; S <= 7; ++S) {
                    A.stage = S;
                    hipLaunchKernelGGL(csl::k_mixdp_stage, grid_dim, dim3(256), 0, s, h->M, A);
                }
                hipLaunchKernelGGL(csl::k_mixdp_decide, dim3(unsigned(B)), dim3(256), 0, s, D);
                launches += 8;
            }
            hip_check(hipGetLastError(), "k_mixdp launch");
            it += batch;
            hip_check(hipMemcpyAsync(back.data(), h->dp_ctl.p, B * sizeof(csl::MixDpCtl), hipMemcpyDeviceToHost, s),
                      "D2H");
            hip_check(hipStreamSynchronize(s), "mixed dopri5");
            running = false;
            for (size_t b = 0; b < B; ++b) running |= back[b].mode != 0;
            if (running && it > it_cap) fail(CSL_ERR_STEP, "the Dormand-Prince iterations did not finish");
        }
        hip_check(hipEventRecord(e1, s), "event record");
        {
            dim3 g(unsigned((h->sz() + 255) / 256), unsigned(B));
            hipLaunchKernelGGL(csl::k_mixdp_final, g, dim3(256), 0, s, h->dp_ctl.p, h->dpX1.p, h->dpX0.p, h->sz());
            hip_check(hipGetLastError(), "k_mixdp_final");
        }
        copy_back(times, h->dp_times.p, B * max_rec, s);
        copy_back(recs, h->dp_recs.p, B * max_rec * CSL_MIXED_RECORD, s);
        if (final_states) {
            h->flat.ensure(B * nf);
            to_flat(h, h->dpX0.p, h->flat.p, B);
            copy_back(final_states, h->flat.p, B * nf, s);
        }
        hip_check(hipStreamSynchronize(s), "mixed dopri5");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        h->last_ms = launches ? ms / double(launches) : 0.0;
        h->last_launches = launches;
        bool over = false;
        for (size_t b = 0; b < B; ++b) {
            const csl::MixDpCtl& c = back[b];
            if (n_rec_out) n_rec_out[b] = size_t(c.saved);
            if (n_acc_out) n_acc_out[b] = size_t(c.n_acc);
            if (n_rej_out) n_rej_out[b] = size_t(c.n_rej);
            over |= size_t(c.saved) > max_rec;
        }
        for (size_t b = 0; b < B; ++b) {
            const int st = back[b].status;
            if (st == 1) fail(CSL_ERR_STEP, "trajectory " + std::to_string(b) + ": step size control failed (500 rejections)");
            if (st == 2) fail(CSL_ERR_STEP, "trajectory " + std::to_string(b) + ": more than max_attempts attempts");
            if (st == 3) fail(CSL_ERR_STEP, "trajectory " + std::to_string(b) + ": step size no longer advances time");
        }
        if (over) fail(CSL_ERR_CAPACITY, "max_rec is smaller than the number of records of some trajectory");
    });
}

int csl_mixed_scale_gradient(csl_mixed* h, size_t B, const double* init, int init_per_replica,
                             const double* t_pulse, const double* pulse_scale, double T0, double T1, double dt,
                             size_t save_interval, int tgrid, const double* target, int target_per_replica,
                             const double* weights, double* loss, double* grad, size_t max_rec, double* recs,
                             double* trecs, size_t* n_rec_out) {
    return guarded([&] {
        check_mixed(h);
        if (B == 0) fail(CSL_ERR_ARG, "batch size must be >= 1");
        if (B > 65535) fail(CSL_ERR_ARG, "at most 65535 trajectories per call");
        if (!init || !t_pulse || !target || !weights || !loss || !grad)
            fail(CSL_ERR_NULL, "null init / t_pulse / target / weights / loss / grad");
        if (save_interval == 0) fail(CSL_ERR_ARG, "save_interval must be >= 1");
        if (weights[CSL_MIXED_RECORD - 1] != 0.0)
            fail(CSL_ERR_ARG, "the energy entry of the record cannot carry a loss weight");
        check_finite(weights, CSL_MIXED_RECORD, "weights");
        auto grid = make_grid(T0, T1, dt, tgrid);
        const size_t n_steps = grid.size() - 1;
        const size_t n_rec = records_for(n_steps, save_interval);
        if (n_rec_out) *n_rec_out = n_rec;
        if ((recs || trecs) && max_rec < n_rec) fail(CSL_ERR_CAPACITY, "max_rec is smaller than the number of records");
        if (n_rec > size_t(std::numeric_limits<int>::max() / CSL_MIXED_RECORD)) fail(CSL_ERR_ARG, "too many records");
        const size_t nf = h->nflat();
        const size_t n_init = init_per_replica ? B * nf : nf;
        check_finite(init, n_init, "initial state");
        check_finite(t_pulse, 2 * B, "t_pulse");
        const size_t n_tgt = (target_per_replica ? B : 1) * n_rec * CSL_MIXED_RECORD;
        check_finite(target, n_tgt, "target");
        hipStream_t s = h->stream;
        h->init.upload(init, n_init, s);
        h->tp.upload(t_pulse, 2 * B, s);
        std::vector<double> sc(2 * B, 1.0);
        if (pulse_scale) {
            check_finite(pulse_scale, 2 * B, "pulse_scale");
            sc.assign(pulse_scale, pulse_scale + 2 * B);
        }
        h->scale.upload(sc.data(), sc.size(), s);
        h->target.upload(target, n_tgt, s);
        h->X.ensure(B * h->sz());
        h->ZX.ensure(B * h->sz());
        hip_check(hipMemsetAsync(h->X.p, 0, B * h->sz() * sizeof(double), s), "memset");
        hip_check(hipMemsetAsync(h->ZX.p, 0, B * h->sz() * sizeof(double), s), "memset");
        to_soa(h, h->init.p, init_per_replica ? nf : 0, h->X.p, B);
        run_mixed(h, B, h->tp.p, h->tp.p, h->scale.p, grid, dt, save_interval, nullptr, true);
        csl::Weights34 W{};
        std::memcpy(W.w, weights, sizeof(W.w));
        h->loss.ensure(B);
        h->grad.ensure(B);
        hipLaunchKernelGGL(csl::k_mixed_loss_grad, dim3(unsigned(B)), dim3(256), 0, s, h->recs.p, h->trecs.p,
                           h->target.p, target_per_replica ? 1 : 0, int(n_rec), int(B), W, h->loss.p, h->grad.p);
        hip_check(hipGetLastError(), "k_mixed_loss_grad");
        copy_back(loss, h->loss.p, B, s);
        copy_back(grad, h->grad.p, B, s);
        copy_back(recs, h->recs.p, n_rec * B * CSL_MIXED_RECORD, s);
        copy_back(trecs, h->trecs.p, n_rec * B * CSL_MIXED_RECORD, s);
        hip_check(hipStreamSynchronize(s), "scale_gradient");
    });
}

int csl_mixed_set_stage_graph(csl_mixed* h, int enable) {
    return guarded([&] {
        check_mixed(h);
        h->stage_graph = enable ? 1 : 0;
    });
}

int csl_mixed_set_persistent(csl_mixed* h, int mode) {
    return guarded([&] {
        check_mixed(h);
        if (mode < -2) fail(CSL_ERR_ARG, "mode must be -2, -1, 0 or a positive workgroup cap");
        h->persist_mode = mode;
    });
}

int csl_mixed_last_path(csl_mixed* h, int* persistent, int* workgroups, size_t* split, double* kernel_ms) {
    return guarded([&] {
        check_mixed(h);
        if (persistent) *persistent = h->last_persist;
        if (workgroups) *workgroups = h->last_persist_grid;
        if (split) *split = h->last_split;
        if (kernel_ms) *kernel_ms = h->last_persist ? h->last_persist_ms : 0.0;
    });
}

int csl_mixed_last_kernel_ms(csl_mixed* h, double* ms, size_t* launches) {
    return guarded([&] {
        if (!h) fail(CSL_ERR_NULL, "null mixed-lattice handle");
        if (ms) *ms = h->last_ms;
        if (launches) *launches = h->last_launches;
    });
}

}  // extern "C"
