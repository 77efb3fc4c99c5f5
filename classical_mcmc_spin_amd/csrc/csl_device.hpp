// csl_device.hpp — device helpers shared by the kernels of csl_kernels.hip
// and csl_persistent.hip: local field from SoA state planes, observables,
// drive subtraction.
#pragma once

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

}  // namespace csl
