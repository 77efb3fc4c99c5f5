// csl_internal.hpp — device-side lattice description and the per-site
// Landau–Lifshitz arithmetic shared by every kernel of libcsl_hip.so.
//
// Device layout (chosen for CDNA4 wave64 execution):
//   * Sites are renumbered sublattice-major:  p = atom * npad + cell, where
//     cell = ((i*d2 + j)*d3 + k) and npad = n_cells rounded up to a multiple
//     of 64.  Every wavefront therefore covers 64 sites of ONE sublattice, so
//     for translation-invariant Hamiltonians the bond matrices, field, onsite
//     matrix, frame and drive direction of a wave are wave-uniform and are
//     read with scalar loads into SGPRs instead of per-lane gathers.
//     Padding slots (cell >= n_cells) carry no spin and are never referenced
//     by a bond.
//   * Spin state is structure-of-arrays [3][NP] (x, y, z planes), FP64.
//   * Bonds are ELL, slot-major: col[k*NP + p] is the device index of the
//     k-th partner of site p, in the reference's bond order
//     (lattice.h:785-803), so sums run in the same order as the reference.
//   * "uniform" handles store one D*D matrix per (sublattice, slot)
//     (mats[(atom*zmax + k)*9]); general handles store one matrix per bond
//     (mats[(k*NP + p)*9]).
#pragma once

#include <hip/hip_runtime.h>

namespace csl {

constexpr int kWave = 64;

// Read-only view of a table in the AMDGPU constant address space (4).  The
// tables a kernel reads this way are never written while it runs, so the
// backend may serve wave-uniform reads with scalar loads (SGPR destinations,
// scalar cache) instead of per-lane vector loads.
template <class T>
using cptr = const T __attribute__((address_space(4)))*;

template <class T>
__device__ __forceinline__ cptr<T> as_const(const T* p) {
    return (cptr<T>)(p);
}

// Force a value the compiler cannot prove wave-uniform into an SGPR.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// A per-lane index the compiler must treat as changed at this point: the
// LDS and global addresses derived from it are recomputed where they are
// used (a shift and an add) instead of being hoisted out of the time loop
// and held in VGPRs (64-bit global addresses of every stage slot) for the
// whole kernel, which spilled them to scratch.
__device__ __forceinline__ int fresh(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

struct LatDev {
    int N;          // real sites
    int n_atoms;
    int n_cells;
    int npad;       // cells per sublattice block, multiple of 64
    int NP;         // n_atoms * npad
    int zmax;       // max bonds per site
    int uniform;    // bond matrices per (sublattice, slot)
    int site_uniform;  // field / onsite per sublattice
    int has_onsite;    // any non-zero onsite matrix (the tables hold 2A)
    int full_coord;    // uniform tables and every site has zmax bonds
    const int* __restrict__ col;        // [zmax*NP]
    const int* __restrict__ cnt;        // [NP] bonds of each site
    const int* __restrict__ cnt_atom;   // [n_atoms] (uniform handles)
    const double* __restrict__ mats;    // see header comment
    const double* __restrict__ field;   // [n_atoms*3] or [NP*3]
    const double* __restrict__ onsite;  // 2A: [n_atoms*9] or [NP*9]
    const double* __restrict__ frames;  // [n_atoms*9]
    const double* __restrict__ afm;     // [n_atoms]
    const double* __restrict__ fdrive;  // [2*n_atoms*3]
    double width, freq, alpha;
};

// Pulse envelope of one pulse, lattice_md.cpp:45-56:
//   amp * exp(-((t - t_p) / (2 w))^2) * cos(freq * (t - t_p))
// Kept out of line: the FP64 exp/cos (with Payne-Hanek argument reduction)
// needs ~40 VGPRs, and inlining it into the persistent kernel would raise
// that kernel's register count for the whole time loop.
__device__ __attribute__((noinline)) double drive_factor(double amp, double t, double tp, double width,
                                               double freq) {
    const double d = t - tp;
    const double a = d / (2.0 * width);
    return amp * exp(-(a * a)) * cos(freq * d);
}

// Matrix of bond slot k of site p (J applied as H += J * S_partner).
__device__ __forceinline__ const double* bond_matrix(const LatDev& L, int atom, int p, int k) {
    return L.uniform ? L.mats + (size_t(atom) * L.zmax + k) * 9
                     : L.mats + (size_t(k) * L.NP + p) * 9;
}

// H = -B + 2 A S  (lattice.h:1655-1666).  The tables hold 2A (doubling is
// exact, so (2A)*S equals the reference's 2.0*A*S); handles without onsite
// terms skip the product.
__device__ __forceinline__ void field_onsite(const LatDev& L, int atom, int p, const double S[3],
                                             double H[3]) {
    if (L.site_uniform) {
        cptr<double> B = as_const(L.field) + atom * 3;
        H[0] = -B[0];
        H[1] = -B[1];
        H[2] = -B[2];
        if (!L.has_onsite) return;
        cptr<double> A2 = as_const(L.onsite) + atom * 9;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
#pragma unroll
            for (int e = 0; e < 3; ++e) H[d] += A2[d * 3 + e] * S[e];
        }
        return;
    }
    const double* B = L.field + size_t(p) * 3;
    const double* A2 = L.onsite + size_t(p) * 9;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double h = -B[d];
        if (L.has_onsite) {
#pragma unroll
            for (int e = 0; e < 3; ++e) h += A2[d * 3 + e] * S[e];
        }
        H[d] = h;
    }
}

// H += J * S_partner, with the row sum formed first (lattice.h:1703-1709)
template <class P>
__device__ __forceinline__ void add_bond(P J, const double Sp[3], double H[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double row = 0.0;
#pragma unroll
        for (int e = 0; e < 3; ++e) row += J[d * 3 + e] * Sp[e];
        H[d] += row;
    }
}

// The same bond term accumulated straight into H (one FMA chain per
// component, no per-row partial sum): fewer FP64 instructions for the
// LDS-resident kernels; the sum is associated differently from add_bond
// (rounding-level difference).
template <class P>
__device__ __forceinline__ void add_bond_fused(P J, const double Sp[3], double H[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int e = 0; e < 3; ++e) H[d] = fma(J[d * 3 + e], Sp[e], H[d]);
}

// dS/dt = H x S, plus Gilbert damping -alpha/|S| S x (H x S)
// (lattice_md.cpp:80-101)
__device__ __forceinline__ void ll_su2(const double H[3], const double S[3], double alpha,
                                       double k[3]) {
    double dx = H[1] * S[2] - H[2] * S[1];
    double dy = H[2] * S[0] - H[0] * S[2];
    double dz = H[0] * S[1] - H[1] * S[0];
    if (alpha > 0.0) {
        const double a = S[1] * dz - S[2] * dy;
        const double b = S[2] * dx - S[0] * dz;
        const double c = S[0] * dy - S[1] * dx;
        const double inv_S = 1.0 / sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
        dx -= alpha * inv_S * a;
        dy -= alpha * inv_S * b;
        dz -= alpha * inv_S * c;
    }
    k[0] = dx;
    k[1] = dy;
    k[2] = dz;
}

// 1/sqrt(x) for x > 0: the hardware reciprocal square root refined by two
// Newton-Raphson steps r <- r + r (1 - x r^2) / 2 (each step roughly doubles
// the number of correct bits), within about an ulp of 1.0 / sqrt(x) at a
// fraction of its cost (the correctly rounded FP64 square root and division
// expand to some thirty instructions).
__device__ __forceinline__ double rsqrt_nr(double x) {
    double r = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double e = fma(-x * r, r, 1.0);
        r = fma(0.5 * r, e, r);
    }
    return r;
}

// ll_su2 with the Gilbert factor 1/|S| from rsqrt_nr (the adjoint kernels:
// their forward map and its transpose use the same factor)
__device__ __forceinline__ void ll_su2_rs(const double H[3], const double S[3], double alpha, double k[3]) {
    double dx = H[1] * S[2] - H[2] * S[1];
    double dy = H[2] * S[0] - H[0] * S[2];
    double dz = H[0] * S[1] - H[1] * S[0];
    if (alpha > 0.0) {
        const double a = S[1] * dz - S[2] * dy;
        const double b = S[2] * dx - S[0] * dz;
        const double c = S[0] * dy - S[1] * dx;
        const double inv_S = rsqrt_nr(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
        dx -= alpha * inv_S * a;
        dy -= alpha * inv_S * b;
        dz -= alpha * inv_S * c;
    }
    k[0] = dx;
    k[1] = dy;
    k[2] = dz;
}

// ll_su2 for a site that may be a padding slot with a zero spin: the Gilbert
// factor 1/|S| is taken as 0 there instead of inf (same bits for |S| > 0)
__device__ __forceinline__ void ll_su2_guarded(const double H[3], const double S[3], double alpha, double k[3]) {
    double dx = H[1] * S[2] - H[2] * S[1];
    double dy = H[2] * S[0] - H[0] * S[2];
    double dz = H[0] * S[1] - H[1] * S[0];
    if (alpha > 0.0) {
        const double a = S[1] * dz - S[2] * dy;
        const double b = S[2] * dx - S[0] * dz;
        const double c = S[0] * dy - S[1] * dx;
        const double n2 = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
        const double inv_S = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
        dx -= alpha * inv_S * a;
        dy -= alpha * inv_S * b;
        dz -= alpha * inv_S * c;
    }
    k[0] = dx;
    k[1] = dy;
    k[2] = dz;
}

// Sum of a double over the 64 lanes of a wavefront (result valid in every lane).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

}  // namespace csl
