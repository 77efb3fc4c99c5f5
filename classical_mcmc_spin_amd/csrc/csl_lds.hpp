// csl_lds.hpp — LDS-resident state of one trajectory and wavefront
// reductions, shared by the one-workgroup-per-trajectory kernels
// (csl_persistent.hip, csl_dopri5.hip).
//
// LDS state layout: the x,y components of every site packed as a 16-byte
// pair, then the z plane, so a neighbour gather is one ds_read_b128 plus one
// ds_read_b64:
//     [ (x,y)[0..NP) | z[0..NP) ]          3*NP doubles per buffer
#pragma once

#include "csl_device.hpp"

namespace csl {

namespace {

// Spin state of one trajectory in LDS (see file header).
struct LdsState {
    double2* xy;
    double* z;
    __device__ __forceinline__ void load(int q, double S[3]) const {
        const double2 v = xy[q];
        S[0] = v.x;
        S[1] = v.y;
        S[2] = z[q];
    }
    __device__ __forceinline__ void store(int q, const double S[3]) const {
        xy[q] = make_double2(S[0], S[1]);
        z[q] = S[2];
    }
};

__device__ __forceinline__ LdsState lds_state(double* base, int NP) {
    return LdsState{reinterpret_cast<double2*>(base), base + 2 * NP};
}

// Local field of site p (no drive) gathering neighbours from LDS; with
// FIELD_ONLY the on-site part is the sublattice field alone and every site
// has ZMAX bonds.
template <int ZMAX, bool UNI, bool FIELD_ONLY = false>
__device__ __forceinline__ void local_field_lds(const LatDev& L, const LdsState& y, int p, int atom,
                                                const int* nb, int cnt, const double S[3], double H[3]) {
    if (FIELD_ONLY) {
        cptr<double> B = as_const(L.field) + atom * 3;
        H[0] = -B[0];
        H[1] = -B[1];
        H[2] = -B[2];
    } else {
        field_onsite(L, atom, p, S, H);
    }
#pragma unroll
    for (int k = 0; k < ZMAX; ++k) {
        if (FIELD_ONLY || k < cnt) {  // FIELD_ONLY tables have zmax bonds at every site
            double Sp[3];
            y.load(nb[k], Sp);
            if (UNI)
                add_bond_fused(as_const(L.mats) + (atom * L.zmax + k) * 9, Sp, H);
            else
                add_bond_fused(L.mats + (size_t(k) * L.NP + p) * 9, Sp, H);
        }
    }
}

// One DPP move of a double across lanes (both 32-bit halves with the same
// control); lanes outside row_mask receive 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int rlo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xf, false);
    const int rhi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(rhi, rlo);
}

// Sum over the 64 lanes of a wavefront with DPP row operations (no LDS
// round trip); the total is valid in lane 63.  Must be called with all lanes
// active.
__device__ __forceinline__ double wave_sum_lane63(double v) {
    v += dpp_f64<0xB1, 0xf>(v);   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E, 0xf>(v);   // quad_perm [2,3,0,1]   -> quad sums
    v += dpp_f64<0x124, 0xf>(v);  // row_ror:4
    v += dpp_f64<0x128, 0xf>(v);  // row_ror:8             -> row (16-lane) sums
    v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3 -> total in row 3
    return v;
}

}  // namespace

}  // namespace csl
