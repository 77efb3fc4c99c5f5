// csl_mixed_device.hpp — device description of a mixed SU(2) + SU(3)
// lattice and the per-site arithmetic shared by the mixed kernels
// (grid-wide stage kernels in csl_mixed.hip, the persistent kernel in
// csl_mixed_persist.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/csl_mixed.h"
#include "csl_internal.hpp"

namespace csl {

// One bond kind on the device (source species gathers partner species).
struct Ell {
    const int* __restrict__ col;        // [zmax][NP_src] partner device index
    const int* __restrict__ cnt;        // [NP_src]
    const int* __restrict__ cnt_atom;   // [A_src]
    const double* __restrict__ mats;    // uniform: [A][zmax][m]; else [zmax][NP_src][m]
    int zmax, uniform;
};

struct MixDev {
    int N2, N3, A2, A3, n_cells, npad, NP2, NP3;
    int site_uniform2, site_uniform3, has_onsite2, has_onsite3;
    const double* __restrict__ field2;   // [A2][3] or [NP2][3]
    const double* __restrict__ onsite2;  // 2A: [A2][9] or [NP2][9]
    const double* __restrict__ field3;   // [A3][8] or [NP3][8]
    const double* __restrict__ onsite3;  // 2A: [A3][64] or [NP3][64]
    Ell e22, e23, e33, e32;
    const double* __restrict__ frames2;  // [A2][9]
    const double* __restrict__ afm2;     // [A2]
    const double* __restrict__ frames3;  // [A3][64]
    const double* __restrict__ fdrive2;  // [2][A2][3]
    const double* __restrict__ fdrive3;  // [2][A3][8]
    const double* __restrict__ eq3;      // [8][NP3] or null
    double amp2, width2, freq2, amp3, width3, freq3;
    double alpha;
    double gamma[8];
    int has_bloch;
    // [A3][8]: offset (doubles) in e23.mats of the SU(2)->SU(3) matrix whose
    // transpose is the SU(3)->SU(2) matrix of each (sublattice, slot), or
    // null when the tables do not pair up that way (the persistent kernel
    // then reads e32.mats)
    const int* __restrict__ j32off;
};

constexpr int kRec = CSL_MIXED_RECORD;

// d_a = f_abc H_b S_c over the 9 non-zero (a<b<c) structure constants
// (cross_prod_SU3_flat, simple_linear_alg.h:78-169)
__device__ __forceinline__ void su3_cross(const double* H, const double* S, double* d) {
    constexpr double h = 0.5;
    constexpr double r = 0.8660254037844386;  // sqrt(3)/2
#pragma unroll
    for (int i = 0; i < 8; ++i) d[i] = 0.0;
    auto tri = [&](int a, int b, int c, double v) {
        d[a] += v * (H[b] * S[c] - H[c] * S[b]);
        d[b] += v * (H[c] * S[a] - H[a] * S[c]);
        d[c] += v * (H[a] * S[b] - H[b] * S[a]);
    };
    tri(0, 1, 2, 1.0);
    tri(0, 3, 6, h);
    tri(0, 4, 5, -h);
    tri(1, 3, 5, h);
    tri(1, 4, 6, h);
    tri(2, 3, 4, h);
    tri(2, 5, 6, -h);
    tri(3, 4, 7, r);
    tri(5, 6, 7, r);
}

__device__ __forceinline__ double envelope(double amp, double t, double tc, double width, double freq) {
    const double d = t - tc;
    const double a = d / (2.0 * width);
    return amp * exp(-(a * a)) * cos(freq * d);
}

// Arguments of k_mixed_persist: trajectories b0 .. b0+B-1 of a batch, one
// workgroup per trajectory at a time (workgroup w runs b0+w, b0+w+grid, ...).
struct MixPersistArgs {
    int b0, B;                 // trajectories of this launch
    int B_rec;                 // batch size: row stride of tp / tp3 / scale / records
    int n_steps, save_interval, n_rec;
    double dt;
    const double* tgrid;       // [n_steps + 1] step start times
    double* X;                 // [B_rec][3 NP2 + 8 NP3] SoA states: initial in, final out; x during the run
    double* G;                 // [grid][2 parities][2][NP3] SU(3) components 6, 7 of the stage values
    const double* tp;          // [B_rec][2] SU(2) pulse centres, or null: no drive
    const double* tp3;         // [B_rec][2] SU(3) pulse centres
    const double* scale;       // [B_rec][2] pulse scales, or null (1)
    double* recs;              // [n_rec][B_rec][34]
    double inv2, inv3, invE;   // record normalisations (1/N2, 1/N3, 1/(N2+N3))
    // work split of more trajectories than workgroups (null: workgroup w
    // runs trajectories w, w + grid, ...): rows [grid][max_pieces][5] of
    // (trajectory, first step, end step, flag to wait for, flag to set),
    // unused rows -1; a piece that continues a trajectory starts from the
    // state the previous piece left in X
    const int* sched;
    int max_pieces;
    unsigned* flags;           // [grid] set once a trajectory's head piece has stored its state
    unsigned* status;          // set when a wait for a flag gives up
};

// Bonds per site of each kind the persistent kernel unrolls (larger tables
// stay on the stage kernels).
constexpr int kMixZ22 = 8, kMixZ23 = 8, kMixZ32 = 8, kMixZ33 = 4;

// Launch shape of the persistent kernel for a lattice: block size and sites
// of each species per thread; block 0 when it does not qualify.  TT: the
// lattice has SU(3)-SU(3) bonds.
struct MixPersistShape {
    int block, spt;
    bool tt;
};
MixPersistShape mixed_persist_shape(int NP2, int NP3, bool tt);
size_t mixed_persist_lds_bytes(int NP2, int NP3, int block);
// Mh: the lattice description on the host, M: its copy in device memory.
// With occ non-null: the co-resident workgroups per CU of that instance
// (nothing is launched).
hipError_t launch_mixed_persist(const MixDev& Mh, const MixDev* M, const MixPersistArgs& A,
                                const MixPersistShape& shape, int grid, hipStream_t st, int* occ);

}  // namespace csl
