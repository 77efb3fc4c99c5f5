// csl_quench.hip — zero-temperature relaxation of a batch of spin states on
// the device (ground-state preparation for the dynamics runs).
//
// Lattice::greedy_quench (lattice.h:3352-3368) repeats deterministic sweeps
// (lattice.h:3334-3349: every spin set antiparallel to its local field,
// S_i = -H_i/|H_i| * spin_length, skipped when |H_i| < 1e-15) until the total
// energy changes by at most rel_tol * (|E_prev| + 1e-18).  The reference
// visits sites in a random order; here a sweep updates the sites colour class
// by colour class of a greedy colouring of the bond graph (no two bonded sites
// share a colour, so all sites of a class update independently and the sweep
// is deterministic).
#include <hip/hip_runtime.h>

#include "csl_device.hpp"

namespace csl {

// one colour class of one sweep for every non-converged trajectory
template <int ZMAX, bool UNI>
__global__ void __launch_bounds__(256) k_quench_class(LatDev L, double* __restrict__ x, const int* __restrict__ color,
                                                      int c, const int* __restrict__ frozen, double len) {
    const int b = blockIdx.y;
    if (frozen[b]) return;
    const int NP = L.NP;
    const int tid = threadIdx.x;
    const int p = blockIdx.x * blockDim.x + tid;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int pw = blockIdx.x * blockDim.x + wave * kWave;
    const int atom = uniform(pw < NP ? pw / L.npad : 0);
    if (!active_site(L, p) || color[p] != c) return;
    double* y = x + size_t(b) * 3 * NP;
    const int cn = UNI ? L.cnt_atom[atom] : L.cnt[p];
    int nb[ZMAX];
#pragma unroll
    for (int k = 0; k < ZMAX; ++k) nb[k] = k < cn ? L.col[size_t(k) * NP + p] : 0;
    const double S[3] = {y[p], y[NP + p], y[2 * NP + p]};
    double H[3];
    local_field_soa<ZMAX, UNI>(L, y, NP, p, atom, nb, cn, S, H);
    const double n = sqrt(H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
    if (n < 1e-15) return;
    const double s = -len / n;
    y[p] = H[0] * s;
    y[NP + p] = H[1] * s;
    y[2 * NP + p] = H[2] * s;
}

// convergence test per trajectory after a sweep (energies E/N in e[])
__global__ void k_quench_check(const double* __restrict__ e, double* __restrict__ e_prev, int* __restrict__ frozen,
                               int* __restrict__ sweeps, int B, int sweep, double rel_tol) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || frozen[b]) return;
    const double E = e[b], Ep = e_prev[b];
    sweeps[b] = sweep + 1;
    if (fabs(E - Ep) <= rel_tol * (fabs(Ep) + 1e-18)) frozen[b] = 1;
    e_prev[b] = E;
}

namespace {
template <int ZMAX>
hipError_t launch_qc(const LatDev& L, double* x, const int* color, int c, const int* frozen, double len, int B,
                     hipStream_t st) {
    const dim3 g(unsigned((L.NP + 255) / 256), unsigned(B));
    if (L.uniform)
        hipLaunchKernelGGL((k_quench_class<ZMAX, true>), g, dim3(256), 0, st, L, x, color, c, frozen, len);
    else
        hipLaunchKernelGGL((k_quench_class<ZMAX, false>), g, dim3(256), 0, st, L, x, color, c, frozen, len);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_quench_class(const LatDev& L, double* x, const int* color, int c, const int* frozen, double len,
                               int B, hipStream_t st) {
    if (L.zmax <= 3) return launch_qc<3>(L, x, color, c, frozen, len, B, st);
    if (L.zmax <= 6) return launch_qc<6>(L, x, color, c, frozen, len, B, st);
    if (L.zmax <= 12) return launch_qc<12>(L, x, color, c, frozen, len, B, st);
    return hipErrorInvalidValue;
}

hipError_t launch_quench_check(const double* e, double* e_prev, int* frozen, int* sweeps, int B, int sweep,
                               double rel_tol, hipStream_t st) {
    hipLaunchKernelGGL(k_quench_check, dim3(unsigned((B + 255) / 256)), dim3(256), 0, st, e, e_prev, frozen, sweeps, B,
                       sweep, rel_tol);
    return hipGetLastError();
}

}  // namespace csl
