// csl_dssf.hip — dynamical structure factor of batched trajectories.
//
//   S_a(q, t) = sum_i exp(-i q . r_i) S_ia(t),   q = 2 pi (h/L1, k/L2, l/L3)
//   S(q, w)   = 1/(B N T) sum_b sum_a |sum_t exp(-i w t) S_a(q, t)|^2
//
// r_i = cell vector (i, j, k) + the atom's position in the cell (in units of
// the cell's lattice vectors, which are orthonormal for the cubic
// pyrochlore cell).  Per sample step the spins of every (trajectory, atom,
// component) are expanded into a complex 3-D grid over the cells, a batched
// 3-D FFT (hipFFT) gives F_a,atom(q), and k_sq_combine sums the atoms with
// their position phases into the S(q, t) buffer; after the run a batched
// 1-D FFT along t and k_sq_power give S(q, w).  The reference post-processes
// saved spin snapshots in Python (util/readers_new/reader_pyrochlore.py:150-270).
#include <hip/hip_runtime.h>

#include "csl_device.hpp"

namespace csl {

// SoA state [B][3][NP] -> complex grids [B][n_atoms][3][n_cells]
__global__ void k_sq_fill(LatDev L, const double* __restrict__ x, int B, double2* __restrict__ out) {
    const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t C = size_t(L.n_cells);
    const size_t total = size_t(B) * L.n_atoms * 3 * C;
    if (id >= total) return;
    const size_t cell = id % C;
    const size_t r = id / C;           // (b, atom, comp)
    const int comp = int(r % 3);
    const size_t ba = r / 3;
    const int atom = int(ba % L.n_atoms);
    const int b = int(ba / L.n_atoms);
    const size_t p = size_t(atom) * L.npad + cell;
    out[id] = make_double2(x[(size_t(b) * 3 + comp) * L.NP + p], 0.0);
}

// F [B][n_atoms][3][Q] -> S(q, t) slot [B][3][Q]: sum_atom exp(-i q.r_atom) F
__global__ void k_sq_combine(const double2* __restrict__ F, int B, int n_atoms, int L1, int L2, int L3,
                             const double* __restrict__ pos, double2* __restrict__ sqt) {
    const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t Q = size_t(L1) * L2 * L3;
    const size_t total = size_t(B) * 3 * Q;
    if (id >= total) return;
    const size_t q = id % Q;
    const size_t r = id / Q;           // (b, comp)
    const int comp = int(r % 3);
    const int b = int(r / 3);
    const int h = int(q / (size_t(L2) * L3));
    const int k = int((q / L3) % L2);
    const int l = int(q % L3);
    const double two_pi = 6.283185307179586;
    const double qx = two_pi * h / L1, qy = two_pi * k / L2, qz = two_pi * l / L3;
    double re = 0.0, im = 0.0;
    for (int a = 0; a < n_atoms; ++a) {
        const double ph = -(qx * pos[3 * a] + qy * pos[3 * a + 1] + qz * pos[3 * a + 2]);
        double s, c;
        sincos(ph, &s, &c);
        const double2 f = F[((size_t(b) * n_atoms + a) * 3 + comp) * Q + q];
        re += c * f.x - s * f.y;
        im += c * f.y + s * f.x;
    }
    sqt[id] = make_double2(re, im);
}

// S(q, w) [Q][T] = scale * sum_{b, comp} |S~[t=w][b][comp][q]|^2
__global__ void k_sq_power(const double2* __restrict__ sw, int T, int B, size_t Q, double scale,
                           double* __restrict__ out) {
    const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= Q * size_t(T)) return;
    const size_t q = id / T;
    const int w = int(id % T);
    double acc = 0.0;
    for (int r = 0; r < B * 3; ++r) {
        const double2 v = sw[(size_t(w) * B * 3 + r) * Q + q];
        acc += v.x * v.x + v.y * v.y;
    }
    out[id] = acc * scale;
}

hipError_t launch_sq_fill(const LatDev& L, const double* x, int B, double2* out, hipStream_t st) {
    const size_t total = size_t(B) * L.n_atoms * 3 * L.n_cells;
    hipLaunchKernelGGL(k_sq_fill, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, L, x, B, out);
    return hipGetLastError();
}

hipError_t launch_sq_combine(const double2* F, int B, int n_atoms, int L1, int L2, int L3, const double* pos,
                             double2* sqt, hipStream_t st) {
    const size_t total = size_t(B) * 3 * L1 * L2 * L3;
    hipLaunchKernelGGL(k_sq_combine, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, F, B, n_atoms, L1, L2,
                       L3, pos, sqt);
    return hipGetLastError();
}

hipError_t launch_sq_power(const double2* sw, int T, int B, size_t Q, double scale, double* out, hipStream_t st) {
    const size_t total = Q * size_t(T);
    hipLaunchKernelGGL(k_sq_power, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, sw, T, B, Q, scale, out);
    return hipGetLastError();
}

}  // namespace csl
