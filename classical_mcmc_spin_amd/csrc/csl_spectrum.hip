// csl_spectrum.hip — 2-D nonlinear spectrum of a pump-probe scan on the GPU.
//
// Follows util/readers_new/reader_pyrochlore.py:760-830:
//   M_NL[tau][t] = M01[tau][t] - M0[t] - M1[tau][t] over the last
//                  domain = n_t / 2 samples,
//   zero-padded to (next power of 2 >= n_tau) x (next power of 2 >= domain),
//   2-D FFT per component, fftshift of both axes, absolute value.
#include <hip/hip_runtime.h>

#include <hipfft/hipfft.h>

#include "../../include/csl_gpu.h"
#include "csl_host.hpp"

namespace csl {

// padded complex grids [D][P_tau][P_t] of M_NL (zeros outside the data)
__global__ void k_nl_fill(const double* __restrict__ M0, const double* __restrict__ M1,
                          const double* __restrict__ M01, int n_tau, int n_t, int D, int domain, int Ptau,
                          int Pt, double2* __restrict__ out) {
    const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t total = size_t(D) * Ptau * Pt;
    if (id >= total) return;
    const int j = int(id % Pt);
    const int i = int((id / Pt) % Ptau);
    const int d = int(id / (size_t(Pt) * Ptau));
    double v = 0.0;
    if (i < n_tau && j < domain) {
        const int t = n_t - domain + j;
        v = M01[(size_t(i) * n_t + t) * D + d] - M0[size_t(t) * D + d] - M1[(size_t(i) * n_t + t) * D + d];
    }
    out[id] = make_double2(v, 0.0);
}

// |FFT| with both axes fftshifted: out[(i*Pt + j)*D + d]
__global__ void k_nl_abs_shift(const double2* __restrict__ F, int D, int Ptau, int Pt, double* __restrict__ out) {
    const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t total = size_t(D) * Ptau * Pt;
    if (id >= total) return;
    const int j = int(id % Pt);
    const int i = int((id / Pt) % Ptau);
    const int d = int(id / (size_t(Pt) * Ptau));
    const int si = (i + Ptau / 2) % Ptau;   // fftshift: output row i holds frequency row si
    const int sj = (j + Pt / 2) % Pt;
    const double2 v = F[(size_t(d) * Ptau + si) * Pt + sj];
    out[(size_t(i) * Pt + j) * D + d] = sqrt(v.x * v.x + v.y * v.y);
}

}  // namespace csl

using namespace cslh;

namespace {
int next_pow2(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}
}  // namespace

extern "C" int csl_nonlinear_spectrum(size_t n_tau, size_t n_t, size_t D, const double* M0, const double* M1,
                                      const double* M01, size_t max_out, double* out, size_t* P_tau,
                                      size_t* P_t) {
    return guarded([&] {
        if (!M0 || !M1 || !M01 || !out) fail(CSL_ERR_NULL, "null argument");
        if (n_tau == 0 || n_t < 2 || D == 0) fail(CSL_ERR_ARG, "empty scan");
        if (n_tau > (1u << 20) || n_t > (1u << 24)) fail(CSL_ERR_ARG, "scan too large");
        const int domain = int(n_t / 2);
        const int Ptau = next_pow2(int(n_tau)), Pt = next_pow2(domain);
        if (P_tau) *P_tau = size_t(Ptau);
        if (P_t) *P_t = size_t(Pt);
        const size_t total = D * size_t(Ptau) * Pt;
        if (max_out < total) fail(CSL_ERR_CAPACITY, "output smaller than the padded spectrum");
        check_finite(M0, n_t * D, "M0");
        check_finite(M1, n_tau * n_t * D, "M1");
        check_finite(M01, n_tau * n_t * D, "M01");
        hipStream_t s = nullptr;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        struct Guard {
            hipStream_t s;
            hipfftHandle p = 0;
            ~Guard() {
                if (p) hipfftDestroy(p);
                if (s) (void)hipStreamDestroy(s);
            }
        } g{s};
        DBuf<double> d0, d1, d01, grid, res;
        d0.upload(M0, n_t * D, s);
        d1.upload(M1, n_tau * n_t * D, s);
        d01.upload(M01, n_tau * n_t * D, s);
        grid.ensure(2 * total);
        res.ensure(total);
        auto* gc = reinterpret_cast<double2*>(grid.p);
        const unsigned nb = unsigned((total + 255) / 256);
        hipLaunchKernelGGL(csl::k_nl_fill, dim3(nb), dim3(256), 0, s, d0.p, d1.p, d01.p, int(n_tau), int(n_t),
                           int(D), domain, Ptau, Pt, gc);
        hip_check(hipGetLastError(), "k_nl_fill");
        int dims[2] = {Ptau, Pt};
        size_t ws = 0;
        if (hipfftCreate(&g.p) != HIPFFT_SUCCESS) fail(CSL_ERR_HIP, "hipfftCreate failed");
        if (hipfftMakePlanMany(g.p, 2, dims, nullptr, 1, Ptau * Pt, nullptr, 1, Ptau * Pt, HIPFFT_Z2Z, int(D), &ws) !=
                HIPFFT_SUCCESS ||
            hipfftSetStream(g.p, s) != HIPFFT_SUCCESS)
            fail(CSL_ERR_HIP, "2-D FFT plan failed");
        if (hipfftExecZ2Z(g.p, reinterpret_cast<hipfftDoubleComplex*>(gc), reinterpret_cast<hipfftDoubleComplex*>(gc),
                          HIPFFT_FORWARD) != HIPFFT_SUCCESS)
            fail(CSL_ERR_HIP, "2-D FFT failed");
        hipLaunchKernelGGL(csl::k_nl_abs_shift, dim3(nb), dim3(256), 0, s, gc, int(D), Ptau, Pt, res.p);
        hip_check(hipGetLastError(), "k_nl_abs_shift");
        copy_back(out, res.p, total, s);
        hip_check(hipStreamSynchronize(s), "nonlinear spectrum");
    });
}
