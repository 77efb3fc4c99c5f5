// mixed_gpu_api_shim.cpp — the reference's mixed_gpu:: host API
// (mixed_lattice_gpu_api.h:25-214) implemented over the C ABI of this
// library.  See include/classical_spin_amd/mixed_lattice_gpu_api.h for the
// behavioural notes.
#include "classical_spin_amd/mixed_lattice_gpu_api.h"

#include <stdexcept>
#include <string>

#include "csl_mixed.h"

namespace mixed_gpu {

struct GPUMixedLatticeDataHandle {
    csl_mixed* h = nullptr;
    size_t N2 = 0, N3 = 0, A2 = 0, A3 = 0;
    bool has_state = false;
};

namespace {

void chk(int rc, const char* where) {
    if (rc != CSL_OK) throw std::runtime_error(std::string(where) + ": " + csl_last_error());
}

// padded per-site slots (SIZE_MAX partners beyond the count,
// mixed_lattice.h:5490-5597) -> CSR
void to_csr(size_t N, size_t maxb, size_t m, const std::vector<size_t>& counts, const std::vector<size_t>& partners,
            const std::vector<double>& mats, std::vector<size_t>& off, std::vector<size_t>& par,
            std::vector<double>& J, const char* what) {
    if (counts.size() != N || partners.size() != N * maxb || mats.size() != N * maxb * m)
        throw std::runtime_error(std::string("create_gpu_mixed_lattice_data: ") + what +
                                 " table sizes do not match the dimensions");
    off.assign(N + 1, 0);
    for (size_t i = 0; i < N; ++i) {
        if (counts[i] > maxb) throw std::runtime_error(std::string("create_gpu_mixed_lattice_data: ") + what +
                                                       " count exceeds the padded width");
        off[i + 1] = off[i] + counts[i];
        for (size_t n = 0; n < counts[i]; ++n) {
            par.push_back(partners[i * maxb + n]);
            J.insert(J.end(), mats.begin() + (i * maxb + n) * m, mats.begin() + (i * maxb + n + 1) * m);
        }
    }
}

}  // namespace

GPUMixedLatticeDataHandle* create_gpu_mixed_lattice_data(
    size_t N2, size_t D2, size_t A2, size_t N3, size_t D3, size_t A3, size_t max22, size_t max33, size_t max23,
    size_t max32, const std::vector<double>& field2, const std::vector<double>& onsite2,
    const std::vector<double>& bil2, const std::vector<size_t>& par2, const std::vector<size_t>& cnt2,
    const std::vector<double>& field3, const std::vector<double>& onsite3, const std::vector<double>& bil3,
    const std::vector<size_t>& par3, const std::vector<size_t>& cnt3, const std::vector<double>& mixed,
    const std::vector<size_t>& /*mixed_partners_SU2 (the source site itself)*/,
    const std::vector<size_t>& mixed_par3, const std::vector<size_t>& mixed_cnt2,
    const std::vector<double>& mixed3, const std::vector<size_t>& mixed_par2, const std::vector<size_t>& mixed_cnt3) {
    if (D2 != 3 || D3 != 8) throw std::runtime_error("create_gpu_mixed_lattice_data: spin dims must be 3 and 8");
    if (!A2 || !A3 || N2 % A2 || N3 % A3 || N2 / A2 != N3 / A3)
        throw std::runtime_error("create_gpu_mixed_lattice_data: the two species must share the unit cells");
    if (field2.size() != N2 * 3 || onsite2.size() != N2 * 9 || field3.size() != N3 * 8 || onsite3.size() != N3 * 64)
        throw std::runtime_error("create_gpu_mixed_lattice_data: field / onsite sizes do not match");
    std::vector<size_t> o22, p22, o33, p33, o23, p23, o32, p32;
    std::vector<double> J22, J33, J23, J32;
    to_csr(N2, max22, 9, cnt2, par2, bil2, o22, p22, J22, "SU(2)");
    to_csr(N3, max33, 64, cnt3, par3, bil3, o33, p33, J33, "SU(3)");
    to_csr(N2, max23, 24, mixed_cnt2, mixed_par3, mixed, o23, p23, J23, "mixed SU(2)");
    to_csr(N3, max32, 24, mixed_cnt3, mixed_par2, mixed3, o32, p32, J32, "mixed SU(3)");
    auto* H = new GPUMixedLatticeDataHandle();
    const int rc = csl_mixed_create(A2, A3, N2 / A2, field2.data(), onsite2.data(), o22.data(), p22.data(), J22.data(),
                                    field3.data(), onsite3.data(), o33.data(), p33.data(), J33.data(), o23.data(),
                                    p23.data(), J23.data(), o32.data(), p32.data(), J32.data(), &H->h);
    if (rc != CSL_OK) {
        delete H;
        chk(rc, "create_gpu_mixed_lattice_data");
    }
    H->N2 = N2;
    H->N3 = N3;
    H->A2 = A2;
    H->A3 = A3;
    return H;
}

void destroy_gpu_mixed_lattice_data(GPUMixedLatticeDataHandle* handle) {
    if (!handle) return;
    csl_mixed_destroy(handle->h);
    delete handle;
}

void set_gpu_pulse_SU2(GPUMixedLatticeDataHandle* handle, const std::vector<double>& drive, double amp, double width,
                       double freq, double t1, double t2) {
    if (!handle) return;
    if (drive.size() != 2 * handle->A2 * 3) throw std::runtime_error("set_gpu_pulse_SU2: drive size mismatch");
    chk(csl_mixed_set_pulse(handle->h, 2, drive.data(), amp, width, freq, t1, t2), "set_gpu_pulse_SU2");
}

void set_gpu_pulse_SU3(GPUMixedLatticeDataHandle* handle, const std::vector<double>& drive, double amp, double width,
                       double freq, double t1, double t2) {
    if (!handle) return;
    if (drive.size() != 2 * handle->A3 * 8) throw std::runtime_error("set_gpu_pulse_SU3: drive size mismatch");
    chk(csl_mixed_set_pulse(handle->h, 3, drive.data(), amp, width, freq, t1, t2), "set_gpu_pulse_SU3");
}

void set_gpu_mixed_spins(GPUMixedLatticeDataHandle* handle, const std::vector<double>& flat) {
    if (!handle) return;
    if (flat.size() != handle->N2 * 3 + handle->N3 * 8) throw std::runtime_error("set_gpu_mixed_spins: size mismatch");
    chk(csl_mixed_set_spins(handle->h, flat.data()), "set_gpu_mixed_spins");
    handle->has_state = true;
}

void get_gpu_mixed_spins(GPUMixedLatticeDataHandle* handle, std::vector<double>& flat) {
    if (!handle || !handle->has_state) return;
    flat.resize(handle->N2 * 3 + handle->N3 * 8);
    chk(csl_mixed_get_spins(handle->h, flat.data()), "get_gpu_mixed_spins");
}

void integrate_mixed_gpu(GPUMixedLatticeDataHandle* handle, double T_start, double T_end, double dt,
                         size_t save_interval, std::vector<std::pair<double, std::vector<double>>>& trajectory,
                         const std::string& method) {
    if (!handle || !handle->has_state) return;
    // every method of gpu::ode::step (mixed_lattice_gpu.cu:469-488 drives the
    // same stepper); unknown names warn and run ssprk53 inside the library
    size_t n_steps = 0, n_rec = 0;
    chk(csl_count_records(T_start, T_end, dt, save_interval, CSL_TGRID_GPU, &n_steps, &n_rec), "integrate_mixed_gpu");
    const size_t nd = handle->N2 * 3 + handle->N3 * 8;
    std::vector<double> times(n_rec), states(n_rec * nd);
    size_t got = 0;
    chk(csl_mixed_integrate(handle->h, T_start, T_end, dt, save_interval, method.c_str(), CSL_TGRID_GPU, n_rec,
                            times.data(),
                            states.data(), nullptr, &got),
        "integrate_mixed_gpu");
    for (size_t r = 0; r < got; ++r)
        trajectory.emplace_back(times[r], std::vector<double>(states.begin() + r * nd, states.begin() + (r + 1) * nd));
}

void step_mixed_gpu(GPUMixedLatticeDataHandle* handle, double t, double dt, const std::string& method) {
    if (!handle || !handle->has_state) return;
    chk(csl_mixed_step(handle->h, t, dt, method.c_str()), "step_mixed_gpu");
}

// compute_magnetization_kernel semantics (gpu_common_helpers.cu:163-210):
// local = sum / N, staggered sign = 1 - 2 ((i % N_atoms) & 1), / N
void compute_magnetization_mixed_gpu(GPUMixedLatticeDataHandle* handle, std::vector<double>& m2,
                                     std::vector<double>& s2, std::vector<double>& m3, std::vector<double>& s3) {
    if (!handle || !handle->has_state) return;
    std::vector<double> x;
    get_gpu_mixed_spins(handle, x);
    auto red = [&](const double* v, size_t N, size_t D, size_t A, std::vector<double>& loc, std::vector<double>& stg) {
        loc.assign(D, 0.0);
        stg.assign(D, 0.0);
        for (size_t i = 0; i < N; ++i) {
            const double sg = ((i % A) & 1) ? -1.0 : 1.0;
            for (size_t d = 0; d < D; ++d) {
                loc[d] += v[i * D + d];
                stg[d] += v[i * D + d] * sg;
            }
        }
        for (size_t d = 0; d < D; ++d) {
            loc[d] /= double(N);
            stg[d] /= double(N);
        }
    };
    red(x.data(), handle->N2, 3, handle->A2, m2, s2);
    red(x.data() + handle->N2 * 3, handle->N3, 8, handle->A3, m3, s3);
}

void normalize_spins_mixed_gpu(GPUMixedLatticeDataHandle* handle, double l2, double l3) {
    if (!handle || !handle->has_state) return;
    chk(csl_mixed_normalize(handle->h, l2, l3), "normalize_spins_mixed_gpu");
}

void get_mixed_lattice_dims(GPUMixedLatticeDataHandle* handle, size_t& N2, size_t& D2, size_t& A2, size_t& N3,
                            size_t& D3, size_t& A3) {
    if (!handle) return;
    N2 = handle->N2;
    D2 = 3;
    A2 = handle->A2;
    N3 = handle->N3;
    D3 = 8;
    A3 = handle->A3;
}

csl_mixed* native_handle(GPUMixedLatticeDataHandle* handle) { return handle ? handle->h : nullptr; }

}  // namespace mixed_gpu
