// gpu_api_shim.cpp — the reference's gpu:: host API (lattice_gpu_api.h:24-216)
// implemented over the C ABI of this library.  See
// include/classical_spin_amd/lattice_gpu_api.h for the behavioural notes.
#include "classical_spin_amd/lattice_gpu_api.h"

#include <iostream>
#include <stdexcept>
#include <string>

#include "csl_gpu.h"

namespace gpu {

struct GPULatticeDataHandle {
    csl_lattice* h = nullptr;
    size_t N = 0, D = 0, n_atoms = 0;
    double t_pulse_1 = 0.0;
    double amp = 0.0;
};

namespace {

void chk(int rc, const char* where) {
    if (rc != CSL_OK) throw std::runtime_error(std::string(where) + ": " + csl_last_error());
}

}  // namespace

GPULatticeDataHandle* create_gpu_lattice_data(size_t lattice_size, size_t spin_dim, size_t N_atoms,
                                              size_t max_bilinear, const std::vector<double>& flat_field,
                                              const std::vector<double>& flat_onsite,
                                              const std::vector<double>& flat_bilinear,
                                              const std::vector<size_t>& flat_partners,
                                              const std::vector<size_t>& num_bilinear_per_site) {
    if (flat_field.size() != lattice_size * spin_dim || flat_onsite.size() != lattice_size * spin_dim * spin_dim ||
        flat_bilinear.size() != lattice_size * max_bilinear * spin_dim * spin_dim ||
        flat_partners.size() != lattice_size * max_bilinear || num_bilinear_per_site.size() != lattice_size)
        throw std::runtime_error("create_gpu_lattice_data: table sizes do not match the dimensions");
    auto* H = new GPULatticeDataHandle();
    const int rc = csl_create_lattice(lattice_size, spin_dim, N_atoms, max_bilinear, flat_field.data(),
                                      flat_onsite.data(), flat_bilinear.data(), flat_partners.data(),
                                      num_bilinear_per_site.data(), &H->h);
    if (rc != CSL_OK) {
        delete H;
        chk(rc, "create_gpu_lattice_data");
    }
    H->N = lattice_size;
    H->D = spin_dim;
    H->n_atoms = N_atoms;
    return H;
}

void destroy_gpu_lattice_data(GPULatticeDataHandle* handle) {
    if (!handle) return;
    if (handle->h) csl_destroy_lattice(handle->h);
    delete handle;
}

void set_gpu_pulse(GPULatticeDataHandle* handle, const std::vector<double>& flat_field_drive,
                   double pulse_amp, double pulse_width, double pulse_freq, double t_pulse_1,
                   double t_pulse_2) {
    if (!handle) return;
    if (flat_field_drive.size() != 2 * handle->n_atoms * handle->D)
        throw std::runtime_error("set_gpu_pulse: field_drive must hold 2*N_atoms*spin_dim values");
    chk(csl_set_pulse(handle->h, flat_field_drive.data(), pulse_amp, pulse_width, pulse_freq, t_pulse_1,
                      t_pulse_2),
        "set_gpu_pulse");
    handle->t_pulse_1 = t_pulse_1;
    handle->amp = pulse_amp;
}

void set_gpu_spins(GPULatticeDataHandle* handle, const std::vector<double>& flat_spins) {
    if (!handle) return;
    chk(csl_set_spins(handle->h, flat_spins.data(), flat_spins.size()), "set_gpu_spins");
}

void get_gpu_spins(GPULatticeDataHandle* handle, std::vector<double>& flat_spins) {
    if (!handle) return;
    flat_spins.resize(handle->N * handle->D);
    const int rc = csl_get_spins(handle->h, flat_spins.data(), flat_spins.size());
    if (rc == CSL_ERR_NOSTATE) {  // lattice_gpu.cu:670-674: no state -> no-op
        flat_spins.clear();
        return;
    }
    chk(rc, "get_gpu_spins");
}

void integrate_gpu(GPULatticeDataHandle* handle, double T_start, double T_end, double dt,
                   size_t save_interval, std::vector<std::pair<double, std::vector<double>>>& trajectory,
                   const std::string& method) {
    if (!handle) return;
    // every method of gpu::ode::step (integrator.cuh:91-458) with its
    // fixed-step loop (integrator.cuh:470-487); unknown names warn and run
    // ssprk53 inside the library, as the reference does
    size_t n_steps = 0, n_rec = 0;
    chk(csl_count_records(T_start, T_end, dt, save_interval, CSL_TGRID_GPU, &n_steps, &n_rec),
        "integrate_gpu");
    const size_t nd = handle->N * handle->D;
    std::vector<double> times(n_rec), states(n_rec * nd);
    size_t got = 0;
    chk(csl_integrate(handle->h, T_start, T_end, dt, save_interval, method.c_str(), CSL_TGRID_GPU, 1, n_rec,
                      times.data(), states.data(), nullptr, nullptr, &got),
        "integrate_gpu");
    for (size_t r = 0; r < got; ++r)
        trajectory.emplace_back(times[r], std::vector<double>(states.begin() + r * nd, states.begin() + (r + 1) * nd));
}

void step_gpu(GPULatticeDataHandle* handle, double t, double dt, const std::string& method) {
    if (!handle) return;
    chk(csl_step(handle->h, t, dt, method.c_str()), "step_gpu");
}

double compute_energy_gpu(GPULatticeDataHandle* handle) {
    if (!handle) return 0.0;
    double E = 0.0;
    chk(csl_energy(handle->h, &E), "compute_energy_gpu");
    return E;
}

void normalize_spins_gpu(GPULatticeDataHandle* handle, double spin_length) {
    if (!handle) return;
    chk(csl_normalize(handle->h, spin_length), "normalize_spins_gpu");
}

void compute_magnetization_gpu(GPULatticeDataHandle* handle, std::vector<double>& mag_local,
                               std::vector<double>& mag_staggered) {
    if (!handle) return;
    mag_local.assign(handle->D, 0.0);
    mag_staggered.assign(handle->D, 0.0);
    chk(csl_magnetization(handle->h, mag_local.data(), mag_staggered.data()), "compute_magnetization_gpu");
}

BatchedMagResult integrate_gpu_batched_ex(GPULatticeDataHandle* handle,
                                          const std::vector<double>& initial_states, bool per_replica,
                                          const std::vector<double>& t_pulse,
                                          const std::vector<double>& pulse_amp,
                                          const std::vector<double>& afm_signs,
                                          const std::vector<double>& sublattice_frames, double T_start,
                                          double T_end, double dt, size_t save_interval) {
    BatchedMagResult r{};
    if (!handle) return r;
    const size_t B = t_pulse.size() / 2;
    if (B == 0) return r;
    chk(csl_set_frames(handle->h, sublattice_frames.empty() ? nullptr : sublattice_frames.data(),
                       afm_signs.empty() ? nullptr : afm_signs.data()),
        "integrate_gpu_batched");
    size_t n_steps = 0, n_rec = 0;
    chk(csl_count_records(T_start, T_end, dt, save_interval, CSL_TGRID_BATCHED, &n_steps, &n_rec),
        "integrate_gpu_batched");
    r.B = B;
    r.spin_dim = handle->D;
    r.times.resize(n_rec);
    r.mag_data.resize(n_rec * B * 3 * handle->D);
    size_t got = 0;
    chk(csl_integrate_batched(handle->h, B, initial_states.data(), per_replica ? 1 : 0, t_pulse.data(),
                              pulse_amp.empty() ? nullptr : pulse_amp.data(), T_start, T_end, dt, save_interval,
                              "rk4", CSL_TGRID_BATCHED, 0, n_rec, r.times.data(), r.mag_data.data(), nullptr,
                              nullptr, &got),
        "integrate_gpu_batched");
    r.n_time_points = got;
    return r;
}

BatchedMagResult integrate_gpu_batched(GPULatticeDataHandle* handle,
                                       const std::vector<double>& flat_initial_state,
                                       const std::vector<double>& tau2_values,
                                       const std::vector<double>& afm_signs,
                                       const std::vector<double>& sublattice_frames, double T_start,
                                       double T_end, double dt, size_t save_interval,
                                       const std::string& method) {
    if (!handle || tau2_values.empty()) return BatchedMagResult{};
    // batched mode is RK4 only; other methods warn and run rk4
    // (lattice_gpu.cu:779-784)
    if (method != "rk4")
        std::cerr << "[integrate_gpu_batched] method='" << method
                  << "' is not supported in batched mode; using rk4." << std::endl;
    // reference semantics: t_pulse_1 and the amplitude are the handle's, only
    // t_pulse_2 varies per replica (lattice_gpu.cu:336-397)
    const size_t B = tau2_values.size();
    std::vector<double> tp(2 * B), am(2 * B, handle->amp);
    for (size_t b = 0; b < B; ++b) {
        tp[2 * b] = handle->t_pulse_1;
        tp[2 * b + 1] = tau2_values[b];
    }
    return integrate_gpu_batched_ex(handle, flat_initial_state, false, tp, am, afm_signs, sublattice_frames,
                                    T_start, T_end, dt, save_interval);
}

void set_gpu_damping(GPULatticeDataHandle* handle, double alpha) {
    if (!handle) return;
    chk(csl_set_damping(handle->h, alpha), "set_gpu_damping");
}

}  // namespace gpu
