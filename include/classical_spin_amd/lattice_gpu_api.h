// lattice_gpu_api.h — drop-in replacement of the reference's GPU host API
// (include/classical_spin/gpu/lattice_gpu_api.h:24-216) implemented over the
// C ABI of libcsl_hip.so (include/csl_gpu.h).
//
// Same namespace, types and signatures, so the reference's Lattice glue
// (lattice.h:5419-5967) compiles and links against libcsl_hip.so unchanged.
// Behavioural differences (documented, not silent):
//   * every method name of gpu::ode::step is accepted (integrator.cuh:91-458);
//     as in the reference an unknown name warns on stderr and runs ssprk53
//     (integrator.cuh:454-458) and integrate_gpu_batched warns and runs rk4
//     for any other method (lattice_gpu.cu:779-784).
//   * device errors throw std::runtime_error with the library's message (the
//     reference checks only at the end of an integration); a null handle
//     stays a no-op as in lattice_gpu.cu:660-710.
//   * compute_energy_gpu returns the energy (the reference throws,
//     lattice_gpu.cu:298-303).
//   * integrate_gpu_batched additionally accepts the per-replica form through
//     integrate_gpu_batched_ex.
#ifndef CLASSICAL_SPIN_AMD_LATTICE_GPU_API_H
#define CLASSICAL_SPIN_AMD_LATTICE_GPU_API_H

#include <cstddef>
#include <string>
#include <utility>
#include <vector>

namespace gpu {

struct GPULatticeDataHandle;

GPULatticeDataHandle* create_gpu_lattice_data(size_t lattice_size, size_t spin_dim, size_t N_atoms,
                                              size_t max_bilinear, const std::vector<double>& flat_field,
                                              const std::vector<double>& flat_onsite,
                                              const std::vector<double>& flat_bilinear,
                                              const std::vector<size_t>& flat_partners,
                                              const std::vector<size_t>& num_bilinear_per_site);

void destroy_gpu_lattice_data(GPULatticeDataHandle* handle);

void set_gpu_pulse(GPULatticeDataHandle* handle, const std::vector<double>& flat_field_drive,
                   double pulse_amp, double pulse_width, double pulse_freq, double t_pulse_1,
                   double t_pulse_2);

void set_gpu_spins(GPULatticeDataHandle* handle, const std::vector<double>& flat_spins);

void get_gpu_spins(GPULatticeDataHandle* handle, std::vector<double>& flat_spins);

void integrate_gpu(GPULatticeDataHandle* handle, double T_start, double T_end, double dt,
                   size_t save_interval, std::vector<std::pair<double, std::vector<double>>>& trajectory,
                   const std::string& method = "ssprk53");

void step_gpu(GPULatticeDataHandle* handle, double t, double dt, const std::string& method = "ssprk53");

double compute_energy_gpu(GPULatticeDataHandle* handle);

void normalize_spins_gpu(GPULatticeDataHandle* handle, double spin_length);

void compute_magnetization_gpu(GPULatticeDataHandle* handle, std::vector<double>& mag_local,
                               std::vector<double>& mag_staggered);

struct BatchedMagResult {
    size_t B;
    size_t spin_dim;
    size_t n_time_points;
    std::vector<double> times;     // [n_time_points]
    std::vector<double> mag_data;  // [n_time_points * B * 3 * spin_dim]
};

BatchedMagResult integrate_gpu_batched(GPULatticeDataHandle* handle,
                                       const std::vector<double>& flat_initial_state,
                                       const std::vector<double>& tau2_values,
                                       const std::vector<double>& afm_signs,
                                       const std::vector<double>& sublattice_frames, double T_start,
                                       double T_end, double dt, size_t save_interval = 1,
                                       const std::string& method = "rk4");

// Per-replica form (no counterpart in the reference): initial states
// [B*N*D] (or [N*D] shared), pulse centres [B*2], pulse amplitudes [B*2].
BatchedMagResult integrate_gpu_batched_ex(GPULatticeDataHandle* handle,
                                          const std::vector<double>& initial_states, bool per_replica,
                                          const std::vector<double>& t_pulse,
                                          const std::vector<double>& pulse_amp,
                                          const std::vector<double>& afm_signs,
                                          const std::vector<double>& sublattice_frames, double T_start,
                                          double T_end, double dt, size_t save_interval = 1);

// Gilbert damping (the reference GPU path has none).
void set_gpu_damping(GPULatticeDataHandle* handle, double alpha);

}  // namespace gpu

#endif
