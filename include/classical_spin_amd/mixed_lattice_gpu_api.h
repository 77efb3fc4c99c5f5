// mixed_lattice_gpu_api.h — drop-in replacement of the reference's mixed
// GPU host API (include/classical_spin/gpu/mixed_lattice_gpu_api.h:25-214)
// implemented over the C ABI of libcsl_hip.so (include/csl_mixed.h).
//
// Same namespace and signatures, so the reference's MixedLattice glue
// (mixed_lattice.h:6041-6560) compiles and links against libcsl_hip.so.
// Behavioural differences (documented, not silent):
//   * every method name of gpu::ode::step is accepted ("rk4" on the fused
//     RK4 stage kernels, the others on the generic stage engine; unknown
//     names warn on stderr and run ssprk53, as the reference does); errors
//     throw std::runtime_error with the library's message; a null handle
//     stays a no-op as in the reference.
//   * the device RHS also carries the CPU path's SU(2) Gilbert and SU(3)
//     Bloch damping; both are off unless csl_mixed_set_damping is called, so
//     a reference caller sees the reference GPU's undamped dynamics.
#ifndef CLASSICAL_SPIN_AMD_MIXED_LATTICE_GPU_API_H
#define CLASSICAL_SPIN_AMD_MIXED_LATTICE_GPU_API_H

#include <cstddef>
#include <string>
#include <utility>
#include <vector>

struct csl_mixed;

namespace mixed_gpu {

struct GPUMixedLatticeDataHandle;

GPUMixedLatticeDataHandle* create_gpu_mixed_lattice_data(
    size_t lattice_size_SU2, size_t spin_dim_SU2, size_t N_atoms_SU2,
    size_t lattice_size_SU3, size_t spin_dim_SU3, size_t N_atoms_SU3,
    size_t max_bilinear_SU2, size_t max_bilinear_SU3, size_t max_mixed_bilinear, size_t max_mixed_bilinear_SU3,
    const std::vector<double>& flat_field_SU2,
    const std::vector<double>& flat_onsite_SU2,
    const std::vector<double>& flat_bilinear_SU2,
    const std::vector<size_t>& flat_partners_SU2,
    const std::vector<size_t>& num_bilinear_per_site_SU2,
    const std::vector<double>& flat_field_SU3,
    const std::vector<double>& flat_onsite_SU3,
    const std::vector<double>& flat_bilinear_SU3,
    const std::vector<size_t>& flat_partners_SU3,
    const std::vector<size_t>& num_bilinear_per_site_SU3,
    const std::vector<double>& flat_mixed_bilinear,
    const std::vector<size_t>& flat_mixed_partners_SU2,
    const std::vector<size_t>& flat_mixed_partners_SU3,
    const std::vector<size_t>& num_mixed_per_site_SU2,
    const std::vector<double>& flat_mixed_bilinear_SU3,
    const std::vector<size_t>& flat_mixed_partners_SU2_from_SU3,
    const std::vector<size_t>& num_mixed_per_site_SU3);

void destroy_gpu_mixed_lattice_data(GPUMixedLatticeDataHandle* handle);

void set_gpu_pulse_SU2(GPUMixedLatticeDataHandle* handle, const std::vector<double>& flat_field_drive,
                       double pulse_amp, double pulse_width, double pulse_freq, double t_pulse_1,
                       double t_pulse_2);

void set_gpu_pulse_SU3(GPUMixedLatticeDataHandle* handle, const std::vector<double>& flat_field_drive,
                       double pulse_amp, double pulse_width, double pulse_freq, double t_pulse_1,
                       double t_pulse_2);

void set_gpu_mixed_spins(GPUMixedLatticeDataHandle* handle, const std::vector<double>& flat_spins);

void get_gpu_mixed_spins(GPUMixedLatticeDataHandle* handle, std::vector<double>& flat_spins);

void integrate_mixed_gpu(GPUMixedLatticeDataHandle* handle, double T_start, double T_end, double dt,
                         size_t save_interval,
                         std::vector<std::pair<double, std::vector<double>>>& trajectory,
                         const std::string& method = "ssprk53");

void step_mixed_gpu(GPUMixedLatticeDataHandle* handle, double t, double dt,
                    const std::string& method = "ssprk53");

void compute_magnetization_mixed_gpu(GPUMixedLatticeDataHandle* handle, std::vector<double>& mag_SU2,
                                     std::vector<double>& mag_staggered_SU2, std::vector<double>& mag_SU3,
                                     std::vector<double>& mag_staggered_SU3);

void normalize_spins_mixed_gpu(GPUMixedLatticeDataHandle* handle, double spin_length_SU2,
                               double spin_length_SU3);

void get_mixed_lattice_dims(GPUMixedLatticeDataHandle* handle, size_t& lattice_size_SU2, size_t& spin_dim_SU2,
                            size_t& N_atoms_SU2, size_t& lattice_size_SU3, size_t& spin_dim_SU3,
                            size_t& N_atoms_SU3);

// Access to the underlying C-ABI handle (damping, batched runs, records).
csl_mixed* native_handle(GPUMixedLatticeDataHandle* handle);

}  // namespace mixed_gpu

#endif
