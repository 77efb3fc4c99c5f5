/*
 * csl_gpu.h — C ABI of the MI355X (gfx950) Landau–Lifshitz spin-dynamics
 * library (libcsl_hip.so).
 *
 * This is the drop-in boundary for the reference's GPU host API
 * (include/classical_spin/gpu/lattice_gpu_api.h, a C++ ABI with std::vector /
 * std::string).  Every entry point below names the reference function it
 * replaces (file:line).  All arguments are plain pointers, sizes and scalars;
 * no C++ or torch types cross this boundary.  A C++ shim with the reference's
 * exact `gpu::` signatures is provided on top of it in
 * include/classical_spin_amd/lattice_gpu_api.h.
 *
 * Conventions
 *   - Every function returns an int status: CSL_OK (0) on success, a
 *     negative CSL_ERR_* code otherwise.  csl_last_error() returns the
 *     message of the last failure on the calling thread.  (The reference
 *     silently no-ops on a null handle, lattice_gpu.cu:660-710, and throws
 *     std::runtime_error on CUDA errors, gpu_common_helpers.cuh:16-31; here
 *     both become status codes.)
 *   - Arrays are host memory, FP64, in the reference's layouts:
 *       spins          AoS [N*D], site index ((i*d2+j)*d3+k)*n_atoms+atom
 *       magnetisation  [t][b][3][D] with m = 0 M_antiferro, 1 M_local,
 *                      2 M_global, each divided by N
 *                      (BatchedMagResult, lattice_gpu_api.h:161-177)
 *   - All state lives on the device in FP64.  The library picks the device
 *     that is current when the handle is created and uses one stream per
 *     handle.  Handles are not thread-safe (as in the reference).
 */
#ifndef CSL_GPU_H
#define CSL_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CSL_OK               0
#define CSL_ERR_NULL        -1   /* null handle / missing required pointer   */
#define CSL_ERR_ARG         -2   /* inconsistent sizes or invalid value      */
#define CSL_ERR_UNSUPPORTED -3   /* method / physics not implemented         */
#define CSL_ERR_NOSTATE     -4   /* spins were never set on the handle       */
#define CSL_ERR_HIP         -5   /* HIP runtime error                        */
#define CSL_ERR_CAPACITY    -6   /* output buffer too small                  */
#define CSL_ERR_NONFINITE   -7   /* NaN / Inf in an input state              */
#define CSL_ERR_STEP        -8   /* adaptive step-size control failed        */

/* time-grid policies for the integrators (the host computes the step times):
 *   CSL_TGRID_GPU     t += dt while t < T1                (integrator.cuh:470-487)
 *   CSL_TGRID_BATCHED t += dt while t < T1 - 1e-12        (lattice_gpu.cu:557-565)
 *   CSL_TGRID_BOOST   t_k = T0 + k*dt while t_k+dt <= T1  (Boost integrate_const,
 *                                                          lattice.h:4034-4039) */
#define CSL_TGRID_GPU     0
#define CSL_TGRID_BATCHED 1
#define CSL_TGRID_BOOST   2

typedef struct csl_lattice csl_lattice;

const char* csl_last_error(void);
int csl_device_count(int* count);
int csl_version(int* major, int* minor);

/* ---- lattice handle -------------------------------------------------- */

/* Replaces gpu::create_gpu_lattice_data (lattice_gpu_api.h:47-57): the
 * reference's zero-padded layout — field [N*D], onsite [N*D*D], bilinear
 * [N*max_bi*D*D], partners [N*max_bi] (size_t), counts [N]. */
int csl_create_lattice(size_t N, size_t D, size_t n_atoms, size_t max_bi,
                       const double* field, const double* onsite,
                       const double* bilinear, const size_t* partners,
                       const size_t* counts, csl_lattice** out);

/* Same lattice from the CSR mirror bi_flat_offset [N+1] / bi_flat_partner /
 * bi_flat_J (lattice.h:868-903). */
int csl_create_lattice_csr(size_t N, size_t D, size_t n_atoms,
                           const double* field, const double* onsite,
                           const size_t* offset, const size_t* partner,
                           const double* J, csl_lattice** out);

/* Replaces gpu::destroy_gpu_lattice_data (lattice_gpu_api.h:62). */
int csl_destroy_lattice(csl_lattice* h);

/* Sizes of a handle: N, D, n_atoms, padded device sites, max bonds/site and
 * whether the bond matrices were compressed to per-(sublattice, slot) form. */
int csl_lattice_info(const csl_lattice* h, size_t* N, size_t* D, size_t* n_atoms,
                     size_t* n_padded, size_t* max_bonds, int* uniform);

/* Replaces gpu::set_gpu_pulse (lattice_gpu_api.h:67-75): field_drive is
 * [2*n_atoms*D] (already frame-rotated, as Lattice::update_gpu_pulse sends
 * it, lattice.h:5502-5522). */
int csl_set_pulse(csl_lattice* h, const double* field_drive, double amp,
                  double width, double freq, double t_pulse_1, double t_pulse_2);

/* Gilbert damping alpha (Lattice::alpha_gilbert, lattice_md.cpp:88-98; the
 * reference GPU path has no damping, lattice_gpu.cu:73-117). */
int csl_set_damping(csl_lattice* h, double alpha);

/* Sublattice frames [n_atoms*D*D] (row-major R, S_global = R S_local) and
 * AFM signs [n_atoms] used by the on-device magnetisation reductions
 * (integrate_gpu_batched arguments, lattice_gpu_api.h:200-214).  Defaults:
 * identity frames, all signs +1. */
int csl_set_frames(csl_lattice* h, const double* frames, const double* afm_signs);

/* Replace gpu::set_gpu_spins / get_gpu_spins (lattice_gpu_api.h:83-97). */
int csl_set_spins(csl_lattice* h, const double* spins, size_t n);
int csl_get_spins(csl_lattice* h, double* spins, size_t n);

/* ---- single trajectory ---------------------------------------------- */

/* Replaces gpu::step_gpu (lattice_gpu_api.h:142-147): one fixed step of
 * the named method.  Methods are those of the reference GPU stepper
 * gpu::ode::step (integrator.cuh:91-458): "euler", "rk2"/"midpoint", "rk4",
 * "rk5"/"rkck54", "dopri5" (fixed step, no error control), "rk54"/"rkf54"
 * (= dopri5), "rk78"/"rkf78", "bulirsch_stoer"/"bs", "ssprk53"; any other
 * name prints the reference's warning on stderr and runs ssprk53.  NULL
 * means "rk4". */
int csl_step(csl_lattice* h, double t, double dt, const char* method);

/* Replaces gpu::integrate_gpu (lattice_gpu_api.h:117-125).  Integrates the
 * handle's spins from T0 to T1 with fixed dt on the given time grid policy,
 * recording every save_interval steps plus the final state.  Any of
 * times [max_rec], states [max_rec*N*D], mags [max_rec*3*D], energy
 * [max_rec] (E/N) may be NULL.  *n_rec receives the number of records.
 * engine: 0 = auto, 1 = grid-wide stage kernels (one launch per RK stage),
 * 2 = persistent single-workgroup kernel (lattice resident in LDS),
 * 3 = the stage-kernel launch sequence captured once into a hipGraph and
 * replayed by later calls with the same shapes, lattice parameters and
 * device buffers (the inputs are refreshed in place before each replay).
 * method: any name of csl_step.  Methods other than "rk4" run on the
 * generic stage engine (one derivative launch and one combination launch
 * per stage; engine 0 or 1, CSL_ERR_UNSUPPORTED for 2 and 3). */
int csl_integrate(csl_lattice* h, double T0, double T1, double dt,
                  size_t save_interval, const char* method, int tgrid, int engine,
                  size_t max_rec, double* times, double* states, double* mags,
                  double* energy, size_t* n_rec);

/* Total energy of the handle's spins (total_energy_flat, lattice.h:1485-1613;
 * the reference GPU entry throws, lattice_gpu.cu:298-303). */
int csl_energy(csl_lattice* h, double* E);

/* Replaces gpu::normalize_spins_gpu (lattice_gpu_api.h:157). */
int csl_normalize(csl_lattice* h, double spin_length);

/* Replaces gpu::compute_magnetization_gpu (lattice_gpu_api.h:162-166):
 * local sum and alternating-sign staggered sum, both /N. */
int csl_magnetization(csl_lattice* h, double* mag_local, double* mag_staggered);

/* Local field H_i without drive for every site, AoS [N*D]
 * (get_local_field_flat, lattice.h:1653-1762) and dS/dt at time t
 * (landau_lifshitz_flat, lattice_md.cpp:39-153) for the handle's spins. */
int csl_local_field(csl_lattice* h, double* H, size_t n);
int csl_rhs(csl_lattice* h, double t, double* dsdt, size_t n);

/* ---- batched trajectories (persistent kernel) ------------------------ */

/* Generalises gpu::integrate_gpu_batched (lattice_gpu_api.h:200-214).
 *   B              number of trajectories
 *   init           [N*D] shared by all (init_per_replica = 0) or [B*N*D]
 *   t_pulse        [B*2] pulse centres per trajectory (reference: t_pulse_1
 *                  shared, t_pulse_2 = tau2[b])
 *   pulse_amp      [B*2] per-trajectory amplitude of pulse 1 / pulse 2
 *                  (0 switches a pulse off; the reference pushes pulse 2 past
 *                  T_end for M0, lattice.h:5688-5692).  NULL = the handle's
 *                  amplitude for both pulses.
 *   width/freq and the drive directions come from csl_set_pulse; frames and
 *   AFM signs from csl_set_frames.
 *   mags           [max_rec*B*3*D] BatchedMagResult layout; energy
 *                  [max_rec*B] (E/N, may be NULL); final_states [B*N*D]
 *                  (may be NULL); times [max_rec].
 * method "rk4" only (the reference falls back to rk4, lattice_gpu.cu:781-784).
 * block: workgroup size of the one-workgroup-per-trajectory kernel
 *        (0 = autotuned); -1 selects the grid-wide stage kernels (the path
 *        lattices too large for one workgroup's LDS always take). */
int csl_integrate_batched(csl_lattice* h, size_t B, const double* init,
                          int init_per_replica, const double* t_pulse,
                          const double* pulse_amp, double T0, double T1, double dt,
                          size_t save_interval, const char* method, int tgrid,
                          int block, size_t max_rec, double* times, double* mags,
                          double* energy, double* final_states, size_t* n_rec);

/* Per-trajectory Hamiltonians for csl_integrate_batched: bond matrices
 * J [B*nnz*D*D] in the handle's CSR bond order (same sparsity) and fields
 * [B*N*D], e.g. one coupling set per parameter start.  They must have the
 * handle's structure (translation invariant when the handle is).  B = 0
 * clears them.  No reference counterpart: the reference runs one
 * Hamiltonian per Lattice (lattice_gpu.cu:336-397 shares the tables across
 * replicas). */
int csl_set_replica_hamiltonians(csl_lattice* h, size_t B, const double* J, const double* field);

/* Number of records an integration will produce (same grid rules). */
int csl_count_records(double T0, double T1, double dt, size_t save_interval,
                      int tgrid, size_t* n_steps, size_t* n_rec);

/* build_pulse_segments (pulse_chunking.h:82-161): the free and
 * pulse-active segments covering [T_start, T_end'] where T_end' is the last
 * point of the grid T_start + k*T_step not after T_end.  Each window
 * [c - window, c + window] around a centre is snapped to the nearest grid
 * points and clipped; overlapping or touching windows merge.  Rows
 * out[3*i + {0, 1, 2}] = (t0, t1, dt_init): contiguous, on the grid;
 * dt_init is T_step for active segments and min(max(T_step,
 * free_dt_factor*T_step), t1 - t0) for free ones.  *n_out receives the
 * count (0 when T_end' <= T_start); CSL_ERR_CAPACITY when it exceeds
 * max_out.  Host-only: needs no device. */
int csl_pulse_segments(double T_start, double T_end, const double* centers, size_t n_centers,
                       double window, double T_step, double free_dt_factor, size_t max_out,
                       double* out, size_t* n_out);

/* ---- 2DCS / pump-probe driver --------------------------------------- */

/* The work of pump_probe_spectroscopy (lattice_md.cpp:683-926) and
 * pump_probe_spectroscopy_gpu_batched (lattice.h:5617-5855) in one batched
 * launch: M0 (pulse at 0), M1(tau) (pulse at tau) and M01(tau) (pulses at 0
 * and tau) for every tau on the grid tau_start + i*tau_step.  field_in is
 * the (global-frame) pulse direction per sublattice [n_atoms*D]; it is
 * rotated by the frames as Lattice::set_pulse does.  Records follow the CPU
 * pulse-drive observer (lattice_md.cpp:532-554): the t = T_start sample is
 * dropped and samples sit on the Boost grid.  Outputs: times [max_rec],
 * M0 [max_rec*3*D], M1 and M01 [max_tau*max_rec*3*D], tau_values
 * [max_tau]. */
int csl_pump_probe(csl_lattice* h, const double* ground, const double* field_in,
                   double amp, double width, double freq,
                   double tau_start, double tau_end, double tau_step,
                   double T0, double T1, double T_step,
                   size_t max_rec, size_t max_tau, size_t* n_rec, size_t* n_tau,
                   double* times, double* M0, double* M1, double* M01,
                   double* tau_values);

/* csl_pump_probe with the reference driver's method and W1 options
 * (pump_probe_spectroscopy(..., method, ..., reuse_m0_for_m1,
 * stationarity_tol, ..., pulse_window_chunking, abs_tol, rel_tol),
 * lattice_md.cpp:683-697).  method "rk4" integrates on the Boost grid as
 * csl_pump_probe; "dopri5" runs every replica with the controlled
 * Dormand-Prince stepper between the observation points (integrate_const
 * with make_controlled(abs_tol, rel_tol), lattice_md.cpp:562-575) through
 * csl_integrate_dopri5, with pulse_window_chunking (the reference default)
 * one integrate_const call per segment of csl_pulse_segments around the
 * replica's pulse centres (window 9 widths, lattice_md.cpp:562-571; for
 * "rk4" the chunked and unchunked runs take the same steps).  When
 * reuse_m0_for_m1 is set and the L_inf norm of dS/dt at the ground state
 * with the drive off (max_dSdt_norm_no_drive, :236-280, evaluated on the
 * device) is <= stationarity_tol, the M1(tau) replicas are not integrated;
 * M1 is synthesised from M0 by the time shift of synthesize_M1_from_M0
 * (:283-317), with the ground-state magnetisations before tau (W1,
 * :729-752).  *used_w1 (may be NULL) receives 1 when that happened,
 * *max_dsdt (may be NULL) the norm when it was evaluated.  The handle's spins
 * hold the ground state afterwards when reuse_m0_for_m1 is set.  Same
 * outputs as csl_pump_probe. */
int csl_pump_probe_ex(csl_lattice* h, const double* ground, const double* field_in,
                      double amp, double width, double freq,
                      double tau_start, double tau_end, double tau_step,
                      double T0, double T1, double T_step,
                      int reuse_m0_for_m1, double stationarity_tol, int pulse_window_chunking,
                      const char* method, double abs_tol, double rel_tol,
                      size_t max_rec, size_t max_tau, size_t* n_rec, size_t* n_tau,
                      double* times, double* M0, double* M1, double* M01,
                      double* tau_values, int* used_w1, double* max_dsdt);

/* Trajectories on lattices of different sizes in one persistent launch:
 * lattice i (handle hs[i], its own Hamiltonian, drive directions, width and
 * frequency) runs B_of[i] trajectories, taken in handle order, from the
 * states init (per lattice B_of[i] states of N_i*3 doubles, reference
 * order, concatenated); t_pulse / pulse_amp [B*2] per trajectory (pulse_amp
 * NULL: each handle's amplitude); records as csl_integrate_batched for
 * B = sum B_of[i]: times [max_rec], mags [max_rec*B*9], energy [max_rec*B]
 * (E/N of the trajectory's lattice; may be NULL).  RK4 on the time grid of
 * tgrid; the handles must share one device and either all or none have
 * translation-invariant tables.  Timing and launch shape are reported on
 * hs[0].  No counterpart in the reference (one lattice per integration). */
int csl_integrate_multi(csl_lattice* const* hs, size_t n_lat, const size_t* B_of, const double* init,
                        const double* t_pulse, const double* pulse_amp, double T0, double T1, double dt,
                        size_t save_interval, int tgrid, size_t max_rec, double* times, double* mags,
                        double* energy, size_t* n_rec);

/* Bytes of the per-GPU scratch region of csl_integrate_batched_noalloc. */
#define CSL_NOALLOC_SCRATCH_BYTES (24u << 20)

/* csl_integrate_batched on the persistent kernel without allocating device
 * memory.  No counterpart in the reference: for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Inputs and outputs are DEVICE pointers on the
 * handle's device: d_init [N*3] (AoS, reference order) or [B*N*3] with
 * init_per_replica, d_t_pulse and d_pulse_amp [B*2]; outputs d_mags
 * [n_rec*B*9], d_energy [n_rec*B] (E/N, may be NULL), d_final_states
 * [B*N*3] (may be NULL); *n_rec receives the record count.  The time grid,
 * the work-split schedule, the seam states and the flags live in a
 * file-scope device array of CSL_NOALLOC_SCRATCH_BYTES per GPU (held until
 * the process ends); a launch that needs more returns CSL_ERR_CAPACITY.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * race.  Lattices that do not fit one workgroup return
 * CSL_ERR_UNSUPPORTED.  Synchronous on the handle's stream; the launch
 * shape is the handle's (autotuned or set) choice for B. */
int csl_integrate_batched_noalloc(csl_lattice* h, size_t B, const double* d_init, int init_per_replica,
                                  const double* d_t_pulse, const double* d_pulse_amp, double T0, double T1,
                                  double dt, size_t save_interval, int tgrid, double* d_mags,
                                  double* d_energy, double* d_final_states, size_t* n_rec);

/* ---- Dormand-Prince 5(4) --------------------------------------------- */

/* Integrates B trajectories with the Dormand-Prince 5(4) pair, one
 * workgroup per trajectory with its own step-size sequence (lattices of up
 * to 2048 sites; CSL_ERR_UNSUPPORTED beyond).  control selects how the
 * reference uses the stepper (integrate_ode_system, lattice.h:4072-4085):
 *   0  fixed steps dt on the time grid of `tgrid`, no error control (the
 *      reference GPU stepper's "dopri5", integrator.cuh:186-228);
 *   1  integrate_adaptive with make_controlled(abs_tol, rel_tol): observer at
 *      T0 and after every accepted step, the last step cut to end at T1
 *      (molecular_dynamics_cpu, lattice_md.cpp:452-453);
 *   2  integrate_const with the controlled stepper: observer on the
 *      integrate_const grid T0 + k dt, adaptive steps inside each interval,
 *      the step size carried from one interval to the next (the pulse
 *      drivers, lattice_md.cpp:570-575).  With segments (may be NULL) each
 *      trajectory b makes one integrate_const call per row j <
 *      n_segments[b] of segments[(b*max_segments + j)*2 + {0, 1}] = (t0, t1),
 *      each starting again from the step hint dt with the observer at t0
 *      (W3 chunking, lattice_md.cpp:562-571; the seams are observed twice,
 *      as the reference's observer sees them).
 * tgrid applies to control 0; control 2 uses the integrate_const grid.
 * Observer call c is recorded when c % save_interval == 0 (lattice_md.cpp:
 * 399).  init / t_pulse / pulse_amp as csl_integrate_batched (t_pulse may be
 * NULL: zeros; pulse_amp NULL: the handle's amplitude); per-replica
 * Hamiltonians apply.  Outputs per trajectory b: times [b*max_rec + r],
 * mags [(b*max_rec + r)*9 + j], energy [b*max_rec + r] (E/N, may be NULL),
 * final_states [b*N*3] (may be NULL), n_rec[b] (records the run produced),
 * n_accepted[b], n_rejected[b] (may be NULL).  When some n_rec[b] exceeds
 * max_rec the first max_rec records are written and CSL_ERR_CAPACITY is
 * returned, so the caller can repeat with a larger buffer (the run is
 * deterministic).  CSL_ERR_STEP: 500 rejections in a row, a step size that no
 * longer advances t, or more than max_attempts attempts (0: 5e7). */
int csl_integrate_dopri5(csl_lattice* h, size_t B, const double* init, int init_per_replica,
                         const double* t_pulse, const double* pulse_amp,
                         double T0, double T1, double dt, int control,
                         double abs_tol, double rel_tol, size_t save_interval, int tgrid,
                         const double* segments, const size_t* n_segments, size_t max_segments,
                         size_t max_attempts, size_t max_rec, double* times, double* mags,
                         double* energy, double* final_states, size_t* n_rec,
                         size_t* n_accepted, size_t* n_rejected);

/* ---- autotuner ------------------------------------------------------- */

/* Times the persistent kernel for B trajectories over n_probe_steps at each
 * candidate workgroup size and stores the fastest for this lattice shape;
 * later csl_integrate_batched calls with block = 0 use it.  *best_block
 * receives the choice, *best_ms its time. */
int csl_autotune_batched(csl_lattice* h, size_t B, size_t n_probe_steps,
                         int* best_block, double* best_ms);

/* ---- timing support -------------------------------------------------- */

/* Average duration (ms) of the dominant kernel of the last
 * csl_integrate_batched / csl_integrate call on this handle, measured with
 * HIP events on the handle's stream, and the number of launches it spans. */
int csl_last_kernel_ms(csl_lattice* h, double* ms, size_t* launches);

/* Launch shape of the last integration on this handle: workgroup size,
 * sites per thread and flags of the one-workgroup-per-trajectory kernel
 * (flags bit 0: two LDS state buffers, one barrier per RK stage; bit 1: the
 * four RK stages specialised at compile time), or block = -1 (sites per
 * thread and flags 0) after the grid-wide stage kernels. */
int csl_last_launch_config(csl_lattice* h, int* block, int* sites_per_thread, int* flags);

/* Fix the persistent launch shape used by this handle for batches of B's
 * size class (1, 2..63, >= 64 trajectories), as the autotuner would;
 * rejects shapes that do not fit.  The choice is per handle. */
int csl_set_launch_config(csl_lattice* h, size_t B, int block, int sites_per_thread, int flags);

/* The persistent launch shapes that fit this handle's lattice (the
 * autotuner's candidate set): workgroup size, sites per thread, flags (as
 * csl_last_launch_config) and the compile-time variant the launcher uses
 * for the shape (bit 0: per-sublattice field with full coordination, bit 1:
 * atom-major site mapping).  *n_out receives the count; CSL_ERR_CAPACITY
 * when it exceeds max_out.  No counterpart in the reference (one launch
 * shape per kernel). */
int csl_persistent_candidates(csl_lattice* h, size_t max_out, int* block, int* sites_per_thread,
                              int* flags, int* variant, size_t* n_out);

/* With enable != 0, csl_integrate_batched calls that run on the grid-wide
 * stage kernels (lattices larger than one workgroup, or block = -1) issue
 * their launch sequence (drive table, 4 stage launches and one record
 * reduction per step) from a hipGraph captured on the first call and
 * replayed by later calls with the same shapes, device buffers and lattice
 * parameters, as engine 3 of csl_integrate does.  Default off.  No
 * counterpart in the reference. */
int csl_set_stage_graph(csl_lattice* h, int enable);

/* Variant bits (as csl_persistent_candidates) of the last persistent launch
 * on this handle. */
int csl_last_launch_variant(csl_lattice* h, int* variant);

/* Partitioned persistent kernel for batched RK4 (csl_integrate_batched with
 * block = 0): each trajectory is spread over P co-resident workgroups, each
 * holding a contiguous range of cells in LDS, with the halo values exchanged
 * through global memory once per RK stage.  No reference counterpart (the
 * reference launches one kernel per stage, lattice_gpu.cu:502-565).
 * parts = 0 (default): used for lattices whose state does not fit one
 * workgroup, with the fewest parts that fit; parts > 0: always, with that
 * many parts (CSL_ERR_UNSUPPORTED when they do not fit); parts = -1: never
 * (grid-wide stage kernels).  Needs translation-invariant bond tables and
 * per-sublattice fields; other lattices take the stage kernels.  The parts
 * of a trajectory wait for each other every stage, so the launch assumes its
 * workgroups are co-resident (it is sized to fit the GPU at once); other
 * work running on the device at the same time can delay a part beyond the
 * bounded wait, and the call then fails with CSL_ERR_HIP. */
int csl_set_partition(csl_lattice* h, int parts);
/* Parts per trajectory and trajectory groups of the last batched launch
 * (0, 0 when it did not use the partitioned kernel). */
int csl_last_partition(csl_lattice* h, int* parts, int* groups);

/* Work split of the one-workgroup-per-trajectory kernel (on by default):
 * when a batch holds more trajectories than workgroups fit the GPU at once,
 * that many workgroups share the trajectory-steps evenly, a trajectory cut
 * between two of them continuing from the state the first one saves (same
 * arithmetic, bit-identical results).  Off: one workgroup per trajectory,
 * run in rounds.  No counterpart in the reference, whose batched GPU path
 * launches grid-wide kernels per RK stage. */
int csl_set_batch_split(csl_lattice* h, int enable);

/* Workgroups of the last persistent launch on this handle (the batch size
 * unless the work was split). */
int csl_last_grid(csl_lattice* h, int* workgroups);

/* The split used for B trajectories of n_steps steps on G workgroups
 * (B > G): rows[(g*pieces_per_wg + i)*5 + {0..4}] = (trajectory, first step,
 * end step, slot loaded (-1: initial state), slot saved (-1: none)) of piece
 * i of workgroup g, -1 rows unused; *pieces_per_wg receives the row stride.
 * Host-only. */
int csl_split_schedule(size_t B, size_t G, size_t n_steps, size_t max_rows, int* rows,
                       size_t* pieces_per_wg);

/* Zero-temperature relaxation of B states in place (states [B*N*3], AoS):
 * Lattice::greedy_quench (lattice.h:3352-3368) — sweeps that set every spin
 * antiparallel to its local field, S = -H/|H| * spin_length (skipped when
 * |H| < 1e-15; deterministic_sweep, lattice.h:3334-3349), until the energy
 * changes by at most rel_tol * (|E_prev| + 1e-18) or max_sweeps.  Sites are
 * visited colour class by colour class of a greedy colouring of the bond
 * graph (the reference draws them at random).  sweeps_done [B] (optional)
 * receives each state's sweep count. */
int csl_quench(csl_lattice* h, size_t B, double* states, size_t max_sweeps, double rel_tol,
               double spin_length, size_t* sweeps_done);

/* Dynamical structure factor of B undriven trajectories (the reference
 * computes it in Python from saved spin snapshots,
 * util/readers_new/reader_pyrochlore.py:150-270):
 *   S_a(q, t) = sum_i exp(-i q.r_i) S_ia(t),  q = 2 pi (h/L1, k/L2, l/L3)
 *   sqw[q * n_t + w] = 1/(B N n_t) sum_b sum_a |sum_t exp(-2 pi i w t/n_t) S_a(q, t)|^2
 * with q = (h*L2 + k)*L3 + l, r_i = cell (i, j, k) + positions[atom]
 * (fractional coordinates of the lattice vectors), t over the n_t =
 * floor(n_steps / sample_interval) samples taken every sample_interval RK4
 * steps from T0 (grid policy tgrid).  L1*L2*L3 must equal the number of
 * cells.  Runs the grid-wide stage kernels; the FFTs are hipFFT.  mags
 * [(n_steps+1)*B*9] (optional) receives M(t) at every step. */
int csl_structure_factor(csl_lattice* h, size_t B, const double* init, int init_per_replica,
                         double T0, double T1, double dt, int tgrid, size_t sample_interval,
                         size_t L1, size_t L2, size_t L3, const double* positions, size_t max_t,
                         double* sqw, size_t* n_t, double* mags);


/* 2-D nonlinear spectrum of a pump-probe scan
 * (util/readers_new/reader_pyrochlore.py:760-830): with domain = n_t/2,
 *   M_NL[tau][t] = M01[tau][n_t-domain+t] - M0[n_t-domain+t] - M1[tau][n_t-domain+t]
 * per component d < D, zero-padded to P_tau x P_t (the next powers of two
 * of n_tau and domain), 2-D FFT, both axes fftshifted;
 * out[(i*P_t + j)*D + d] = |FFT|.  Inputs (host): M0 [n_t*D],
 * M1 / M01 [n_tau*n_t*D]; max_out >= P_tau*P_t*D.  No handle needed. */
int csl_nonlinear_spectrum(size_t n_tau, size_t n_t, size_t D, const double* M0, const double* M1,
                           const double* M01, size_t max_out, double* out, size_t* P_tau,
                           size_t* P_t);

/* Gradient of a trajectory loss by the discrete adjoint of RK4.  The
 * reference has no gradient path; this entry is for fitting Hamiltonian
 * parameters to a target M(t) (the reference's MD path is the forward model:
 * lattice_md.cpp:39-153 RHS, Boost runge_kutta4 steps, lattice_md.cpp:383-436
 * observer).
 * For each trajectory b (initial state init [N*3] shared or [B*N*3] with
 * init_per_replica; pulses t_pulse / pulse_amp [B*2] as in
 * csl_integrate_batched; the handle's per-replica Hamiltonians when set with
 * csl_set_replica_hamiltonians for this B), integrate T0 -> T1 (grid policy
 * tgrid), record m = [M_anti, M_local, M_global] at every step and after the
 * last, and return
 *   loss[b]   = sum_n sum_j weights[j] (m_bj(t_n) - target[n*9 + j])^2
 *   grad[b*P + p] = d loss[b] / d theta_p
 * for P = n_params <= 8 parameters entering the Hamiltonian linearly:
 * dJ [P*nnz*D*D] in the handle's CSR bond order and dB [P*N*D] per site
 * (both translation invariant when the handle's tables are).  n_target must be
 * n_steps + 1.  The states are checkpointed every ckpt_interval steps
 * (0 = ceil(sqrt(n_steps))) and segments are recomputed on the way back.
 * Optional outputs: records [n_target*B*9] (the forward m), grad_init
 * [B*N*3] (d loss / d initial state).  Requires translation-invariant bond
 * tables whose reverse bonds carry J^T (the Lattice expansion,
 * lattice.h:790-803), at most 6 bonds per site and 3*NP <= 6144 doubles
 * (one workgroup per trajectory). */
int csl_adjoint_gradient(csl_lattice* h, size_t B, const double* init, int init_per_replica,
                         const double* t_pulse, const double* pulse_amp, double T0, double T1,
                         double dt, int tgrid, const double* target, size_t n_target,
                         const double* weights, size_t n_params, const double* dJ,
                         const double* dB, size_t ckpt_interval, double* loss, double* grad,
                         double* records, double* grad_init);

#ifdef __cplusplus
}
#endif
#endif /* CSL_GPU_H */
