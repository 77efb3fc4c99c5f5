/*
 * csl_mixed.h — C ABI of the MI355X mixed SU(2) + SU(3) Landau–Lifshitz RK4
 * path (libcsl_hip.so).
 *
 * Replaces the mixed_gpu:: host API of the reference
 * (include/classical_spin/gpu/mixed_lattice_gpu_api.h:25-214) behind a C
 * ABI: POD pointers, size_t counts and int status codes (the CSL_* codes and
 * csl_last_error() of csl_gpu.h).  A C++ shim with the reference's exact
 * mixed_gpu:: signatures is in include/classical_spin_amd/mixed_lattice_gpu_api.h.
 *
 * Physics (MixedLattice::landau_lifshitz, src/core/mixed_lattice_md.cpp:433-605):
 *   SU(2) site i (3 components):
 *     H_i = -B_i + 2 A_i S_i + sum J S_p (SU(2) partners)
 *                            + sum K lambda_q (SU(3) partners, K is 3x8) - drive_SU2(t)
 *     dS/dt = H x S  +  (alpha/|S|) (S (S.H) - H |S|^2)
 *   SU(3) site j (8 Gell-Mann components):
 *     H_j = -B_j + 2 A_j lambda_j + sum J lambda_q + sum K^T S_p - drive_SU3(t)
 *     dlambda_a/dt = f_abc H_b lambda_c - Gamma_a (lambda_a - lambda_eq,a)
 *   drive_s(t) = f_1(t) d_1[atom] + f_2(t) d_2[atom],
 *     f_k = amp_s exp(-((t - t_k)/(2 w_s))^2) cos(freq_s (t - t_k))
 *
 * State layout at the boundary: the reference flat state [N2*3 | N3*8]
 * (mixed_lattice.h:3878-3900).  Tables are given in the CSR form of the
 * reference's nested per-site lists (site i's bonds at off[i] .. off[i+1]):
 *   22: SU(2)->SU(2) 3x3   33: SU(3)->SU(3) 8x8
 *   23: SU(2)->SU(3) 3x8   32: SU(3)->SU(2) 8x3  (row-major)
 * Site index = cell * n_atoms + atom for both species (same cells).
 *
 * Records (one per saved step, per trajectory) are 34 doubles:
 *   [M2_anti(3), M2_local(3), M2_global(3), M3_anti(8), M3_local(8),
 *    M3_global(8), E/(N2+N3)]
 * as the mixed MD observer computes them (mixed_lattice_md.cpp:1249-1281):
 * M2_anti uses the SU(2) frames and AFM signs, M3_anti the alternating sign
 * of the SU(3) site index, globals the sublattice frames; all divided by the
 * species' site count.
 */
#ifndef CSL_MIXED_H
#define CSL_MIXED_H

#include <stddef.h>

#include "csl_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CSL_MIXED_RECORD 34

typedef struct csl_mixed csl_mixed;

/* create_gpu_mixed_lattice_data (mixed_lattice_gpu_api.h:66-87), CSR form.
 * field2 [N2*3], onsite2 [N2*9], field3 [N3*8], onsite3 [N3*64] per site;
 * N2 = n_atoms2*n_cells, N3 = n_atoms3*n_cells. */
int csl_mixed_create(size_t n_atoms2, size_t n_atoms3, size_t n_cells,
                     const double* field2, const double* onsite2,
                     const size_t* off22, const size_t* par22, const double* J22,
                     const double* field3, const double* onsite3,
                     const size_t* off33, const size_t* par33, const double* J33,
                     const size_t* off23, const size_t* par23, const double* J23,
                     const size_t* off32, const size_t* par32, const double* J32,
                     csl_mixed** out);

/* destroy_gpu_mixed_lattice_data (:90). */
int csl_mixed_destroy(csl_mixed* h);

/* get_mixed_lattice_dims (:203-212) plus the compression flags of the four
 * bond tables (1 = one matrix per (sublattice, slot)). */
int csl_mixed_info(csl_mixed* h, size_t* N2, size_t* N3, int* uniform22, int* uniform23,
                   int* uniform33, int* uniform32);

/* Sublattice frames (SU(2) [n_atoms2*9], SU(3) [n_atoms3*64], row-major) and
 * SU(2) AFM signs [n_atoms2] for the records.  Default: identity, +1. */
int csl_mixed_set_frames(csl_mixed* h, const double* frames2, const double* afm2,
                         const double* frames3);

/* set_gpu_pulse_SU2 / set_gpu_pulse_SU3 (:95-113).  species = 2 or 3;
 * drive [2*n_atoms*D] are the two pulses' per-atom directions already in the
 * storage frame (MixedLattice::set_pulse_SU2/SU3 store F^T B,
 * mixed_lattice_md.cpp:114-263). */
int csl_mixed_set_pulse(csl_mixed* h, int species, const double* drive, double amp,
                        double width, double freq, double t1, double t2);

/* SU(2) Gilbert alpha and SU(3) Bloch rates gamma[8] (NULL = 0) towards the
 * equilibrium eq [N3*8] (NULL = 0) (mixed_lattice_md.cpp:570-605). */
int csl_mixed_set_damping(csl_mixed* h, double alpha, const double* gamma8, const double* eq);

/* set_gpu_mixed_spins / get_gpu_mixed_spins (:122-139): flat [N2*3 + N3*8]. */
int csl_mixed_set_spins(csl_mixed* h, const double* flat);
int csl_mixed_get_spins(csl_mixed* h, double* flat);

/* One record (34 doubles, see above) of the handle's state. */
int csl_mixed_record(csl_mixed* h, double* out34);

/* Total energy (mixed_lattice.h:1783-1900) of the handle's state. */
int csl_mixed_energy(csl_mixed* h, double* energy);

/* dS/dt (flat [N2*3 + N3*8]) of the handle's state at time t, drive included. */
int csl_mixed_rhs(csl_mixed* h, double t, double* dsdt);

/* normalize_spins_mixed_gpu (:196-200). */
int csl_mixed_normalize(csl_mixed* h, double spin_length2, double spin_length3);

/* step_mixed_gpu (:166-171): one fixed step of any method name of
 * csl_step (mixed_lattice_gpu.cu:456-467 steps through the same
 * gpu::ode::step); "rk4" runs the fused RK4 stage kernels, the others the
 * generic stage engine. */
int csl_mixed_step(csl_mixed* h, double t, double dt, const char* method);

/* integrate_mixed_gpu (:150-157) from the handle's state, which is advanced,
 * with any method name of csl_step.  tgrid: CSL_TGRID_*.  Outputs (any may be NULL): times [max_rec],
 * states [max_rec*(N2*3+N3*8)], recs [max_rec*34]. */
int csl_mixed_integrate(csl_mixed* h, double T0, double T1, double dt, size_t save_interval,
                        const char* method, int tgrid, size_t max_rec, double* times,
                        double* states, double* recs, size_t* n_rec);

/* B trajectories in one call: init [N2*3+N3*8] shared or [B*(...)]
 * (init_per_replica), pulse centres t_pulse [B*2] shared by both species,
 * pulse_scale [B*2] multiplying each pulse of both species (NULL = 1).
 * recs [max_rec*B*34], final_states [B*(N2*3+N3*8)] (may be NULL). */
int csl_mixed_integrate_batched(csl_mixed* h, size_t B, const double* init, int init_per_replica,
                                const double* t_pulse, const double* pulse_scale, double T0,
                                double T1, double dt, size_t save_interval, const char* method,
                                int tgrid, size_t max_rec, double* times, double* recs,
                                double* final_states, size_t* n_rec);

/* Dormand-Prince 5(4) for B mixed trajectories (MixedLattice::
 * integrate_ode_system with "dopri5", mixed_lattice.h:3960-4010), each with
 * its own step-size sequence; inputs as csl_mixed_integrate_batched.
 * control, tolerances, save_interval, tgrid (control 0), segments /
 * n_segments / max_segments (control 2, W3 chunking of single_pulse_drive,
 * mixed_lattice.h:4841-4870) and max_attempts as csl_integrate_dopri5 of
 * the SU(2) path: 0 fixed steps, 1 integrate_adaptive (molecular_dynamics,
 * mixed_lattice_md.cpp:1314), 2 integrate_const with step control (pulse
 * drivers).  Outputs per trajectory b: times [b*max_rec + r], recs
 * [(b*max_rec + r)*34], final_states [b*(N2*3+N3*8)] (may be NULL), n_rec,
 * n_accepted, n_rejected [B] (the last two may be NULL).  CSL_ERR_CAPACITY
 * when some trajectory has more than max_rec records (the first max_rec
 * are written); CSL_ERR_STEP when the step control fails. */
int csl_mixed_integrate_dopri5(csl_mixed* h, size_t B, const double* init, int init_per_replica,
                               const double* t_pulse, const double* pulse_scale, double T0, double T1,
                               double dt, int control, double abs_tol, double rel_tol,
                               size_t save_interval, int tgrid, const double* segments,
                               const size_t* n_segments, size_t max_segments, size_t max_attempts,
                               size_t max_rec, double* times, double* recs, double* final_states,
                               size_t* n_rec, size_t* n_accepted, size_t* n_rejected);

/* Gradient of the loss on the records w.r.t. a common scale theta of the
 * Fe-Tm (SU(2)-SU(3)) couplings, at theta = 1, for B trajectories set up as
 * in csl_mixed_integrate_batched: loss_b = sum_t sum_j w_j (rec_j(t) -
 * target_j(t))^2 over the records of trajectory b, grad_b = d loss_b /
 * d theta.  The derivative is the exact one of the discrete RK4 scheme,
 * computed forward (tangent-linear: d state / d theta advanced with the state
 * in the same stage launches; for one parameter it equals what the discrete
 * adjoint gives, without checkpoints).  target is [n_rec][34] shared
 * (target_per_replica 0) or [n_rec][B][34]; weights [34] with weights[33]
 * (energy) = 0.  Optional outputs recs and trecs (the tangent records,
 * d rec / d theta) are [max_rec][B][34].  No counterpart in the reference
 * (which has no gradient path); forward model as csl_mixed_integrate_batched. */
int csl_mixed_scale_gradient(csl_mixed* h, size_t B, const double* init, int init_per_replica,
                             const double* t_pulse, const double* pulse_scale, double T0, double T1, double dt,
                             size_t save_interval, int tgrid, const double* target, int target_per_replica,
                             const double* weights, double* loss, double* grad, size_t max_rec, double* recs,
                             double* trecs, size_t* n_rec);

/* Average time of one stage launch of the last integration on this handle
 * (HIP events on the handle's stream) and the number of stage launches; when
 * the last csl_mixed_integrate_batched ran on the persistent kernel, the
 * device time of the whole integration and *launches = 1. */
int csl_mixed_last_kernel_ms(csl_mixed* h, double* ms, size_t* launches);

/* Path of csl_mixed_integrate_batched ("rk4").  mode 0 (default): the
 * persistent kernel when the lattice qualifies (translation-invariant bond
 * tables of all four kinds, per-sublattice fields and onsite matrices, at
 * most 2048 sites of each species) — one workgroup integrates one trajectory
 * over the whole time grid; with more trajectories than co-resident
 * workgroups, each workgroup runs an equal share of the trajectory-steps
 * (whole trajectories plus the head or the rest of a trajectory split with
 * a neighbouring workgroup); -1: always the grid-wide stage kernels (one
 * launch per RK stage, as mixed_lattice_gpu.cu:278-347); -2: the stage
 * kernels' workgroups of the whole batch in one cooperative launch over the
 * time grid (a grid-wide barrier per stage; for a few trajectories of a
 * large lattice, where one workgroup per trajectory is slower), the stage
 * kernels when they cannot all be resident; n > 0: the persistent kernel
 * with at most n workgroups (tests of the work split).
 * The stage-graph mode (csl_mixed_set_stage_graph) keeps the stage kernels.
 * No counterpart in the reference. */
int csl_mixed_set_persistent(csl_mixed* h, int mode);

/* Path of the last csl_mixed_integrate_batched: *persistent = 1 when the
 * persistent kernel ran, 2 the cooperative one, 0 the stage kernels,
 * *workgroups its grid, *split the trajectories run
 * in two pieces by two workgroups, *kernel_ms the device time of the
 * persistent launch (HIP events).  Null outputs are skipped.  No
 * counterpart in the reference. */
int csl_mixed_last_path(csl_mixed* h, int* persistent, int* workgroups, size_t* split, double* kernel_ms);

/* With enable != 0, csl_mixed_integrate_batched issues its launch sequence
 * (4 stage launches and one record reduction per step) from a hipGraph
 * captured on the first call and replayed by later calls with the same
 * shapes, time grid, device buffers and lattice parameters.  Default off.
 * No counterpart in the reference. */
int csl_mixed_set_stage_graph(csl_mixed* h, int enable);

#ifdef __cplusplus
}
#endif
#endif
