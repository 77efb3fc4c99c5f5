/*
 * oracle.h — CPU FP64 restatement of the reference's Landau–Lifshitz path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (and the "port" CPU baseline timed beside it).  Nothing in
 * the product library links it, loads it or calls it.
 *
 * Every function restates a reference function; the citation (path:line,
 * relative to the reference repository root) is given next to each
 * declaration.  Pinning of each function (which closed form / invariant /
 * brute-force check covers it) is listed in the header of oracle.cpp.
 *
 * Conventions (all as in the reference):
 *   - site index  ((i*d2 + j)*d3 + k)*n_atoms + atom        lattice.h:1057-1059
 *   - state is AoS FP64 [N*D]
 *   - H_i = -B_i + 2 A_i S_i + sum_n J_n S_p(n) - drive(t)   lattice.h:1653-1762,
 *                                                            lattice_md.cpp:178-191
 *   - dS/dt = H x S (SU(2)); d_a = f_abc H_b S_c (SU(3))     lattice_md.cpp:39-153
 *   - magnetisation records are [M_antiferro, M_local, M_global] (3*D doubles)
 */
#ifndef CSL_ORACLE_H
#define CSL_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OracleLattice OracleLattice;

/* Lattice(uc, d1, d2, d3) bond expansion + build_flat_bilinear_tables
 * (lattice.h:601-855, :868-903).  Bonds are given in unit-cell insertion
 * order (UnitCell::set_bilinear_interaction, unitcell.h:166-172). */
OracleLattice* orc_lattice_new(int spin_dim, int n_atoms, int d1, int d2, int d3,
                               int n_bonds, const int* bond_src, const int* bond_partner,
                               const int* bond_offset, const double* bond_J,
                               const double* field_atom, const double* onsite_atom,
                               const double* frames, const double* afm_signs);
/* Lattice from explicit reference tables (the CSR mirror plus per-site field
 * and onsite matrices), for Hamiltonians that are not translation invariant.
 * dims = (d1, d2, d3); N = n_atoms*d1*d2*d3. */
OracleLattice* orc_lattice_from_csr(int spin_dim, int n_atoms, int d1, int d2, int d3,
                                    const size_t* offset, const size_t* partner,
                                    const double* J, const double* field_site,
                                    const double* onsite_site, const double* frames,
                                    const double* afm_signs);
void   orc_lattice_free(OracleLattice* L);
size_t orc_lattice_size(const OracleLattice* L);
size_t orc_total_bonds(const OracleLattice* L);
/* CSR mirror bi_flat_offset / bi_flat_partner / bi_flat_J (lattice.h:868-903). */
void   orc_flat_tables(const OracleLattice* L, size_t* offset, size_t* partner, double* J);
/* Gilbert damping alpha (lattice.h:563). */
void   orc_set_alpha(OracleLattice* L, double alpha);
/* Lattice::set_pulse: field_drive[p] = R_atom * field_in_p (lattice_md.cpp:210-224). */
void   orc_set_pulse(OracleLattice* L, const double* field_in1, double t1,
                     const double* field_in2, double t2,
                     double amp, double width, double freq);
/* Drive reset done at the end of *_pulse_drive (lattice_md.cpp:578-583). */
void   orc_clear_pulse(OracleLattice* L);
/* get_local_field_flat (lattice.h:1653-1762), no drive. */
void   orc_local_field(const OracleLattice* L, const double* state, size_t site, double* H);
/* landau_lifshitz_flat (lattice_md.cpp:39-153). */
void   orc_rhs(const OracleLattice* L, const double* state, double* dsdt, double t);
/* total_energy_flat (lattice.h:1485-1613). */
double orc_energy(const OracleLattice* L, const double* state);
/* [M_antiferro(staggered), M_local, M_global] as the pulse-drive observer
 * assembles them (lattice_md.cpp:533-554, lattice.h:4461-4512). */
void   orc_mags(const OracleLattice* L, const double* state, double* out);
/* One Boost.Odeint runge_kutta4 do_step (explicit_generic_rk, 4 stages). */
void   orc_rk4_step(const OracleLattice* L, double* state, double t, double dt);
/* Number of integrate_const steps and the k-th grid time (Boost semantics:
 * t_k = t0 + k*dt while t_k + dt - t1 <= eps). */
size_t orc_boost_num_steps(double t0, double t1, double dt);
/* molecular_dynamics_cpu with method "rk4" (lattice_md.cpp:341-453):
 * observer every save_interval integrate_const calls, records t, mags, E/N. */
size_t orc_integrate_md(const OracleLattice* L, double* state, double T0, double T1,
                        double dt, size_t save_interval, size_t max_rec,
                        double* times, double* mags, double* energy_per_site);
/* The reference GPU path's integrate_gpu (lattice_gpu.cu:688-702) with any
 * method of gpu::ode::step (integrator.cuh:91-458), restated operation by
 * operation: loop t += dt while t < T1 (integrator.cuh:470-487), observer
 * every save_interval steps and after the loop; records times, full states
 * [n_rec*N*D], mags [n_rec*3*D], E/N (any may be NULL).  An unknown method
 * runs ssprk53 (integrator.cuh:454-458) and the call returns 0 (1 when the
 * name is known).  *n_rec receives the number of observer calls. */
int    orc_integrate_gpu(const OracleLattice* L, double* state, double T0, double T1, double dt,
                         size_t save_interval, const char* method, size_t max_rec, size_t* n_rec,
                         double* times, double* states, double* mags, double* energy_per_site);
/* Dormand-Prince 5(4) (Boost.Odeint runge_kutta_dopri5, make_controlled(abs,
 * rel)) as integrate_ode_system uses it (lattice.h:4072-4085):
 *   control 0: fixed steps on the integrate_const grid (no error control,
 *              the reference GPU stepper's "dopri5", integrator.cuh:186-228);
 *   control 1: integrate_adaptive, observer after every accepted step
 *              (molecular_dynamics_cpu, lattice_md.cpp:452-453);
 *   control 2: integrate_const with the controlled stepper, observer on the
 *              grid T0 + k dt (pulse drivers, lattice_md.cpp:570-575); with
 *              n_seg > 0 one integrate_const call per row (t0, t1) of segs
 *              (W3 chunking, lattice_md.cpp:562-571).
 * Records as orc_integrate_md.  Returns the number of records, or -1 after
 * 500 rejections in a row. */
long   orc_integrate_dopri5(const OracleLattice* L, double* state, double T0, double T1,
                            double dt, int control, double abs_tol, double rel_tol,
                            size_t save_interval, const double* segs, size_t n_seg,
                            size_t max_rec, double* times, double* mags,
                            double* energy_per_site, size_t* n_accepted, size_t* n_rejected);
/* The same stepper on y' = lambda*y + cos(omega*t) (pinning tests):
 * y(T1), step counts, and for one step of size dt from (T0, y0) the embedded
 * error estimate and the propagated value in err_first_step[0..1]. */
long   orc_dopri5_scalar(double lambda, double omega, double y0, double T0, double T1,
                         double dt, int control, double abs_tol, double rel_tol,
                         double* y_end, size_t* n_accepted, size_t* n_rejected,
                         double* err_first_step);
/* single_pulse_drive / double_pulse_drive, CPU branch, method "rk4"
 * (lattice_md.cpp:493-586, :589-680).  Returns the number of records. */
size_t orc_single_pulse_drive(OracleLattice* L, const double* init, const double* field_in,
                              double t_B, double amp, double width, double freq,
                              double T0, double T1, double step, int chunking,
                              size_t max_rec, double* times, double* mags);
size_t orc_double_pulse_drive(OracleLattice* L, const double* init,
                              const double* field_in1, double t_B1,
                              const double* field_in2, double t_B2,
                              double amp, double width, double freq,
                              double T0, double T1, double step, int chunking,
                              size_t max_rec, double* times, double* mags);
/* max_dSdt_norm_no_drive (lattice_md.cpp:236-280). */
double orc_max_dsdt_no_drive(const OracleLattice* L, const double* state);
/* pump_probe_spectroscopy (lattice_md.cpp:683-926) minus HDF5: M0, M1(tau),
 * M01(tau) for every tau; outer OpenMP over tau when n_outer > 1 (W2).
 * Output blocks are [n_tau][max_rec][3*D].  Returns the number of tau points. */
int    orc_pump_probe(const OracleLattice* L, const double* ground, const double* field_in,
                      double amp, double width, double freq,
                      double tau_start, double tau_end, double tau_step,
                      double T0, double T1, double T_step,
                      int reuse_m0, double stationarity_tol, int n_outer, int chunking,
                      size_t max_rec, size_t max_tau, size_t* n_rec,
                      double* times, double* M0, double* M1, double* M01,
                      double* tau_values, int* used_w1);
/* build_pulse_segments (pulse_chunking.h:82-161): out = [t0, t1, dt_init]*n. */
size_t orc_build_pulse_segments(double T0, double T1, const double* centers, size_t n_centers,
                                double window, double T_step, double free_dt_factor,
                                size_t max_seg, double* out);
/* Lattice::greedy_quench (lattice.h:3352-3368): deterministic sweeps
 * (deterministic_sweep, lattice.h:3334-3349) setting S_i = -H_i/|H_i| *
 * spin_length (skipped when |H_i| < 1e-15) until |E - E_prev| <= rel_tol *
 * (|E_prev| + 1e-18) or max_sweeps.  The reference draws the sites of a
 * sweep from its Lehmer generator; here every sweep visits the N sites in
 * the given order (order[N], a permutation).  Returns the sweeps done. */
size_t orc_greedy_quench(const OracleLattice* L, double* state, const size_t* order, double spin_length,
                         double rel_tol, size_t max_sweeps);
/* OpenMP thread count actually available to the oracle. */
int    orc_max_threads(void);
/* Threads of the oracle's OpenMP regions from now on (n >= 1). */
void   orc_set_threads(int n);

/* Tangent-linear (forward-mode) RK4: propagates (state, dstate) together for
 * a directional derivative of the flow w.r.t. the initial state and a
 * parameter perturbation of the bond matrices dJ (per unit-cell bond,
 * [n_bonds*D*D]) and of the per-atom field dB ([n_atoms*D]).  Used only to
 * check adjoint gradients. */
size_t orc_tangent_md(OracleLattice* L, double* state, double* dstate,
                      const double* dJ, const double* dB,
                      double T0, double T1, double dt, size_t max_rec,
                      double* mags, double* dmags);

/* ---- mixed SU(2) + SU(3) lattice (MixedLattice) -------------------------
 * Expansion: build_sublattice per species (mixed_lattice.h:591-700) and the
 * SU(2)-SU(3) bilinears (mixed_lattice.h:773-800: J [3x8] at the SU(2)
 * source, J^T at the SU(3) partner).  State [N2*3 | N3*8]. */
typedef struct MixedOracle MixedOracle;
MixedOracle* orc_mixed_new(int d1, int d2, int d3,
                           int A2, const double* field2, const double* onsite2,
                           const double* frames2, const double* afm2,
                           int nb2, const int* src2, const int* par2, const int* off2, const double* J2,
                           int A3, const double* field3, const double* onsite3, const double* frames3,
                           int nb3, const int* src3, const int* par3, const int* off3, const double* J3,
                           int nbm, const int* srcm, const int* parm, const int* offm, const double* Jm);
void   orc_mixed_free(MixedOracle* M);
void   orc_mixed_sizes(const MixedOracle* M, size_t* N2, size_t* N3, size_t* nnz22,
                       size_t* nnz33, size_t* nnz23);
void   orc_mixed_tables(const MixedOracle* M, size_t* off22, size_t* par22, double* J22,
                        size_t* off33, size_t* par33, double* J33, size_t* off23,
                        size_t* par23, double* J23, size_t* off32, size_t* par32, double* J32);
/* SU(2) Gilbert alpha (mixed_lattice_md.cpp:570-578), SU(3) Bloch rates
 * gamma[8] towards eq [N3*8] (:596-605; NULL = zeros). */
void   orc_mixed_set_damping(MixedOracle* M, double alpha, const double* gamma8, const double* eq);
/* set_pulse_SU2 / set_pulse_SU3 (mixed_lattice_md.cpp:114-263; F^T B). */
void   orc_mixed_set_pulse(MixedOracle* M, int species, const double* field_in1, double t1,
                           const double* field_in2, double t2, double amp, double width,
                           double freq);
/* get_local_field_SU2/SU3_flat_into without drive (mixed_lattice_md.cpp:716-1105). */
void   orc_mixed_local_field(const MixedOracle* M, const double* x, int species, size_t site,
                             double* H);
/* MixedLattice::landau_lifshitz (mixed_lattice_md.cpp:433-605). */
void   orc_mixed_rhs(const MixedOracle* M, const double* x, double* dxdt, double t);
/* total_energy_flat (mixed_lattice.h:1783-1900). */
double orc_mixed_energy(const MixedOracle* M, const double* x);
/* MD observer record (mixed_lattice_md.cpp:1249-1281): [M2_anti, M2_local,
 * M2_global, M3_anti, M3_local, M3_global, E/(N2+N3)] = 34 doubles. */
void   orc_mixed_record(const MixedOracle* M, const double* x, double* out34);
/* Boost RK4 + integrate_const over [T0, T1]; records every step. */
size_t orc_mixed_integrate_md(const MixedOracle* M, double* state, double T0, double T1,
                              double dt, size_t max_rec, double* times, double* recs);
/* integrate_mixed_gpu (mixed_lattice_gpu.cu:469-488, :679-693) with any
 * method of gpu::ode::step, as orc_integrate_gpu; records of 34 doubles. */
int    orc_mixed_integrate_gpu(const MixedOracle* M, double* state, double T0, double T1, double dt,
                               size_t save_interval, const char* method, size_t max_rec, size_t* n_rec,
                               double* times, double* states, double* recs);
/* Dormand-Prince 5(4) over the flat mixed state with the controls of
 * orc_integrate_dopri5 (mixed_lattice.h:3960-4010); records of 34 doubles
 * every save_interval-th observer call.  Returns the number of records, or
 * -1 after 500 rejections in a row. */
long   orc_mixed_integrate_dopri5(const MixedOracle* M, double* state, double T0, double T1,
                                  double dt, int control, double abs_tol, double rel_tol,
                                  size_t save_interval, const double* segs, size_t n_seg,
                                  size_t max_rec, double* times, double* recs,
                                  size_t* n_accepted, size_t* n_rejected);

#ifdef __cplusplus
}
#endif
#endif
