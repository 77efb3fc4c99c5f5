/* csl_h5.h — HDF5 output of the spin-dynamics drivers (libcsl_h5.so).
 *
 * Writes the two files the reference's drivers produce, with the same group,
 * dataset and attribute names, shapes and HDF5 types, so that the reference's
 * readers (the util/readers_new scripts) open them unchanged:
 *
 *   pump_probe_spectroscopy.h5   HDF5PumpProbeWriter used by
 *                                Lattice::pump_probe_spectroscopy
 *                                (include/classical_spin/io/hdf5_io.h:729-1000,
 *                                 src/core/lattice_md.cpp:771-787)
 *   trajectory.h5                HDF5MDWriter used by molecular_dynamics
 *                                (include/classical_spin/io/hdf5_io.h:83-300,
 *                                 src/core/lattice_md.cpp:363-373)
 *
 * The library is separate from libcsl_hip.so so that the integrator carries
 * no HDF5 dependency; it links the HDF5 C library.  Arrays are row-major
 * float64 in the layouts the csl_gpu.h drivers return: magnetisation records
 * are [n_times][3][spin_dim] with the three observables in the order
 * M_antiferro, M_local, M_global; spin snapshots are [n][lattice_size][spin_dim]
 * in the reference's site order.
 *
 * Every function returns 0 on success and a negative code on error, with the
 * message available from csl_h5_last_error() (per thread).  A handle is used
 * by one thread at a time.
 */
#ifndef CSL_H5_H
#define CSL_H5_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CSL_H5_OK 0
#define CSL_H5_ERR_NULL (-1)
#define CSL_H5_ERR_ARG (-2)
#define CSL_H5_ERR_IO (-3)
#define CSL_H5_ERR_STATE (-4)

typedef struct csl_h5_file csl_h5_file;

/* Arguments of the HDF5PumpProbeWriter constructor (hdf5_io.h:729-750),
 * written as /metadata attributes by write_metadata (hdf5_io.h:855-915):
 * counts as NATIVE_HSIZE, reals as NATIVE_DOUBLE (spin_length rounded to
 * float first, as the reference passes it), integration_method and
 * T_zero_quench ("true"/"false") as strings, tau_steps as NATIVE_INT,
 * quench_sweeps only when T_zero_quench is set.  pulse_field_direction
 * [n_atoms][spin_dim] and positions [lattice_size][3] become /metadata
 * datasets when not NULL. */
typedef struct {
    size_t lattice_size, spin_dim, n_atoms, dim1, dim2, dim3;
    double spin_length;
    double pulse_amp, pulse_width, pulse_freq;
    double T_start, T_end, T_step;
    const char* integration_method;
    double tau_start, tau_end, tau_step;
    double ground_state_energy;
    double ground_state_magnetization; /* |M_local| of the ground state */
    double Temp_start, Temp_end;
    size_t n_anneal;
    int T_zero_quench;
    size_t quench_sweeps;
    const double* pulse_field_direction;
    const double* positions;
} csl_h5_pump_probe_meta;

const char* csl_h5_last_error(void);

/* Creates (truncates) the pump-probe file with groups /metadata,
 * /reference and /tau_scan, the metadata above and /tau_scan/tau_values
 * [n_tau], tau_i = tau_start + i * tau_step for
 * n_tau = int(|(tau_end - tau_start) / tau_step|) + 1 (hdf5_io.h:751-780);
 * *n_tau receives that count. */
int csl_h5_pump_probe_create(const char* path, const csl_h5_pump_probe_meta* meta,
                             size_t* n_tau, csl_h5_file** out);

/* /reference/times [n_times] and /reference/{M_antiferro, M_local,
 * M_global} [n_times][spin_dim] from M0 [n_times][3][spin_dim]
 * (write_reference_trajectory, hdf5_io.h:785-808).  Once per file. */
int csl_h5_pump_probe_write_reference(csl_h5_file* f, size_t n_times, const double* times,
                                      const double* M0);

/* Group /tau_scan/tau_<tau_idx> with attribute tau_value and the datasets
 * M1_{antiferro,local,global} and M01_{antiferro,local,global}
 * [n_times][spin_dim] from M1, M01 [n_times][3][spin_dim]
 * (write_tau_trajectory, hdf5_io.h:813-838).  Groups may be written in any
 * order, once each. */
int csl_h5_pump_probe_write_tau(csl_h5_file* f, size_t tau_idx, double tau_value,
                                size_t n_times, const double* M1, const double* M01);

/* Creates (truncates) an MD trajectory file (HDF5MDWriter constructor,
 * hdf5_io.h:84-195): /metadata attributes lattice_size, spin_dim, n_atoms,
 * save_interval (NATIVE_HSIZE), dt_initial, T_start, T_end, spin_length
 * (NATIVE_DOUBLE), integration_method, creation_time, code_version,
 * file_format (fixed-length NUL-terminated strings), datasets
 * /metadata/dimensions [3] and, when positions [lattice_size][3] is given,
 * /metadata/positions; and the extendible /trajectory datasets times [n],
 * magnetization_{antiferro,local,global} [n][spin_dim] and
 * spins [n][lattice_size][spin_dim], chunked as the reference chunks them
 * and compressed with deflate level `deflate` (the reference uses 6; 0
 * stores them uncompressed). */
int csl_h5_md_create(const char* path, size_t lattice_size, size_t spin_dim, size_t n_atoms,
                     size_t dim1, size_t dim2, size_t dim3, const char* method,
                     double dt_initial, double T_start, double T_end, size_t save_interval,
                     double spin_length, const double* positions, int deflate,
                     csl_h5_file** out);

/* Appends n_steps records (HDF5MDWriter::write_flat_step, hdf5_io.h:290-352,
 * for a block of steps at once): times [n_steps], mags [n_steps][3][spin_dim],
 * spins [n_steps][lattice_size][spin_dim]. */
int csl_h5_md_write_steps(csl_h5_file* f, size_t n_steps, const double* times,
                          const double* mags, const double* spins);

/* Number of records written so far to an MD file. */
int csl_h5_md_steps(csl_h5_file* f, size_t* n_steps);

/* Flushes and closes any handle (NULL is accepted). */
int csl_h5_close(csl_h5_file* f);

/* ---- read-back (used by tests and by callers that reopen a file) ----- */

/* Shape of dataset `name` in the file at `path`: *ndims (<= 4) and dims. */
int csl_h5_dataset_shape(const char* path, const char* name, int* ndims, size_t dims[4]);

/* Reads float64 dataset `name` into out (capacity `cap` doubles). */
int csl_h5_read_doubles(const char* path, const char* name, double* out, size_t cap);

/* Reads a scalar attribute `attr` of object `obj` (group or dataset) as a
 * double (floating or integer attributes) or, for string attributes, into
 * buf (capacity cap bytes, NUL-terminated). */
int csl_h5_read_attr_double(const char* path, const char* obj, const char* attr, double* value);
int csl_h5_read_attr_string(const char* path, const char* obj, const char* attr, char* buf, size_t cap);

#ifdef __cplusplus
}
#endif

#endif /* CSL_H5_H */
