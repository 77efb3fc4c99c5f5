// HDF5 writers for the pump-probe and MD trajectory files (include/csl_h5.h).
//
// Built with the HDF5 C API into libcsl_h5.so, separate from the HIP library.
// The file layouts follow include/classical_spin/io/hdf5_io.h:729-1000
// (HDF5PumpProbeWriter) and :83-300 (HDF5MDWriter): group, dataset and
// attribute names, ranks, HDF5 types, chunk shapes and deflate level.
// MD records are appended a block of steps at a time (one hyperslab write per
// dataset per block) rather than one step at a time.

#include "csl_h5.h"

#include <hdf5.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// Turns HDF5's own error stack printing off once; errors come back as codes.
void quiet_hdf5() {
    static bool done = false;
    if (!done) {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        done = true;
    }
}

// Closes an HDF5 identifier of any kind when it leaves scope.
struct Id {
    hid_t id = -1;
    Id() = default;
    explicit Id(hid_t v) : id(v) {}
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
    Id(Id&& o) noexcept : id(o.id) { o.id = -1; }
    Id& operator=(Id&& o) noexcept {
        if (this != &o) { reset(); id = o.id; o.id = -1; }
        return *this;
    }
    ~Id() { reset(); }
    void reset() {
        if (id >= 0) {
            switch (H5Iget_type(id)) {
                case H5I_FILE: H5Fclose(id); break;
                case H5I_GROUP: H5Gclose(id); break;
                case H5I_DATASET: H5Dclose(id); break;
                case H5I_DATASPACE: H5Sclose(id); break;
                case H5I_DATATYPE: H5Tclose(id); break;
                case H5I_ATTR: H5Aclose(id); break;
                case H5I_GENPROP_LST: H5Pclose(id); break;
                default: break;
            }
        }
        id = -1;
    }
    bool ok() const { return id >= 0; }
    operator hid_t() const { return id; }
};

bool attr_scalar(hid_t obj, const char* name, hid_t type, const void* value) {
    Id space(H5Screate(H5S_SCALAR));
    if (!space.ok()) return false;
    Id attr(H5Acreate2(obj, name, type, space, H5P_DEFAULT, H5P_DEFAULT));
    return attr.ok() && H5Awrite(attr, type, value) >= 0;
}

bool attr_double(hid_t obj, const char* name, double v) {
    return attr_scalar(obj, name, H5T_NATIVE_DOUBLE, &v);
}

bool attr_hsize(hid_t obj, const char* name, size_t v) {
    hsize_t h = static_cast<hsize_t>(v);
    return attr_scalar(obj, name, H5T_NATIVE_HSIZE, &h);
}

// Fixed-length C string of size+1 bytes (hdf5_io.h:391-396).
bool attr_string(hid_t obj, const char* name, const std::string& s) {
    Id type(H5Tcopy(H5T_C_S1));
    if (!type.ok() || H5Tset_size(type, s.size() + 1) < 0) return false;
    return attr_scalar(obj, name, type, s.c_str());
}

// Contiguous float64 dataset of the given shape, written whole.
bool write_dataset(hid_t grp, const char* name, int rank, const hsize_t* dims, const double* data) {
    Id space(H5Screate_simple(rank, dims, nullptr));
    if (!space.ok()) return false;
    Id ds(H5Dcreate2(grp, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
    return ds.ok() && H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
}

// Column `m` of [n][3][d] records as a dense [n][d] block.
std::vector<double> take_mag(const double* rec, size_t n, size_t d, int m) {
    std::vector<double> out(n * d);
    for (size_t t = 0; t < n; ++t)
        for (size_t c = 0; c < d; ++c) out[t * d + c] = rec[(t * 3 + m) * d + c];
    return out;
}

const char* const kMagNames[3] = {"antiferro", "local", "global"};

}  // namespace

struct csl_h5_file {
    enum Kind { PumpProbe, Trajectory } kind;
    Id file;
    size_t spin_dim = 0;
    size_t lattice_size = 0;
    // trajectory datasets: times, mag x3, spins
    Id ds_times, ds_mag[3], ds_spins;
    size_t steps = 0;
    std::vector<bool> tau_written;
    bool reference_written = false;
};

extern "C" {

const char* csl_h5_last_error(void) { return g_err.c_str(); }

int csl_h5_pump_probe_create(const char* path, const csl_h5_pump_probe_meta* meta,
                             size_t* n_tau_out, csl_h5_file** out) {
    if (!path || !meta || !out || !meta->integration_method)
        return fail(CSL_H5_ERR_NULL, "csl_h5_pump_probe_create: null argument");
    if (meta->spin_dim == 0) return fail(CSL_H5_ERR_ARG, "csl_h5_pump_probe_create: spin_dim must be > 0");
    if (!(meta->tau_step != 0.0) || !std::isfinite(meta->tau_start) || !std::isfinite(meta->tau_end))
        return fail(CSL_H5_ERR_ARG, "csl_h5_pump_probe_create: tau_step must be non-zero and the range finite");
    const double span = std::fabs((meta->tau_end - meta->tau_start) / meta->tau_step);
    if (!(span < 2147483646.0)) return fail(CSL_H5_ERR_ARG, "csl_h5_pump_probe_create: too many delays");
    // hdf5_io.h:769-773 and :905-906: int(|(tau_end - tau_start) / tau_step|) + 1
    const int n_tau = static_cast<int>(span) + 1;
    quiet_hdf5();
    auto* f = new csl_h5_file{};
    f->kind = csl_h5_file::PumpProbe;
    f->spin_dim = meta->spin_dim;
    f->tau_written.assign(static_cast<size_t>(n_tau), false);
    f->file = Id(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT));
    if (!f->file.ok()) {
        delete f;
        return fail(CSL_H5_ERR_IO, std::string("cannot create ") + path);
    }
    Id md(H5Gcreate2(f->file, "/metadata", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
    Id ref(H5Gcreate2(f->file, "/reference", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
    Id scan(H5Gcreate2(f->file, "/tau_scan", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
    char stamp[64] = {0};
    const std::time_t now = std::time(nullptr);
    std::strftime(stamp, sizeof(stamp), "%Y-%m-%dT%H:%M:%S", std::localtime(&now));
    const int tau_steps = n_tau;
    bool ok = md.ok() && ref.ok() && scan.ok();
    // attribute order of write_metadata (hdf5_io.h:866-910)
    ok = ok && attr_string(md, "creation_time", stamp) &&
         attr_string(md, "experiment_type", "pump_probe_spectroscopy") &&
         attr_string(md, "code_version", "ClassicalSpin_Cpp v1.0") &&
         attr_string(md, "file_format", "HDF5_PumpProbe_v1.0") &&
         attr_hsize(md, "lattice_size", meta->lattice_size) && attr_hsize(md, "spin_dim", meta->spin_dim) &&
         attr_hsize(md, "n_atoms", meta->n_atoms) && attr_hsize(md, "dim1", meta->dim1) &&
         attr_hsize(md, "dim2", meta->dim2) && attr_hsize(md, "dim3", meta->dim3) &&
         attr_double(md, "spin_length", static_cast<double>(static_cast<float>(meta->spin_length))) &&
         attr_double(md, "pulse_amp", meta->pulse_amp) && attr_double(md, "pulse_width", meta->pulse_width) &&
         attr_double(md, "pulse_freq", meta->pulse_freq) && attr_double(md, "T_start", meta->T_start) &&
         attr_double(md, "T_end", meta->T_end) && attr_double(md, "T_step", meta->T_step) &&
         attr_string(md, "integration_method", meta->integration_method) &&
         attr_double(md, "tau_start", meta->tau_start) && attr_double(md, "tau_end", meta->tau_end) &&
         attr_double(md, "tau_step", meta->tau_step) &&
         attr_scalar(md, "tau_steps", H5T_NATIVE_INT, &tau_steps) &&
         attr_double(md, "ground_state_energy", meta->ground_state_energy) &&
         attr_double(md, "ground_state_magnetization", meta->ground_state_magnetization) &&
         attr_double(md, "Temp_start", meta->Temp_start) && attr_double(md, "Temp_end", meta->Temp_end) &&
         attr_hsize(md, "n_anneal", meta->n_anneal) &&
         attr_string(md, "T_zero_quench", meta->T_zero_quench ? "true" : "false");
    if (ok && meta->T_zero_quench) ok = attr_hsize(md, "quench_sweeps", meta->quench_sweeps);
    if (ok && meta->pulse_field_direction && meta->n_atoms) {
        const hsize_t d[2] = {meta->n_atoms, meta->spin_dim};
        ok = write_dataset(md, "pulse_field_direction", 2, d, meta->pulse_field_direction);
    }
    if (ok && meta->positions && meta->lattice_size) {
        const hsize_t d[2] = {meta->lattice_size, 3};
        ok = write_dataset(md, "positions", 2, d, meta->positions);
    }
    std::vector<double> taus(static_cast<size_t>(n_tau));
    for (int i = 0; i < n_tau; ++i) taus[static_cast<size_t>(i)] = meta->tau_start + i * meta->tau_step;
    const hsize_t tdim[1] = {static_cast<hsize_t>(n_tau)};
    ok = ok && write_dataset(scan, "tau_values", 1, tdim, taus.data());
    if (!ok) {
        delete f;
        return fail(CSL_H5_ERR_IO, std::string("csl_h5_pump_probe_create: HDF5 write failed for ") + path);
    }
    if (n_tau_out) *n_tau_out = static_cast<size_t>(n_tau);
    *out = f;
    return CSL_H5_OK;
}

int csl_h5_pump_probe_write_reference(csl_h5_file* f, size_t n_times, const double* times,
                                      const double* M0) {
    if (!f || (n_times && (!times || !M0)))
        return fail(CSL_H5_ERR_NULL, "csl_h5_pump_probe_write_reference: null argument");
    if (f->kind != csl_h5_file::PumpProbe || !f->file.ok())
        return fail(CSL_H5_ERR_STATE, "csl_h5_pump_probe_write_reference: not an open pump-probe file");
    if (f->reference_written)
        return fail(CSL_H5_ERR_STATE, "csl_h5_pump_probe_write_reference: reference already written");
    Id ref(H5Gopen2(f->file, "/reference", H5P_DEFAULT));
    const hsize_t ndim[1] = {static_cast<hsize_t>(n_times)};
    bool ok = ref.ok() && write_dataset(ref, "times", 1, ndim, times);
    const hsize_t mdim[2] = {static_cast<hsize_t>(n_times), static_cast<hsize_t>(f->spin_dim)};
    for (int m = 0; m < 3 && ok; ++m) {
        std::vector<double> col = take_mag(M0, n_times, f->spin_dim, m);
        ok = write_dataset(ref, (std::string("M_") + kMagNames[m]).c_str(), 2, mdim, col.data());
    }
    if (!ok) return fail(CSL_H5_ERR_IO, "csl_h5_pump_probe_write_reference: HDF5 write failed");
    f->reference_written = true;
    return CSL_H5_OK;
}

int csl_h5_pump_probe_write_tau(csl_h5_file* f, size_t tau_idx, double tau_value,
                                size_t n_times, const double* M1, const double* M01) {
    if (!f || (n_times && (!M1 || !M01))) return fail(CSL_H5_ERR_NULL, "csl_h5_pump_probe_write_tau: null argument");
    if (f->kind != csl_h5_file::PumpProbe || !f->file.ok())
        return fail(CSL_H5_ERR_STATE, "csl_h5_pump_probe_write_tau: not an open pump-probe file");
    if (tau_idx >= f->tau_written.size())
        return fail(CSL_H5_ERR_ARG, "csl_h5_pump_probe_write_tau: tau_idx " + std::to_string(tau_idx) +
                                        " >= n_tau " + std::to_string(f->tau_written.size()));
    if (f->tau_written[tau_idx])
        return fail(CSL_H5_ERR_STATE, "csl_h5_pump_probe_write_tau: tau_" + std::to_string(tau_idx) + " already written");
    const std::string name = "/tau_scan/tau_" + std::to_string(tau_idx);
    Id grp(H5Gcreate2(f->file, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
    bool ok = grp.ok() && attr_double(grp, "tau_value", tau_value);
    const hsize_t dims[2] = {static_cast<hsize_t>(n_times), static_cast<hsize_t>(f->spin_dim)};
    for (int which = 0; which < 2 && ok; ++which) {
        const double* src = which == 0 ? M1 : M01;
        for (int m = 0; m < 3 && ok; ++m) {
            std::vector<double> col = take_mag(src, n_times, f->spin_dim, m);
            const std::string ds = std::string(which == 0 ? "M1_" : "M01_") + kMagNames[m];
            ok = write_dataset(grp, ds.c_str(), 2, dims, col.data());
        }
    }
    if (!ok) return fail(CSL_H5_ERR_IO, "csl_h5_pump_probe_write_tau: HDF5 write failed for " + name);
    f->tau_written[tau_idx] = true;
    return CSL_H5_OK;
}

int csl_h5_md_create(const char* path, size_t lattice_size, size_t spin_dim, size_t n_atoms,
                     size_t dim1, size_t dim2, size_t dim3, const char* method,
                     double dt_initial, double T_start, double T_end, size_t save_interval,
                     double spin_length, const double* positions, int deflate,
                     csl_h5_file** out) {
    if (!path || !method || !out) return fail(CSL_H5_ERR_NULL, "csl_h5_md_create: null argument");
    if (lattice_size == 0 || spin_dim == 0)
        return fail(CSL_H5_ERR_ARG, "csl_h5_md_create: lattice_size and spin_dim must be > 0");
    if (deflate < 0 || deflate > 9) return fail(CSL_H5_ERR_ARG, "csl_h5_md_create: deflate level must be 0..9");
    quiet_hdf5();
    auto* f = new csl_h5_file{};
    f->kind = csl_h5_file::Trajectory;
    f->spin_dim = spin_dim;
    f->lattice_size = lattice_size;
    f->file = Id(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT));
    if (!f->file.ok()) {
        delete f;
        return fail(CSL_H5_ERR_IO, std::string("cannot create ") + path);
    }
    Id traj(H5Gcreate2(f->file, "/trajectory", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
    Id md(H5Gcreate2(f->file, "/metadata", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
    char stamp[64] = {0};
    const std::time_t now = std::time(nullptr);
    std::strftime(stamp, sizeof(stamp), "%Y-%m-%dT%H:%M:%S", std::localtime(&now));
    bool ok = traj.ok() && md.ok();
    ok = ok && attr_hsize(md, "lattice_size", lattice_size) && attr_hsize(md, "spin_dim", spin_dim) &&
         attr_hsize(md, "n_atoms", n_atoms) && attr_string(md, "integration_method", method) &&
         attr_double(md, "dt_initial", dt_initial) && attr_double(md, "T_start", T_start) &&
         attr_double(md, "T_end", T_end) && attr_hsize(md, "save_interval", save_interval) &&
         // the reference passes spin_length as float and stores it as double
         attr_double(md, "spin_length", static_cast<double>(static_cast<float>(spin_length))) &&
         attr_string(md, "creation_time", stamp) &&
         attr_string(md, "code_version", "ClassicalSpin_Cpp v1.0") &&
         attr_string(md, "file_format", "HDF5_MD_v1.0");
    if (ok) {
        const hsize_t d3[1] = {3};
        const hsize_t dims[3] = {dim1, dim2, dim3};
        Id space(H5Screate_simple(1, d3, nullptr));
        Id ds(H5Dcreate2(md, "dimensions", H5T_NATIVE_HSIZE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
        ok = ds.ok() && H5Dwrite(ds, H5T_NATIVE_HSIZE, H5S_ALL, H5S_ALL, H5P_DEFAULT, dims) >= 0;
    }
    if (ok && positions) {
        const hsize_t pd[2] = {lattice_size, 3};
        ok = write_dataset(md, "positions", 2, pd, positions);
    }
    // extendible datasets, chunks as hdf5_io.h:154-192
    auto make_ext = [&](const char* name, int rank, const hsize_t* chunk, Id& dst) {
        hsize_t cur[3] = {0, chunk[1], chunk[2]};
        hsize_t mx[3] = {H5S_UNLIMITED, chunk[1], chunk[2]};
        Id space(H5Screate_simple(rank, cur, mx));
        Id plist(H5Pcreate(H5P_DATASET_CREATE));
        if (!space.ok() || !plist.ok() || H5Pset_chunk(plist, rank, chunk) < 0) return false;
        if (deflate > 0 && H5Pset_deflate(plist, static_cast<unsigned>(deflate)) < 0) return false;
        dst = Id(H5Dcreate2(traj, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, plist, H5P_DEFAULT));
        return dst.ok();
    };
    const hsize_t c_time[3] = {1000, 0, 0};
    const hsize_t c_mag[3] = {100, spin_dim, 0};
    const hsize_t c_spin[3] = {1, lattice_size, spin_dim};
    ok = ok && make_ext("times", 1, c_time, f->ds_times);
    for (int m = 0; m < 3 && ok; ++m)
        ok = make_ext((std::string("magnetization_") + kMagNames[m]).c_str(), 2, c_mag, f->ds_mag[m]);
    ok = ok && make_ext("spins", 3, c_spin, f->ds_spins);
    if (!ok) {
        delete f;
        return fail(CSL_H5_ERR_IO, std::string("csl_h5_md_create: HDF5 write failed for ") + path);
    }
    *out = f;
    return CSL_H5_OK;
}

namespace {

// Extends `ds` along dimension 0 to `total` rows and writes `n` rows of
// `row` elements at row offset `at`.
bool append_rows(hid_t ds, int rank, size_t at, size_t n, const hsize_t* tail, const double* data) {
    hsize_t size[3] = {at + n, tail[0], tail[1]};
    if (H5Dset_extent(ds, size) < 0) return false;
    Id fspace(H5Dget_space(ds));
    hsize_t off[3] = {at, 0, 0};
    hsize_t cnt[3] = {n, tail[0], tail[1]};
    if (!fspace.ok() || H5Sselect_hyperslab(fspace, H5S_SELECT_SET, off, nullptr, cnt, nullptr) < 0) return false;
    Id mspace(H5Screate_simple(rank, cnt, nullptr));
    return mspace.ok() && H5Dwrite(ds, H5T_NATIVE_DOUBLE, mspace, fspace, H5P_DEFAULT, data) >= 0;
}

}  // namespace

int csl_h5_md_write_steps(csl_h5_file* f, size_t n_steps, const double* times,
                          const double* mags, const double* spins) {
    if (!f) return fail(CSL_H5_ERR_NULL, "csl_h5_md_write_steps: null handle");
    if (f->kind != csl_h5_file::Trajectory || !f->file.ok())
        return fail(CSL_H5_ERR_STATE, "csl_h5_md_write_steps: not an open trajectory file");
    if (n_steps == 0) return CSL_H5_OK;
    if (!times || !mags || !spins) return fail(CSL_H5_ERR_NULL, "csl_h5_md_write_steps: null array");
    const size_t d = f->spin_dim;
    const hsize_t t_time[2] = {0, 0};
    const hsize_t t_mag[2] = {d, 0};
    const hsize_t t_spin[2] = {f->lattice_size, d};
    bool ok = append_rows(f->ds_times, 1, f->steps, n_steps, t_time, times);
    for (int m = 0; m < 3 && ok; ++m) {
        std::vector<double> col = take_mag(mags, n_steps, d, m);
        ok = append_rows(f->ds_mag[m], 2, f->steps, n_steps, t_mag, col.data());
    }
    ok = ok && append_rows(f->ds_spins, 3, f->steps, n_steps, t_spin, spins);
    if (!ok) return fail(CSL_H5_ERR_IO, "csl_h5_md_write_steps: HDF5 write failed");
    f->steps += n_steps;
    return CSL_H5_OK;
}

int csl_h5_md_steps(csl_h5_file* f, size_t* n_steps) {
    if (!f || !n_steps) return fail(CSL_H5_ERR_NULL, "csl_h5_md_steps: null argument");
    if (f->kind != csl_h5_file::Trajectory) return fail(CSL_H5_ERR_STATE, "csl_h5_md_steps: not a trajectory file");
    *n_steps = f->steps;
    return CSL_H5_OK;
}

int csl_h5_close(csl_h5_file* f) {
    if (!f) return CSL_H5_OK;
    bool ok = true;
    if (f->file.ok()) ok = H5Fflush(f->file, H5F_SCOPE_LOCAL) >= 0;
    delete f;  // closes datasets before the file
    return ok ? CSL_H5_OK : fail(CSL_H5_ERR_IO, "csl_h5_close: flush failed");
}

int csl_h5_dataset_shape(const char* path, const char* name, int* ndims, size_t dims[4]) {
    if (!path || !name || !ndims || !dims) return fail(CSL_H5_ERR_NULL, "csl_h5_dataset_shape: null argument");
    quiet_hdf5();
    Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT));
    if (!file.ok()) return fail(CSL_H5_ERR_IO, std::string("cannot open ") + path);
    Id ds(H5Dopen2(file, name, H5P_DEFAULT));
    if (!ds.ok()) return fail(CSL_H5_ERR_ARG, std::string("no dataset ") + name);
    Id space(H5Dget_space(ds));
    const int r = H5Sget_simple_extent_ndims(space);
    if (r < 0 || r > 4) return fail(CSL_H5_ERR_ARG, std::string("unsupported rank for ") + name);
    hsize_t h[4] = {0, 0, 0, 0};
    H5Sget_simple_extent_dims(space, h, nullptr);
    *ndims = r;
    for (int i = 0; i < 4; ++i) dims[i] = static_cast<size_t>(h[i]);
    return CSL_H5_OK;
}

int csl_h5_read_doubles(const char* path, const char* name, double* out, size_t cap) {
    if (!path || !name || !out) return fail(CSL_H5_ERR_NULL, "csl_h5_read_doubles: null argument");
    quiet_hdf5();
    Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT));
    if (!file.ok()) return fail(CSL_H5_ERR_IO, std::string("cannot open ") + path);
    Id ds(H5Dopen2(file, name, H5P_DEFAULT));
    if (!ds.ok()) return fail(CSL_H5_ERR_ARG, std::string("no dataset ") + name);
    Id space(H5Dget_space(ds));
    const hssize_t n = H5Sget_simple_extent_npoints(space);
    if (n < 0 || static_cast<size_t>(n) > cap)
        return fail(CSL_H5_ERR_ARG, std::string("buffer too small for ") + name);
    if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
        return fail(CSL_H5_ERR_IO, std::string("read failed for ") + name);
    return CSL_H5_OK;
}

int csl_h5_read_attr_double(const char* path, const char* obj, const char* attr, double* value) {
    if (!path || !obj || !attr || !value) return fail(CSL_H5_ERR_NULL, "csl_h5_read_attr_double: null argument");
    quiet_hdf5();
    Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT));
    if (!file.ok()) return fail(CSL_H5_ERR_IO, std::string("cannot open ") + path);
    Id a(H5Aopen_by_name(file, obj, attr, H5P_DEFAULT, H5P_DEFAULT));
    if (!a.ok()) return fail(CSL_H5_ERR_ARG, std::string("no attribute ") + obj + ":" + attr);
    Id type(H5Aget_type(a));
    const H5T_class_t cls = H5Tget_class(type);
    if (cls == H5T_FLOAT) {
        if (H5Aread(a, H5T_NATIVE_DOUBLE, value) < 0) return fail(CSL_H5_ERR_IO, "attribute read failed");
    } else if (cls == H5T_INTEGER) {
        hsize_t v = 0;
        if (H5Aread(a, H5T_NATIVE_HSIZE, &v) < 0) return fail(CSL_H5_ERR_IO, "attribute read failed");
        *value = static_cast<double>(v);
    } else {
        return fail(CSL_H5_ERR_ARG, std::string("attribute is not numeric: ") + attr);
    }
    return CSL_H5_OK;
}

int csl_h5_read_attr_string(const char* path, const char* obj, const char* attr, char* buf, size_t cap) {
    if (!path || !obj || !attr || !buf || cap == 0)
        return fail(CSL_H5_ERR_NULL, "csl_h5_read_attr_string: null argument");
    quiet_hdf5();
    Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT));
    if (!file.ok()) return fail(CSL_H5_ERR_IO, std::string("cannot open ") + path);
    Id a(H5Aopen_by_name(file, obj, attr, H5P_DEFAULT, H5P_DEFAULT));
    if (!a.ok()) return fail(CSL_H5_ERR_ARG, std::string("no attribute ") + obj + ":" + attr);
    Id type(H5Aget_type(a));
    if (H5Tget_class(type) != H5T_STRING || H5Tis_variable_str(type) > 0)
        return fail(CSL_H5_ERR_ARG, std::string("attribute is not a fixed-length string: ") + attr);
    const size_t sz = H5Tget_size(type);
    if (sz > cap) return fail(CSL_H5_ERR_ARG, std::string("buffer too small for ") + attr);
    std::vector<char> tmp(sz + 1, 0);
    if (H5Aread(a, type, tmp.data()) < 0) return fail(CSL_H5_ERR_IO, "attribute read failed");
    std::memcpy(buf, tmp.data(), sz);
    buf[sz < cap ? sz : cap - 1] = 0;
    return CSL_H5_OK;
}

}  // extern "C"
