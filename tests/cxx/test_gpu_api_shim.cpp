// Drives the gpu:: drop-in API (include/classical_spin_amd/lattice_gpu_api.h)
// the way the reference's Lattice glue does (lattice.h:5429-5522, 5527-5602,
// 5617-5750): padded tables, pulse upload, spins, integrate_gpu with full
// state records (rk4, the glue's default dopri5 and the API default
// ssprk53), step_gpu, integrate_gpu_batched with tau2 per replica.
//
// Inputs (binary, little-endian doubles / size_t) are written by
// tests/test_cxx_shim.py; outputs are written back for comparison with the
// oracle.  Usage: test_gpu_api_shim <in.bin> <out.bin>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "classical_spin_amd/lattice_gpu_api.h"

namespace {

template <class T>
std::vector<T> read_vec(FILE* f) {
    unsigned long long n = 0;
    if (std::fread(&n, sizeof(n), 1, f) != 1) throw std::runtime_error("short read");
    std::vector<T> v(n);
    if (n && std::fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("short read");
    return v;
}

template <class T>
void write_vec(FILE* f, const std::vector<T>& v) {
    unsigned long long n = v.size();
    std::fwrite(&n, sizeof(n), 1, f);
    if (n) std::fwrite(v.data(), sizeof(T), n, f);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    try {
        FILE* in = std::fopen(argv[1], "rb");
        if (!in) throw std::runtime_error("cannot open input");
        auto dims = read_vec<size_t>(in);        // N, D, N_atoms, max_bi
        auto field = read_vec<double>(in);
        auto onsite = read_vec<double>(in);
        auto bil = read_vec<double>(in);
        auto partners = read_vec<size_t>(in);
        auto counts = read_vec<size_t>(in);
        auto drive = read_vec<double>(in);       // [2*N_atoms*D]
        auto pulse = read_vec<double>(in);       // amp, width, freq, t1, t2
        auto spins = read_vec<double>(in);
        auto tgrid = read_vec<double>(in);       // T0, T1, dt
        auto tau2 = read_vec<double>(in);
        auto afm = read_vec<double>(in);
        auto frames = read_vec<double>(in);
        std::fclose(in);

        gpu::GPULatticeDataHandle* h = gpu::create_gpu_lattice_data(dims[0], dims[1], dims[2], dims[3], field,
                                                                    onsite, bil, partners, counts);
        gpu::set_gpu_pulse(h, drive, pulse[0], pulse[1], pulse[2], pulse[3], pulse[4]);
        gpu::set_gpu_spins(h, spins);
        const double E0 = gpu::compute_energy_gpu(h);

        std::vector<std::pair<double, std::vector<double>>> traj;
        gpu::integrate_gpu(h, tgrid[0], tgrid[1], tgrid[2], 10, traj, "rk4");
        std::vector<double> t_out, s_out;
        for (auto& r : traj) {
            t_out.push_back(r.first);
            s_out.insert(s_out.end(), r.second.begin(), r.second.end());
        }
        std::vector<double> ml, ms;
        gpu::compute_magnetization_gpu(h, ml, ms);

        // one step_gpu with the Dormand-Prince tableau from the state the
        // integration left on the device
        gpu::step_gpu(h, 0.0, 0.01, "dopri5");
        std::vector<double> after_step;
        gpu::get_gpu_spins(h, after_step);

        // molecular_dynamics_gpu with the glue's default method "dopri5"
        // (spin_config.h md_integrator; lattice.h:5527-5602)
        gpu::set_gpu_spins(h, spins);
        std::vector<std::pair<double, std::vector<double>>> traj_dp;
        gpu::integrate_gpu(h, tgrid[0], tgrid[1], tgrid[2], 10, traj_dp, "dopri5");
        std::vector<double> t_dp, s_dp;
        for (auto& r : traj_dp) {
            t_dp.push_back(r.first);
            s_dp.insert(s_dp.end(), r.second.begin(), r.second.end());
        }
        // the API default (ssprk53) with save_interval 25
        gpu::set_gpu_spins(h, spins);
        std::vector<std::pair<double, std::vector<double>>> traj_ss;
        gpu::integrate_gpu(h, tgrid[0], tgrid[1], tgrid[2], 25, traj_ss);
        std::vector<double> s_ss;
        for (auto& r : traj_ss) s_ss.insert(s_ss.end(), r.second.begin(), r.second.end());

        // batched mode with a non-rk4 method warns and runs rk4 (lattice_gpu.cu:779-784)
        gpu::BatchedMagResult bm = gpu::integrate_gpu_batched(h, spins, tau2, afm, frames, tgrid[0], tgrid[1],
                                                              tgrid[2], 1, "dopri5");
        gpu::destroy_gpu_lattice_data(h);
        gpu::destroy_gpu_lattice_data(nullptr);  // reference: null is a no-op

        FILE* out = std::fopen(argv[2], "wb");
        if (!out) throw std::runtime_error("cannot open output");
        write_vec(out, std::vector<double>{E0, 0.0});
        write_vec(out, t_out);
        write_vec(out, s_out);
        write_vec(out, ml);
        write_vec(out, ms);
        write_vec(out, std::vector<double>{double(bm.B), double(bm.spin_dim), double(bm.n_time_points)});
        write_vec(out, bm.times);
        write_vec(out, bm.mag_data);
        write_vec(out, after_step);
        write_vec(out, t_dp);
        write_vec(out, s_dp);
        write_vec(out, s_ss);
        std::fclose(out);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "test_gpu_api_shim: %s\n", e.what());
        return 1;
    }
    return 0;
}
