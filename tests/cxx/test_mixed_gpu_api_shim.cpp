// Drives the mixed_gpu:: drop-in API
// (include/classical_spin_amd/mixed_lattice_gpu_api.h) the way the
// reference's MixedLattice glue does (mixed_lattice.h:6160-6360, 6400-6520):
// tables padded to the per-species maximum (mixed partners SIZE_MAX beyond a
// site's count), both pulses, spins, integrate_mixed_gpu with the glue's
// default method "dopri5" and with "rk4", step_mixed_gpu, magnetisations,
// normalisation and the dimension query.
//
// Inputs and outputs are binary vectors written / read by
// tests/test_cxx_shim.py.  Usage: test_mixed_gpu_api_shim <in.bin> <out.bin>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "classical_spin/gpu/mixed_lattice_gpu_api.h"

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

void flatten(const std::vector<std::pair<double, std::vector<double>>>& traj, std::vector<double>& t,
             std::vector<double>& s) {
    for (const auto& r : traj) {
        t.push_back(r.first);
        s.insert(s.end(), r.second.begin(), r.second.end());
    }
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
        auto dims = read_vec<size_t>(in);  // N2, A2, N3, A3, max22, max33, max23, max32
        auto field2 = read_vec<double>(in);
        auto onsite2 = read_vec<double>(in);
        auto bil2 = read_vec<double>(in);
        auto par2 = read_vec<size_t>(in);
        auto cnt2 = read_vec<size_t>(in);
        auto field3 = read_vec<double>(in);
        auto onsite3 = read_vec<double>(in);
        auto bil3 = read_vec<double>(in);
        auto par3 = read_vec<size_t>(in);
        auto cnt3 = read_vec<size_t>(in);
        auto mixed = read_vec<double>(in);
        auto mixed_src2 = read_vec<size_t>(in);
        auto mixed_par3 = read_vec<size_t>(in);
        auto mixed_cnt2 = read_vec<size_t>(in);
        auto mixed3 = read_vec<double>(in);
        auto mixed_par2_from3 = read_vec<size_t>(in);
        auto mixed_cnt3 = read_vec<size_t>(in);
        auto drive2 = read_vec<double>(in);
        auto pulse2 = read_vec<double>(in);  // amp, width, freq, t1, t2
        auto drive3 = read_vec<double>(in);
        auto pulse3 = read_vec<double>(in);
        auto spins = read_vec<double>(in);
        auto tg = read_vec<double>(in);  // T0, T1, dt
        std::fclose(in);

        auto* h = mixed_gpu::create_gpu_mixed_lattice_data(
            dims[0], 3, dims[1], dims[2], 8, dims[3], dims[4], dims[5], dims[6], dims[7], field2, onsite2, bil2,
            par2, cnt2, field3, onsite3, bil3, par3, cnt3, mixed, mixed_src2, mixed_par3, mixed_cnt2, mixed3,
            mixed_par2_from3, mixed_cnt3);
        mixed_gpu::set_gpu_pulse_SU2(h, drive2, pulse2[0], pulse2[1], pulse2[2], pulse2[3], pulse2[4]);
        mixed_gpu::set_gpu_pulse_SU3(h, drive3, pulse3[0], pulse3[1], pulse3[2], pulse3[3], pulse3[4]);
        mixed_gpu::set_gpu_mixed_spins(h, spins);
        std::vector<double> back;
        mixed_gpu::get_gpu_mixed_spins(h, back);

        // molecular_dynamics_gpu default method (mixed_lattice.h:6306-6344)
        std::vector<std::pair<double, std::vector<double>>> traj_dp, traj_rk;
        mixed_gpu::integrate_mixed_gpu(h, tg[0], tg[1], tg[2], 5, traj_dp, "dopri5");
        std::vector<double> t_dp, s_dp, t_rk, s_rk;
        flatten(traj_dp, t_dp, s_dp);
        mixed_gpu::set_gpu_mixed_spins(h, spins);
        mixed_gpu::integrate_mixed_gpu(h, tg[0], tg[1], tg[2], 5, traj_rk, "rk4");
        flatten(traj_rk, t_rk, s_rk);

        // one step_mixed_gpu with ssprk53 from the initial state at T0
        mixed_gpu::set_gpu_mixed_spins(h, spins);
        mixed_gpu::step_mixed_gpu(h, tg[0], tg[2], "ssprk53");
        std::vector<double> stepped;
        mixed_gpu::get_gpu_mixed_spins(h, stepped);
        std::vector<double> m2, s2, m3, s3;
        mixed_gpu::compute_magnetization_mixed_gpu(h, m2, s2, m3, s3);
        mixed_gpu::normalize_spins_mixed_gpu(h, 1.0, 1.0);
        std::vector<double> normed;
        mixed_gpu::get_gpu_mixed_spins(h, normed);
        size_t N2 = 0, D2 = 0, A2 = 0, N3 = 0, D3 = 0, A3 = 0;
        mixed_gpu::get_mixed_lattice_dims(h, N2, D2, A2, N3, D3, A3);
        mixed_gpu::destroy_gpu_mixed_lattice_data(h);
        mixed_gpu::destroy_gpu_mixed_lattice_data(nullptr);  // null is a no-op

        FILE* out = std::fopen(argv[2], "wb");
        if (!out) throw std::runtime_error("cannot open output");
        write_vec(out, back);
        write_vec(out, t_dp);
        write_vec(out, s_dp);
        write_vec(out, t_rk);
        write_vec(out, s_rk);
        write_vec(out, stepped);
        write_vec(out, m2);
        write_vec(out, s2);
        write_vec(out, m3);
        write_vec(out, s3);
        write_vec(out, normed);
        write_vec(out, std::vector<double>{double(N2), double(D2), double(A2), double(N3), double(D3), double(A3)});
        std::fclose(out);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "test_mixed_gpu_api_shim: %s\n", e.what());
        return 1;
    }
    return 0;
}
