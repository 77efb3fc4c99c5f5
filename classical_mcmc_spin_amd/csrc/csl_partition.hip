// csl_partition.hip — the partitioned persistent batched RK4 kernel, for
// lattices whose state does not fit one workgroup's LDS (e.g. the 65 536-site
// pyrochlore lattice: 1.5 MB of FP64 state per trajectory).
//
// The grid-wide stage kernels re-stream the state through HBM four times per
// step (lattice_gpu.cu:502-565 launches one kernel per stage).  Here each
// trajectory is spread over P co-resident workgroups ("parts"), each owning a
// contiguous range of cells (all sublattices): the RK4 state x and the
// accumulator of a part's own sites live in VGPRs for the whole launch, the
// stage values of its own sites and of a halo of remote neighbours live in
// LDS.  Only the values other parts read cross the chip, once per stage,
// through an exchange buffer in global memory (two parities) and one version
// counter per part: a part publishes version v+1 after writing the stage
// values that version carries, and waits for version v of every part it
// exchanges with before it reads that version's halo.  With the wait set
// symmetric (parts exchanged with in either direction), a part that writes
// version v+1 into the parity of version v-1 has seen version v of all its
// readers, so none of them still reads v-1.
//
// Per stage:  wait for version v -> copy the halo of v into LDS -> barrier ->
// field, LL right-hand side and RK4 update of the own sites (stage values
// kept in VGPRs) -> barrier -> own stage values into LDS and, where read by
// other parts, into the exchange buffer -> barrier -> publish v+1 (one
// release store by one thread).
// DBUF (when two LDS buffers of the own stage values fit): stage s reads
// buffer s&1 and writes the other one as each site is done, without the
// barrier between reads and writes and without holding the new values.
//
// Own sites of a part are laid out sublattice-major in LDS (slot = atom *
// cpad + local cell, cpad a multiple of 64), so every wave covers 64 sites of
// one sublattice and reads that sublattice's bond matrices, field and drive
// directions with scalar loads (translation-invariant tables only).  Records
// are per-part sums of the observables (reduced over the parts by a second
// kernel in a fixed order).
#include "csl_device.hpp"
#include "csl_lds.hpp"

namespace csl {

namespace {

template <int SPT>
constexpr int part_max_block() {
    return SPT <= 4 ? 1024 : 512;
}

// Iterations of one bounded wait for a version (each one an agent-scope
// load and a short sleep); a wait that runs out flags the launch.
constexpr int kPartSpin = 1 << 22;

// Undriven local field of a site whose partners' stage values are in LDS:
// H = -B + 2A S + sum_k J_k S_k (lattice.h:1655-1709), the bond terms as one
// FMA chain per component; slot() maps a partition slot (own q < n_loc,
// halo n_loc + h) to its LDS slot.
template <int ZMAX, class Slot>
__device__ __forceinline__ void part_field(const LatDev& L, const LdsState& Y, int atom, int cnt_a,
                                           const unsigned* nbp, Slot slot, const double S[3], double H[3]) {
    cptr<double> Bu = as_const(L.field) + atom * 3;
    H[0] = -Bu[0];
    H[1] = -Bu[1];
    H[2] = -Bu[2];
    if (L.has_onsite) {
        cptr<double> A2 = as_const(L.onsite) + atom * 9;
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int e = 0; e < 3; ++e) H[d] += A2[d * 3 + e] * S[e];
    }
#pragma unroll
    for (int k = 0; k < ZMAX; ++k) {
        if (k < cnt_a) {
            const int q = int((unsigned(fresh(int(nbp[k >> 1]))) >> (16 * (k & 1))) & 0xffffu);
            double Sp[3];
            Y.load(slot(q), Sp);
            add_bond_fused(as_const(L.mats) + (atom * L.zmax + k) * 9, Sp, H);
        }
    }
}

}  // namespace

template <int SPT, int ZMAX, bool DBUF>
__global__ void __launch_bounds__(part_max_block<SPT>()) k_partitioned(LatDev L0, PartArgs A) {
    extern __shared__ double smem[];
    const int T = blockDim.x;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = uniform(tid >> 6);
    const int nwaves = T >> 6;
    // LDS slots: own stage values [one or two buffers of n_loc] | halo
    const int hb = (DBUF ? 2 : 1) * A.n_loc;
    const int n_tot = hb + A.n_halo_max;
    const LdsState Y = lds_state(smem, n_tot);
    double* drv = smem + 3 * size_t(n_tot);  // [2 parity][3 stage times][2 pulses]
    double* red = drv + 12;                  // [nwaves][10]
    int* ctl = reinterpret_cast<int*>(red + nwaves * 10);  // [0]: a wait gave up

    // group and part of this workgroup; with xcd_map the P parts of a group
    // sit on one XCD (workgroups are dealt to the 8 XCDs round-robin)
    int grp, part;
    if (A.xcd_map) {
        const int per = int(gridDim.x) / 8;
        const int xcd = int(blockIdx.x) % 8, j = int(blockIdx.x) / 8;
        grp = xcd * (per / A.P) + j / A.P;
        part = j % A.P;
    } else {
        grp = int(blockIdx.x) / A.P;
        part = int(blockIdx.x) % A.P;
    }
    grp = uniform(grp);
    part = uniform(part);
    const int NP = L0.NP;
    const int nw = (ZMAX + 1) / 2;

    // own sites: slot q = tid + m*T; sublattice q / cpad (wave-uniform)
    unsigned nbp[SPT][(ZMAX + 1) / 2];
    int atm[SPT], cnt[SPT];
    unsigned act = 0, expm = 0;
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        const int q = tid + m * T;
        const int a = uniform((wave * kWave + m * T) / A.cpad);
        atm[m] = a;
        cnt[m] = uniform(as_const(L0.cnt_atom)[a]);
        const int c = q - a * A.cpad;
        if (c < A.c_per && part * A.c_per + c < L0.n_cells) act |= 1u << m;
        if (A.exported[size_t(part) * A.n_loc + q]) expm |= 1u << m;
#pragma unroll
        for (int j = 0; j < nw; ++j) nbp[m][j] = A.col[(size_t(part) * nw + j) * A.n_loc + q];
    }
    auto cell_of = [&](int m) { return part * A.c_per + (tid + m * T - atm[m] * A.cpad); };
    const int n_halo = A.n_halo[part];
    const int* hsrc = A.halo_src + size_t(part) * A.n_halo_max;
    const int* nbrs = A.nbr + size_t(part) * A.n_nbr_max;
    unsigned* ver = A.ver + size_t(grp) * A.P;
    double* exch = A.exch + size_t(grp) * 6 * size_t(NP);
    unsigned v = 0;  // versions this part has published (the same count in every part of the group)

    if (tid == 0) ctl[0] = 0;
    __syncthreads();

    // Wait until every part exchanged with has published version `want`.
    // The lanes of wave 0 poll one part each with relaxed agent-scope loads;
    // the barrier then orders every wave's halo loads after the observed
    // versions.  False when a wait gave up here or anywhere in the launch
    // (uniform over the workgroup).
    auto wait_for = [&](unsigned want) -> bool {
        if (wave == 0 && A.n_nbr_max > 0) {
            if (tid < A.n_nbr_max) {
                const int o = nbrs[tid];
                if (o >= 0) {
                    bool ok = false;
                    for (int it = 0; it < kPartSpin; ++it) {
                        if (__hip_atomic_load(ver + o, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= want) {
                            ok = true;
                            break;
                        }
                        if ((it & 255) == 255 &&
                            __hip_atomic_load(A.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                            break;
                        __builtin_amdgcn_s_sleep(1);
                    }
                    if (!ok) {
                        __hip_atomic_store(A.status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ctl[0] = 1;
                    }
                }
            }
        }
        __syncthreads();
        return ctl[0] == 0;
    };
    // halo of version `want` -> LDS slots n_loc .. (agent-scope loads: they
    // read the values at the device's coherence point, not a stale copy in
    // this XCD's L2 from two stages ago)
    auto import_halo = [&](unsigned want) {
        const double* src = exch + size_t(want & 1u) * 3 * size_t(NP);
        for (int h = tid; h < n_halo; h += T) {
            const int d = hsrc[h];
            const double val[3] = {__hip_atomic_load(src + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                                   __hip_atomic_load(src + NP + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                                   __hip_atomic_load(src + 2 * NP + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)};
            Y.store(hb + h, val);
        }
    };
    // the own stage values y of version v+1: LDS, and the exchange buffer
    // where another part reads them; then publish v+1
    // (two LDS buffers: the values were stored by put() during the stage).
    // Exports are agent-scope stores: written through to the device's
    // coherence point, so no release has to write the XCD's L2 back.
    auto put = [&](int m, int buf, const double y[3]) {
        Y.store(fresh(buf * A.n_loc + tid + m * T), y);
        if ((expm >> m) & 1u) {
            double* dst = exch + size_t((v + 1u) & 1u) * 3 * size_t(NP);
            const int d = fresh(atm[m] * L0.npad + cell_of(m));
            __hip_atomic_store(dst + d, y[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(dst + NP + d, y[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(dst + 2 * NP + d, y[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    };
    // Publication: every wave waits until its exports are complete (an
    // agent-scope store completes once it is performed at the coherence
    // point), the barrier orders all of them before thread 0's version
    // store, and a reader issues its halo loads only after it has seen the
    // version.  An agent-scope release fence would in addition write back
    // every dirty line of the XCD's L2 (about 6 us per stage on MI355X).
    auto publish = [&](const double (*yv)[3]) {
        if (yv) {
#pragma unroll
            for (int m = 0; m < SPT; ++m) put(m, 0, yv[m]);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        ++v;
        if (tid == 0) __hip_atomic_store(ver + part, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };

    double x[SPT][3], acc[SPT][3];
    const double dt = A.dt;
    const double b16 = (1.0 / 6.0) * dt, b13 = (1.0 / 3.0) * dt;

    for (int b = grp; b < A.B; b += A.groups) {
        LatDev Lb = L0;
        Lb.mats += size_t(b) * A.mats_stride;
        Lb.field += size_t(b) * A.field_stride;
        const double tp1 = A.t_pulse[2 * b], tp2 = A.t_pulse[2 * b + 1];
        const double am1 = A.amp[2 * b], am2 = A.amp[2 * b + 1];
        const bool drive_on = (am1 != 0.0) || (am2 != 0.0);
        const double alpha = Lb.alpha;
        const bool want_e = A.want_energy != 0;

        // drive envelopes of the three stage times of step n, both pulses,
        // into parity slot n&1 (read during step n; the barriers of step n-1
        // separate the write from the reads of step n-2)
        auto write_drive = [&](int n) {
            if (drive_on && wave == nwaves - 1 && lane < 6) {
                const int st = lane >> 1, pl = lane & 1;
                const double t = A.tgrid[n];
                const double ts = st == 0 ? t : (st == 1 ? t + 0.5 * dt : t + 1.0 * dt);
                drv[(n & 1) * 6 + lane] = drive_factor(pl ? am2 : am1, ts, pl ? tp2 : tp1, Lb.width, Lb.freq);
            }
        };
        // wave sums of the observables of the owned sites into red[]
        auto reduce_obs = [&](const LatDev& Lr, const double* e) {
            double pt[10];
#pragma unroll
            for (int j = 0; j < 10; ++j) pt[j] = 0.0;
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                if ((act >> m) & 1u) {
                    cptr<double> R = as_const(Lr.frames) + atm[m] * 9;
                    const double sg = as_const(Lr.afm)[atm[m]];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const double g = R[d * 3 + 0] * x[m][0] + R[d * 3 + 1] * x[m][1] + R[d * 3 + 2] * x[m][2];
                        pt[d] += sg * g;
                        pt[3 + d] += x[m][d];
                        pt[6 + d] += g;
                    }
                    pt[9] += e[m];
                }
            }
#pragma unroll
            for (int j = 0; j < 10; ++j) {
                if (j == 9 && !want_e) break;
                const double s = wave_sum_lane63(pt[j]);
                if (lane == kWave - 1) red[wave * 10 + j] = s;
            }
        };
        // the part's sums of the recorded step ri (after a barrier)
        auto record = [&](int ri) {
            if (wave == 0 && lane < 10 && (lane < 9 || want_e)) {
                double s = 0.0;
                for (int w = 0; w < nwaves; ++w) s += red[w * 10 + lane];
                A.part_obs[((size_t(b) * A.n_rec + ri) * A.P + part) * 10 + lane] = s;
            }
        };

        // initial state: VGPRs, and version v+1 of the stage values
        __syncthreads();  // the previous trajectory is done with LDS
        {
            const double* init = A.init + size_t(b) * A.init_stride;
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                x[m][0] = x[m][1] = x[m][2] = 0.0;
                acc[m][0] = acc[m][1] = acc[m][2] = 0.0;
                if ((act >> m) & 1u) {
                    const size_t i = size_t(cell_of(m)) * L0.n_atoms + atm[m];
                    x[m][0] = init[3 * i];
                    x[m][1] = init[3 * i + 1];
                    x[m][2] = init[3 * i + 2];
                }
            }
            write_drive(0);
            publish(x);
        }

        // One RK4 stage of step n; false when a wait gave up.
        auto stage = [&](auto s_, int par, bool rec, int ri) -> bool {
            constexpr int s = decltype(s_)::value;
            if (!wait_for(v)) return false;
            import_halo(v);
            __syncthreads();
            const LatDev L = refresh(Lb);
            constexpr int st = s == 0 ? 0 : (s == 3 ? 2 : 1);
            const double f1 = drive_on ? drv[par * 6 + st * 2 + 0] : 0.0;
            const double f2 = drive_on ? drv[par * 6 + st * 2 + 1] : 0.0;
            const bool drive_now = drive_on && (f1 != 0.0 || f2 != 0.0);
            const double cdt = (s == 2) ? 1.0 * dt : 0.5 * dt;
            const double bdt = (s == 0 || s == 3) ? b16 : b13;
            const bool obs = (s == 0) && rec;
            // two buffers: stage s reads buffer s&1 and writes the other one
            constexpr int cur = DBUF ? (s & 1) : 0, nxt = DBUF ? 1 - (s & 1) : 0;
            const int n_loc = A.n_loc;
            auto slot = [&](int q) {
                if constexpr (!DBUF) return q;
                else if constexpr (cur == 1) return q + n_loc;
                else return q < n_loc ? q : q + n_loc;
            };
            double e[SPT], yn[DBUF ? 1 : SPT][3];
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                double S[3], H[3];
                if (s == 0) {
                    S[0] = x[m][0];
                    S[1] = x[m][1];
                    S[2] = x[m][2];
                } else {
                    Y.load(fresh(cur * n_loc + tid + m * T), S);
                }
                part_field<ZMAX>(L, Y, atm[m], cnt[m], nbp[m], slot, S, H);
                e[m] = (obs && want_e) ? site_energy(L, atm[m], 0, S, H) : 0.0;
                if (drive_now) sub_drive(L, atm[m], f1, f2, H);
                double k[3];
                ll_su2_guarded(H, S, alpha, k);
                double* y = yn[DBUF ? 0 : m];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    if (s == 0) acc[m][d] = x[m][d] + bdt * k[d];
                    else if (s < 3) acc[m][d] = acc[m][d] + bdt * k[d];
                    else x[m][d] = acc[m][d] + bdt * k[d];
                    y[d] = (s < 3) ? x[m][d] + cdt * k[d] : x[m][d];
                }
                if (DBUF) put(m, nxt, y);
            }
            // stage 0: x still holds the state at t_n (the record's state)
            if (obs) reduce_obs(L, e);
            // one buffer: every read of this stage's LDS values must be done
            // before the new values replace them; records: red[] complete
            if (!DBUF || obs) __syncthreads();
            if (obs) record(ri);
            publish(DBUF ? nullptr : yn);
            return true;
        };

        bool alive = true;
        for (int n = 0; n < A.n_steps && alive; ++n) {
            const int par = n & 1;
            const bool rec = (n % A.save_interval) == 0;
            const int ri = n / A.save_interval;
            if (n + 1 < A.n_steps) write_drive(n + 1);
            alive = stage(std::integral_constant<int, 0>{}, par, rec, ri) &&
                    stage(std::integral_constant<int, 1>{}, par, rec, ri) &&
                    stage(std::integral_constant<int, 2>{}, par, rec, ri) &&
                    stage(std::integral_constant<int, 3>{}, par, rec, ri);
        }
        if (!alive || !wait_for(v)) return;
        // final record: the state after the last step (lattice_gpu.cu:567-568)
        import_halo(v);
        __syncthreads();
        {
            const LatDev L = refresh(Lb);
            double e[SPT];
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                double H[3];
                part_field<ZMAX>(L, Y, atm[m], cnt[m], nbp[m], [&](int q) { return q < A.n_loc ? q : q - A.n_loc + hb; },
                                 x[m], H);
                e[m] = want_e ? site_energy(L, atm[m], 0, x[m], H) : 0.0;
            }
            reduce_obs(L, e);
        }
        __syncthreads();
        record(A.n_rec - 1);
        if (A.final_states) {
            double* out = A.final_states + size_t(b) * L0.N * 3;
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                if ((act >> m) & 1u) {
                    const size_t i = size_t(cell_of(m)) * L0.n_atoms + atm[m];
                    out[3 * i] = x[m][0];
                    out[3 * i + 1] = x[m][1];
                    out[3 * i + 2] = x[m][2];
                }
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_part_reduce(const double* __restrict__ part_obs, int P, int B, int n_rec,
                                                     double inv_n, double* __restrict__ mags,
                                                     double* __restrict__ energy) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;  // (b, ri, j)
    const size_t total = size_t(B) * n_rec * 10;
    if (i >= total) return;
    const int j = int(i % 10);
    const size_t br = i / 10;
    const int ri = int(br % n_rec), b = int(br / n_rec);
    const double* src = part_obs + br * P * 10 + j;
    double s = 0.0;
    for (int p = 0; p < P; ++p) s += src[size_t(p) * 10];
    if (j < 9) {
        if (mags) mags[(size_t(ri) * B + b) * 9 + j] = s * inv_n;
    } else if (energy) {
        energy[size_t(ri) * B + b] = s * inv_n;
    }
}

namespace {

template <int SPT, int ZMAX, bool DBUF>
hipError_t launch_part_t(const LatDev& L, const PartArgs& A, int block, size_t lds, hipStream_t st, int* occ) {
    auto kern = k_partitioned<SPT, ZMAX, DBUF>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    if (occ) return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, kern, block, lds);
    hipLaunchKernelGGL(kern, dim3(unsigned(A.groups * A.P)), dim3(block), lds, st, L, A);
    return hipGetLastError();
}

template <int SPT, bool DBUF>
hipError_t launch_part_zd(const LatDev& L, const PartArgs& A, int block, size_t lds, hipStream_t st, int* occ) {
    if (L.zmax <= 3) return launch_part_t<SPT, 3, DBUF>(L, A, block, lds, st, occ);
    if (L.zmax <= 6) return launch_part_t<SPT, 6, DBUF>(L, A, block, lds, st, occ);
    if (L.zmax <= 12) return launch_part_t<SPT, 12, DBUF>(L, A, block, lds, st, occ);
    return hipErrorInvalidValue;
}

template <int SPT>
hipError_t launch_part_z(const LatDev& L, const PartArgs& A, int block, size_t lds, hipStream_t st, int* occ) {
    return A.dbuf ? launch_part_zd<SPT, true>(L, A, block, lds, st, occ)
                  : launch_part_zd<SPT, false>(L, A, block, lds, st, occ);
}

}  // namespace

bool partitioned_spt_supported(int spt) {
    return spt == 1 || spt == 2 || spt == 3 || spt == 4 || spt == 6 || spt == 8;
}

int partitioned_block_limit(int spt) {
    return partitioned_spt_supported(spt) ? (spt <= 4 ? part_max_block<1>() : part_max_block<8>()) : 0;
}

size_t partitioned_lds_bytes(int n_loc, int n_halo_max, int block, bool dbuf) {
    return sizeof(double) * (3 * size_t((dbuf ? 2 : 1) * n_loc + n_halo_max) + 12 + size_t(block / kWave) * 10) + 16;
}

hipError_t launch_partitioned(const LatDev& L, const PartArgs& A, int block, int spt, hipStream_t st, int* occ) {
    const size_t lds = partitioned_lds_bytes(A.n_loc, A.n_halo_max, block, A.dbuf != 0);
    switch (spt) {
        case 1: return launch_part_z<1>(L, A, block, lds, st, occ);
        case 2: return launch_part_z<2>(L, A, block, lds, st, occ);
        case 3: return launch_part_z<3>(L, A, block, lds, st, occ);
        case 4: return launch_part_z<4>(L, A, block, lds, st, occ);
        case 6: return launch_part_z<6>(L, A, block, lds, st, occ);
        case 8: return launch_part_z<8>(L, A, block, lds, st, occ);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_part_reduce(const double* part_obs, int P, int B, int n_rec, double inv_n, double* mags,
                              double* energy, hipStream_t st) {
    const size_t total = size_t(B) * n_rec * 10;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_part_reduce, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, part_obs, P, B, n_rec,
                       inv_n, mags, energy);
    return hipGetLastError();
}

}  // namespace csl
