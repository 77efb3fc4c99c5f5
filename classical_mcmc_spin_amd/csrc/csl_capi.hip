// csl_capi.hip — host implementation of the C ABI declared in include/csl_gpu.h.
//
// Owns the device copies of one lattice (handle), converts the reference's
// table layouts into the device layout described in csl_internal.hpp,
// computes the integrator time grids on the host and drives the kernels of
// csl_kernels.hip on one stream per handle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include <hipfft/hipfft.h>

#include "../../include/csl_gpu.h"
#include "csl_host.hpp"
#include "csl_kernels.hpp"
#include "csl_tableau.hpp"

// Scratch of csl_integrate_batched_noalloc: one region per GPU, in the code
// object's device memory, so that the entry point allocates nothing.  Host
// code uses the address hipGetSymbolAddress returns for the current device
// (scratch_ptr), never the array's name.
static __device__ __align__(256) unsigned char scratch[CSL_NOALLOC_SCRATCH_BYTES];

namespace {

using namespace cslh;

// Address of `scratch` on the current device, looked up once per device.
unsigned char* scratch_ptr() {
    static std::mutex mu;
    static std::map<int, unsigned char*> addr;
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) fail(CSL_ERR_HIP, "no device address for the scratch region");
    addr[dev] = static_cast<unsigned char*>(p);
    return addr[dev];
}


struct TuneVal {
    int block, spt;
    int flags;  // csl::kPersistDoubleBuffer | csl::kPersistUnrollStages
    bool dbuf() const { return flags & csl::kPersistDoubleBuffer; }
};

// batch-size class of the launch-shape cache: 1 trajectory, a few, >= 64
int batch_class(size_t B) { return B >= 64 ? 2 : (B > 1 ? 1 : 0); }

}  // namespace

// Partition of a lattice over the parts of the partitioned kernel
// (csl_partition.hip; the field meanings are those of csl::PartArgs).
struct PartPlan {
    int P = 0;  // 0: no plan built
    int c_per = 0, cpad = 0, n_loc = 0, n_halo_max = 0, n_nbr_max = 0, block = 0, spt = 0, dbuf = 0;
    DBuf<unsigned> col, exported;
    DBuf<int> halo_src, n_halo, nbr;
};

struct csl_lattice {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t N = 0, D = 3, n_atoms = 1;
    csl::LatDev L{};
    size_t lds_limit = 65536;
    int n_cu = 1;
    // tables (device)
    DBuf<int> col, cnt, cnt_atom;
    DBuf<double> mats, field, onsite, frames, afm, fdrive;
    // host copies of per-atom data
    std::vector<double> h_frames, h_afm, h_fdrive;
    double amp = 0.0, width = 1.0, freq = 0.0, t1 = 0.0, t2 = 0.0;
    // state
    DBuf<double> state;  // AoS [N*3]
    bool has_state = false;
    // workspaces
    DBuf<double> X, ACC, Y1, Y2, partials, fac, tgrid, rec_mags, rec_energy, rec_states, aux;
    DBuf<double> KS;              // stage derivatives of the generic fixed-step engine [S][3][NP]
    DBuf<double> bt_init, bt_tp, bt_amp, bt_final;
    DBuf<double> dp_times;        // Dormand-Prince records: times [B][max_rec]
    // persistent-kernel work split (batches larger than the co-resident workgroups)
    DBuf<int> ps_sched;
    DBuf<double> ps_saved;
    DBuf<unsigned> ps_flags, ps_status;
    DBuf<csl::LatDev> ps_lats;          // lattices of a multi-lattice batch
    DBuf<int> ps_latof;
    DBuf<size_t> ps_initoff;
    int split_batches = 1;
    // partitioned kernel: 0 = for lattices that do not fit one workgroup,
    // P > 0 = always with P parts, -1 = never (csl_set_partition)
    int part_req = 0;
    PartPlan part;
    DBuf<double> part_exch, part_obs;
    DBuf<unsigned> part_ver;
    int last_parts = 0, last_groups = 0;
    // adjoint work buffers (stage inputs, field cotangents, checkpoints, ...)
    DBuf<double> adj_Y, adj_U, adj_ckpt, adj_lam, adj_lossn, adj_part;
    hipStream_t stream2 = nullptr;  // second stream of the adjoint (parameter terms)
    int stage_graph = 0;          // batched stage-kernel path replayed from a hipGraph
    int last_grid = 0;
    DBuf<long long> dp_counters;  // [B][4]
    DBuf<double> dp_segs;         // control-2 segments [B][max_seg][2]
    DBuf<int> dp_nseg;            // [B]
    // CSR -> device slot mapping of the handle's tables (for per-replica tables)
    std::vector<size_t> h_off, h_slot_of, h_par;
    // per-trajectory Hamiltonians (csl_set_replica_hamiltonians)
    DBuf<double> rep_mats, rep_field;
    size_t rep_B = 0, rep_mats_stride = 0, rep_field_stride = 0;
    // launch shapes chosen by the autotuner / csl_set_launch_config for this
    // handle's lattice, per batch-size class (batch_class)
    std::map<int, TuneVal> tune;
    double last_ms = 0.0;
    int last_block = 0, last_spt = 0, last_dbuf = 0, last_variant = 0;
    size_t last_launches = 0;
    // stage-kernel launch sequence captured as a hipGraph (engine 3)
    hipGraphExec_t graph_exec = nullptr;
    std::vector<unsigned char> graph_key;
    ~csl_lattice() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (stream2) (void)hipStreamDestroy(stream2);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

csl_lattice* check_handle(csl_lattice* h) {
    if (!h) fail(CSL_ERR_NULL, "null lattice handle");
    hip_check(hipSetDevice(h->device), "hipSetDevice");
    return h;
}

// Build device tables from the reference CSR mirror (lattice.h:868-903).
void build_tables(csl_lattice* h, size_t N, size_t D, size_t n_atoms, const double* field,
                  const double* onsite, const std::vector<size_t>& off,
                  const std::vector<size_t>& partner, const double* J) {
    if (D != 3) fail(CSL_ERR_UNSUPPORTED, "spin_dim != 3 is handled by the mixed-lattice API");
    if (n_atoms == 0 || N == 0 || N % n_atoms) fail(CSL_ERR_ARG, "lattice_size must be a multiple of N_atoms");
    if (N > size_t(std::numeric_limits<int>::max() / 4)) fail(CSL_ERR_ARG, "lattice too large");
    h->N = N;
    h->D = D;
    h->n_atoms = n_atoms;
    const int n_cells = int(N / n_atoms);
    const int npad = (n_cells + csl::kWave - 1) / csl::kWave * csl::kWave;
    const int NP = int(n_atoms) * npad;
    auto dev = [&](size_t i) { return int(i % n_atoms) * npad + int(i / n_atoms); };
    int zmax = 0;
    for (size_t i = 0; i < N; ++i) {
        if (off[i + 1] < off[i]) fail(CSL_ERR_ARG, "bond offsets must be non-decreasing");
        zmax = std::max<int>(zmax, int(off[i + 1] - off[i]));
    }
    if (zmax > 12) fail(CSL_ERR_UNSUPPORTED, "more than 12 bonds per site");
    for (size_t k = 0; k < off[N]; ++k)
        if (partner[k] >= N) fail(CSL_ERR_ARG, "bond partner index out of range");
    for (size_t q = 0; q < N * D * (D + 1) + off[N] * D * D; ++q) {
        const double v = q < N * D ? field[q] : q < N * D * (D + 1) ? onsite[q - N * D] : J[q - N * D * (D + 1)];
        if (!std::isfinite(v)) fail(CSL_ERR_NONFINITE, "non-finite Hamiltonian coefficient");
    }
    // Translation invariance of the bond matrices per (sublattice, slot).
    // The reference's per-site bond order (lattice.h:785-803) depends on
    // which neighbour visits first, which changes on cells where a bond
    // wraps; the majority slot sequence of each sublattice is the canonical
    // one and sites whose bonds are a permutation of it are re-ordered to it
    // (slot_of[k] = position of CSR bond k in the device slot order).  Only
    // the order of that site's FP64 bond sum changes.
    std::vector<size_t> slot_of(off[N]);
    for (size_t k = 0; k < off[N]; ++k) slot_of[k] = k;  // identity (CSR order) by default
    bool uni = true;
    std::vector<size_t> canon(n_atoms, 0);  // representative site of each sublattice
    {
        auto same_seq = [&](size_t a, size_t b) {
            const size_t ca = off[a + 1] - off[a];
            return ca == off[b + 1] - off[b] &&
                   std::memcmp(J + off[a] * 9, J + off[b] * 9, ca * 9 * sizeof(double)) == 0;
        };
        for (size_t at = 0; at < n_atoms && uni; ++at) {
            // majority sequence: Boyer-Moore vote over the sites of this sublattice
            size_t cand = at, votes = 0;
            for (size_t i = at; i < N; i += n_atoms) {
                if (votes == 0) { cand = i; votes = 1; }
                else if (same_seq(i, cand)) ++votes;
                else --votes;
            }
            canon[at] = cand;
            const size_t c = off[cand + 1] - off[cand];
            for (size_t i = at; i < N && uni; i += n_atoms) {
                if (same_seq(i, cand)) {
                    for (size_t q = 0; q < c; ++q) slot_of[off[i] + q] = q;
                    continue;
                }
                if (off[i + 1] - off[i] != c) { uni = false; break; }
                std::vector<char> used(c, 0);
                for (size_t q = 0; q < c && uni; ++q) {  // canonical slot q takes an equal unused bond
                    bool found = false;
                    for (size_t r = 0; r < c; ++r)
                        if (!used[r] && std::memcmp(J + (off[i] + r) * 9, J + (off[cand] + q) * 9,
                                                    9 * sizeof(double)) == 0) {
                            used[r] = 1;
                            slot_of[off[i] + r] = q;
                            found = true;
                            break;
                        }
                    if (!found) uni = false;
                }
            }
        }
        if (!uni)
            for (size_t i = 0; i < N; ++i)
                for (size_t k = off[i]; k < off[i + 1]; ++k) slot_of[k] = k - off[i];
    }
    bool site_uni = true;
    for (size_t i = 0; i < N && site_uni; ++i) {
        const size_t r = i % n_atoms;
        if (std::memcmp(field + i * 3, field + r * 3, 3 * sizeof(double)) ||
            std::memcmp(onsite + i * 9, onsite + r * 9, 9 * sizeof(double)))
            site_uni = false;
    }
    std::vector<int> col(size_t(zmax) * NP, 0), cnt(NP, 0), cnt_atom(n_atoms, 0);
    std::vector<double> mats;
    if (uni) mats.assign(n_atoms * size_t(zmax) * 9, 0.0);
    else mats.assign(size_t(zmax) * NP * 9, 0.0);
    for (size_t i = 0; i < N; ++i) {
        const int p = dev(i);
        const int c = int(off[i + 1] - off[i]);
        cnt[p] = c;
        for (int k = 0; k < zmax; ++k) col[size_t(k) * NP + p] = p;  // unused slots point at the site itself
        for (int k = 0; k < c; ++k) {
            const size_t q = slot_of[off[i] + k];
            col[q * NP + p] = dev(partner[off[i] + k]);
            const double* src = J + (off[i] + k) * 9;
            double* dst = uni ? &mats[((i % n_atoms) * zmax + q) * 9] : &mats[(q * NP + p) * 9];
            std::memcpy(dst, src, 9 * sizeof(double));
        }
    }
    for (size_t a = 0; a < n_atoms; ++a) cnt_atom[a] = int(off[canon[a] + 1] - off[canon[a]]);
    h->h_off = off;
    h->h_par.assign(partner.begin(), partner.begin() + off[N]);
    h->h_slot_of = slot_of;
    h->rep_B = 0;
    std::vector<double> f, o;
    if (site_uni) {
        f.assign(field, field + n_atoms * 3);
        o.assign(onsite, onsite + n_atoms * 9);
    } else {
        f.assign(size_t(NP) * 3, 0.0);
        o.assign(size_t(NP) * 9, 0.0);
        for (size_t i = 0; i < N; ++i) {
            std::memcpy(&f[size_t(dev(i)) * 3], field + i * 3, 3 * sizeof(double));
            std::memcpy(&o[size_t(dev(i)) * 9], onsite + i * 9, 9 * sizeof(double));
        }
    }
    // the kernels use 2A (lattice.h:1660-1665 multiplies by 2.0; exact)
    bool has_onsite = false;
    for (double& v : o) {
        has_onsite |= (v != 0.0);
        v *= 2.0;
    }
    hipStream_t s = h->stream;
    h->col.upload(col.data(), col.size(), s);
    h->cnt.upload(cnt.data(), cnt.size(), s);
    h->cnt_atom.upload(cnt_atom.data(), cnt_atom.size(), s);
    h->mats.upload(mats.data(), mats.size(), s);
    h->field.upload(f.data(), f.size(), s);
    h->onsite.upload(o.data(), o.size(), s);
    h->h_frames.assign(n_atoms * 9, 0.0);
    for (size_t a = 0; a < n_atoms; ++a)
        for (int d = 0; d < 3; ++d) h->h_frames[a * 9 + d * 4] = 1.0;
    h->h_afm.assign(n_atoms, 1.0);
    h->h_fdrive.assign(2 * n_atoms * 3, 0.0);
    h->frames.upload(h->h_frames.data(), h->h_frames.size(), s);
    h->afm.upload(h->h_afm.data(), h->h_afm.size(), s);
    h->fdrive.upload(h->h_fdrive.data(), h->h_fdrive.size(), s);
    hip_check(hipStreamSynchronize(s), "table upload");

    csl::LatDev& L = h->L;
    L.N = int(N);
    L.n_atoms = int(n_atoms);
    L.n_cells = n_cells;
    L.npad = npad;
    L.NP = NP;
    L.zmax = std::max(zmax, 1);
    L.uniform = uni ? 1 : 0;
    L.site_uniform = site_uni ? 1 : 0;
    L.has_onsite = has_onsite ? 1 : 0;
    bool full = uni && zmax > 0;
    for (size_t a = 0; a < n_atoms; ++a) full = full && cnt_atom[a] == zmax;
    L.full_coord = full ? 1 : 0;
    L.col = h->col.p;
    L.cnt = h->cnt.p;
    L.cnt_atom = h->cnt_atom.p;
    L.mats = h->mats.p;
    L.field = h->field.p;
    L.onsite = h->onsite.p;
    L.frames = h->frames.p;
    L.afm = h->afm.p;
    L.fdrive = h->fdrive.p;
    L.width = 1.0;
    L.freq = 0.0;
    L.alpha = 0.0;
}

csl_lattice* new_handle() {
    auto* h = new csl_lattice();
    hip_check(hipGetDevice(&h->device), "hipGetDevice");
    hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
    int lds = 0, ncu = 0;
    hip_check(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device),
              "lds attribute");
    hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device),
              "cu attribute");
    h->lds_limit = size_t(lds);
    h->n_cu = ncu;
    return h;
}

// Wave-sized units of sites (64 site slots of which at least one is a real
// site) on the busiest of a CU's four SIMDs for a persistent launch shape
// (the waves of a workgroup are dealt to the SIMDs in turn).
int simd_share(const csl::LatDev& L, int block, int spt, int flags) {
    const int nw = block / 64;
    if (csl::persistent_variant(L, block, spt, flags) & csl::kVariantAtomMajor)
        return csl::atom_major_max_load(nw, L.n_atoms, L.npad / 64);
    int load[4] = {0, 0, 0, 0};
    for (int w = 0; w < nw; ++w)
        for (int m = 0; m < spt; ++m) {
            bool any = false;
            for (int p = w * 64 + m * block; p < std::min(w * 64 + m * block + 64, L.NP) && !any; ++p)
                any = (p % L.npad) < L.n_cells;
            load[w & 3] += any ? 1 : 0;
        }
    return std::max(std::max(load[0], load[1]), std::max(load[2], load[3]));
}

// Candidate persistent launch shapes for this lattice (block multiple of 64).
std::vector<TuneVal> persistent_candidates(const csl_lattice* h) {
    std::vector<TuneVal> c;
    const int NP = h->L.NP;
    for (int spt : {1, 2, 3, 4, 6}) {
        int block = (NP + spt - 1) / spt;
        block = (block + 63) / 64 * 64;
        if (block > csl::persistent_block_limit(spt) || block < 64) continue;
        if (block * spt - NP >= 64 * spt) continue;  // a whole idle wave per site slot
        for (int fl = 0; fl < 4; ++fl)
            if (csl::persistent_lds_bytes(h->L, block, fl & csl::kPersistDoubleBuffer) <= h->lds_limit)
                c.push_back({block, spt, fl});
    }
    // atom-major shapes with more waves than the exact tiling when that
    // lowers the busiest SIMD's share (the units of a sublattice dealt
    // unevenly over its waves): for each sites-per-thread value the block
    // with the lowest share, after the shapes above
    const int fl = csl::kPersistDoubleBuffer | csl::kPersistUnrollStages;
    const int n_atoms = h->L.n_atoms;
    if (h->L.npad % 64 == 0 && n_atoms > 0) {
        const int U = h->L.npad / 64;
        for (int spt : {1, 2, 3, 4, 6}) {
            int best = 0, best_load = 1 << 30;
            for (int wpa = (U + spt - 1) / spt; wpa <= U; ++wpa) {
                const int block = n_atoms * wpa * 64;
                if (block > csl::persistent_block_limit(spt)) break;
                if (!(csl::persistent_variant(h->L, block, spt, fl) & csl::kVariantAtomMajor)) continue;
                if (csl::persistent_lds_bytes(h->L, block, true) > h->lds_limit) continue;
                const int load = csl::atom_major_max_load(block / 64, n_atoms, U);
                if (load < best_load) {
                    best_load = load;
                    best = block;
                }
            }
            if (best == 0) continue;
            bool seen = false;
            for (const auto& v : c) seen = seen || (v.block == best && v.spt == spt);
            if (!seen) c.push_back({best, spt, fl});
        }
    }
    return c;
}

TuneVal pick_persistent(const csl_lattice* h, size_t B, int block_req) {
    auto cands = persistent_candidates(h);
    if (cands.empty()) fail(CSL_ERR_UNSUPPORTED, "lattice does not fit the persistent kernel");
    if (block_req > 0) {
        for (auto& c : cands)
            if (c.block == block_req) return c;
        fail(CSL_ERR_ARG, "requested block size is not a valid persistent launch shape");
    }
    auto it = h->tune.find(batch_class(B));
    if (it != h->tune.end()) return it->second;
    // heuristic: two LDS buffers (one barrier per stage) and compile-time
    // stages, the shape whose busiest SIMD carries the fewest wave-sized
    // units of sites (a stage ends when that SIMD is done), ties: the fewest
    // sites per thread
    const int fl = csl::kPersistDoubleBuffer | csl::kPersistUnrollStages;
    const TuneVal* best = nullptr;
    int best_share = 0;
    for (const auto& c : cands) {
        if (c.flags != fl) continue;
        const int sh = simd_share(h->L, c.block, c.spt, c.flags);
        if (!best || sh < best_share || (sh == best_share && c.spt < best->spt)) {
            best = &c;
            best_share = sh;
        }
    }
    if (best) return *best;
    for (auto& c : cands)
        if (c.flags == csl::kPersistDoubleBuffer) return c;
    return cands.front();
}

struct BatchOut {
    std::vector<double> times;
    size_t n_steps = 0, n_rec = 0;
};

// Device memory of one persistent launch supplied by the caller instead of
// the handle's buffers (csl_integrate_batched_noalloc): outputs, and scratch
// carved out of a preallocated region for the time grid, the work-split
// schedule, the seam states and the flags.
struct ExtBufs {
    double* mags;              // [n_rec][B][9] or null
    double* energy;            // [n_rec][B] or null
    double* final_states;      // [B][N*3] or null
    unsigned char* scratch;    // preallocated device region
    size_t scratch_bytes;
};

// Byte layout of the scratch region of one launch: time grid, status word,
// flags, schedule, seam states; every part 256-byte aligned.
struct ScratchLayout {
    size_t tgrid, status, flags, sched, saved, total;
};

ScratchLayout scratch_layout(size_t n_grid, size_t G, size_t sched_ints, size_t NP) {
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    ScratchLayout l{};
    l.tgrid = 0;
    l.status = up(l.tgrid + n_grid * sizeof(double));
    l.flags = up(l.status + sizeof(unsigned));
    l.sched = up(l.flags + G * sizeof(unsigned));
    l.saved = up(l.sched + sched_ints * sizeof(int));
    l.total = up(l.saved + G * 3 * NP * sizeof(double));
    return l;
}

// Runs the persistent kernel; device outputs are left in h->rec_mags /
// h->rec_energy / h->bt_final, or in ext's buffers (then no device memory is
// allocated and the launch is not timed).
void run_persistent(csl_lattice* h, size_t B, const double* d_init, size_t init_stride,
                    const double* d_tp, const double* d_amp, const std::vector<double>& grid,
                    double dt, size_t si, bool want_energy, bool want_final, const TuneVal& tv,
                    double* ms_out, const ExtBufs* ext = nullptr) {
    const size_t n_steps = grid.size() - 1;
    const size_t n_rec = records_for(n_steps, si);
    if (n_steps > size_t(std::numeric_limits<int>::max())) fail(CSL_ERR_ARG, "too many steps");
    hipStream_t s = h->stream;
    csl::PersistArgs A{};
    if (ext) {
        A.mags = ext->mags;
        A.energy = want_energy ? ext->energy : nullptr;
        A.final_states = want_final ? ext->final_states : nullptr;
    } else {
        h->tgrid.upload(grid.data(), grid.size(), s);
        h->rec_mags.ensure(n_rec * B * 9);
        if (want_energy) h->rec_energy.ensure(n_rec * B);
        if (want_final) h->bt_final.ensure(B * h->N * 3);
        A.tgrid = h->tgrid.p;
        A.mags = h->rec_mags.p;
        A.energy = want_energy ? h->rec_energy.p : nullptr;
        A.final_states = want_final ? h->bt_final.p : nullptr;
    }
    A.B = int(B);
    A.init = d_init;
    A.init_stride = init_stride;
    A.t_pulse = d_tp;
    A.amp = d_amp;
    A.n_steps = int(n_steps);
    A.save_interval = int(si);
    A.n_rec = int(n_rec);
    A.dt = dt;
    csl::LatDev L = h->L;
    if (h->rep_B) {
        if (B != h->rep_B) fail(CSL_ERR_ARG, "batch size differs from the per-replica Hamiltonians set");
        L.mats = h->rep_mats.p;
        L.field = h->rep_field.p;
        A.mats_stride = h->rep_mats_stride;
        A.field_stride = h->rep_field_stride;
    }
    // More trajectories than workgroups that fit the GPU at once: B
    // workgroups would run in rounds, the last one mostly idle (513 = 2*256 + 1
    // takes three rounds).  Instead G co-resident workgroups share the B * n
    // trajectory-steps evenly (McNaughton's wrap-around): workgroup g takes
    // the contiguous range [g*W, (g+1)*W) of the trajectory-major step order,
    // W = ceil(B*n/G) >= n, so at most one trajectory is cut between g and
    // g+1.  g runs the first part of that trajectory before anything else and
    // saves the state; g+1 runs the second part last, after its own whole
    // trajectories, by which time the state is long there.
    const int occ = h->split_batches ? csl::persistent_occupancy(L, tv.block, tv.spt, tv.flags) : 0;
    const size_t G = size_t(std::max(occ, 0)) * size_t(h->n_cu);
    A.sched = nullptr;
    h->last_grid = int(B);
    const bool split = G > 0 && B > G && n_steps > 0 && G <= size_t(std::numeric_limits<int>::max());
    size_t max_p = 1;
    const std::vector<int> sched = split ? split_schedule(B, G, n_steps, max_p) : std::vector<int>();
    unsigned* status = nullptr;
    if (ext) {
        // everything in the caller's preallocated scratch; cleared by a fill
        // kernel (no memset nodes for callers that capture the launch)
        const ScratchLayout l = scratch_layout(grid.size(), split ? G : 0, sched.size(), size_t(h->L.NP));
        if (l.total > ext->scratch_bytes)
            fail(CSL_ERR_CAPACITY, "the launch needs " + std::to_string(l.total) + " bytes of scratch, more than the " +
                                       std::to_string(ext->scratch_bytes) + " available");
        unsigned char* base = ext->scratch;
        hip_check(hipMemcpyAsync(base + l.tgrid, grid.data(), grid.size() * sizeof(double), hipMemcpyHostToDevice, s),
                  "H2D");
        A.tgrid = reinterpret_cast<const double*>(base + l.tgrid);
        status = reinterpret_cast<unsigned*>(base + l.status);
        if (split) {
            hip_check(hipMemcpyAsync(base + l.sched, sched.data(), sched.size() * sizeof(int), hipMemcpyHostToDevice, s),
                      "H2D");
            A.sched = reinterpret_cast<const int*>(base + l.sched);
            A.saved = reinterpret_cast<double*>(base + l.saved);
            A.flags = reinterpret_cast<unsigned*>(base + l.flags);
        }
        // status and flags are contiguous words: one fill clears both
        hip_check(csl::launch_fill_u32(status, 0u, (l.sched - l.status) / sizeof(unsigned), s), "fill");
    } else if (split) {
        h->ps_sched.upload(sched.data(), sched.size(), s);
        h->ps_saved.ensure(G * 3 * size_t(h->L.NP));
        h->ps_flags.ensure(G);
        h->ps_status.ensure(1);
        hip_check(hipMemsetAsync(h->ps_flags.p, 0, G * sizeof(unsigned), s), "memset");
        hip_check(hipMemsetAsync(h->ps_status.p, 0, sizeof(unsigned), s), "memset");
        A.sched = h->ps_sched.p;
        A.saved = h->ps_saved.p;
        A.flags = h->ps_flags.p;
        status = h->ps_status.p;
    }
    if (split) {
        A.max_pieces = int(max_p);
        A.grid = int(G);
        A.status = status;
        h->last_grid = int(G);
    }
    if (ext) {
        hip_check(csl::launch_persistent(L, A, tv.block, tv.spt, tv.flags, s), "k_persistent launch");
        hip_check(hipStreamSynchronize(s), "k_persistent");
        if (A.sched) {
            unsigned st = 0;
            hip_check(hipMemcpy(&st, status, sizeof(unsigned), hipMemcpyDeviceToHost), "D2H");
            if (st) fail(CSL_ERR_HIP, "a workgroup of the split batch waited too long for a trajectory's state");
        }
        h->last_ms = 0.0;
        h->last_block = tv.block;
        h->last_spt = tv.spt;
        h->last_dbuf = tv.flags;
        h->last_variant = csl::persistent_variant(L, tv.block, tv.spt, tv.flags);
        h->last_launches = 1;
        return;
    }
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, s), "event record");
    hip_check(csl::launch_persistent(L, A, tv.block, tv.spt, tv.flags, s), "k_persistent launch");
    hip_check(hipEventRecord(e1, s), "event record");
    hip_check(hipEventSynchronize(e1), "k_persistent");
    if (A.sched) {
        unsigned st = 0;
        hip_check(hipMemcpy(&st, h->ps_status.p, sizeof(unsigned), hipMemcpyDeviceToHost), "D2H");
        if (st) fail(CSL_ERR_HIP, "a workgroup of the split batch waited too long for a trajectory's state");
    }
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->last_ms = ms;
    h->last_block = tv.block;
    h->last_spt = tv.spt;
    h->last_dbuf = tv.flags;
    h->last_variant = csl::persistent_variant(L, tv.block, tv.spt, tv.flags);
    h->last_launches = 1;
    if (ms_out) *ms_out = ms;
}

// Host-side partition of the lattice into P parts (csl_partition.hip):
// part w owns the cells [w*c_per, (w+1)*c_per) of every sublattice, LDS
// slot atom*cpad + (cell - w*c_per); partners owned by another part get a
// halo slot n_loc + h.  False when P parts do not fit (an empty part, LDS,
// 16-bit slots, more than 64 parts exchanged with, no thread shape).  Two
// LDS buffers of the own stage values when they fit.
struct PartHost {
    int P = 0, c_per = 0, cpad = 0, n_loc = 0, n_halo_max = 0, n_nbr_max = 0, block = 0, spt = 0, dbuf = 0;
    std::vector<unsigned> col, exported;
    std::vector<int> halo_src, n_halo, nbr;
};

bool plan_partition(const csl_lattice* h, int P, PartHost& o) {
    const csl::LatDev& L = h->L;
    const int na = int(h->n_atoms), nc = L.n_cells, npad = L.npad, NP = L.NP, Z = L.zmax;
    if (P < 1 || P > nc || Z < 1) return false;
    const int c_per = (nc + P - 1) / P;
    if ((P - 1) * c_per >= nc) return false;  // an empty part
    const int cpad = (c_per + csl::kWave - 1) / csl::kWave * csl::kWave;
    const long n_loc = long(na) * cpad;
    if (n_loc > 65535) return false;
    // thread shape: block = 64*w threads, n_loc = block * spt
    const int waves = int(n_loc / csl::kWave);
    int block = 0, spt = 0;
    for (int w = 16; w >= 1 && !block; --w) {
        if (waves % w) continue;
        const int sp = waves / w;
        if (csl::partitioned_spt_supported(sp) && w * csl::kWave <= csl::partitioned_block_limit(sp)) {
            block = w * csl::kWave;
            spt = sp;
        }
    }
    if (!block) return false;
    // device-order partner table from the CSR mirror (the slot order of build_tables)
    const size_t N = h->N;
    auto dev = [&](size_t i) { return int(i % h->n_atoms) * npad + int(i / h->n_atoms); };
    std::vector<int> col(size_t(Z) * NP, -1), cnt(NP, 0);
    for (size_t i = 0; i < N; ++i) {
        const int p = dev(i);
        cnt[p] = int(h->h_off[i + 1] - h->h_off[i]);
        for (size_t k = h->h_off[i]; k < h->h_off[i + 1]; ++k)
            col[h->h_slot_of[k] * size_t(NP) + p] = dev(h->h_par[k]);
    }
    // words per site of the packed partner table: the kernel is compiled
    // for the bucketed bond count (3, 6 or 12)
    const int Zb = csl::zmax_bucket(Z);
    const int nw = (Zb + 1) / 2;
    o = PartHost{};
    o.col.assign(size_t(P) * nw * n_loc, 0u);
    o.exported.assign(size_t(P) * n_loc, 0u);
    std::vector<std::vector<int>> halo(P);
    std::vector<std::vector<char>> link(P, std::vector<char>(P, 0));
    std::vector<int> stamp(NP, -1), hidx(NP, 0);
    for (int w = 0; w < P; ++w) {
        for (int a = 0; a < na; ++a)
            for (int c = 0; c < cpad; ++c) {
                const int q = a * cpad + c, cell = w * c_per + c;
                const bool real = c < c_per && cell < nc;
                const int p = a * npad + cell;
                for (int k = 0; k < Zb; ++k) {
                    int slot = q;  // padding slots and unused bond slots read the site itself
                    if (real && k < cnt[p]) {
                        const int pd = col[size_t(k) * NP + p];
                        const int ap = pd / npad, cp = pd - ap * npad, wp = cp / c_per;
                        if (wp == w) {
                            slot = ap * cpad + (cp - w * c_per);
                        } else {
                            if (stamp[pd] != w) {
                                stamp[pd] = w;
                                hidx[pd] = int(halo[w].size());
                                halo[w].push_back(pd);
                            }
                            slot = int(n_loc) + hidx[pd];
                            link[w][wp] = link[wp][w] = 1;
                            o.exported[size_t(wp) * n_loc + ap * cpad + (cp - wp * c_per)] = 1u;
                        }
                    }
                    if (slot > 65535) return false;
                    o.col[(size_t(w) * nw + k / 2) * n_loc + q] |= unsigned(slot) << (16 * (k & 1));
                }
            }
        if (n_loc + long(halo[w].size()) > 65535) return false;
    }
    int hmax = 0, nmax = 0;
    for (int w = 0; w < P; ++w) {
        hmax = std::max(hmax, int(halo[w].size()));
        int c = 0;
        for (int u = 0; u < P; ++u) c += link[w][u];
        nmax = std::max(nmax, c);
    }
    if (nmax > 64) return false;
    if (csl::partitioned_lds_bytes(int(n_loc), hmax, block, false) > h->lds_limit) return false;
    o.dbuf = csl::partitioned_lds_bytes(int(n_loc), hmax, block, true) <= h->lds_limit ? 1 : 0;
    o.P = P;
    o.c_per = c_per;
    o.cpad = cpad;
    o.n_loc = int(n_loc);
    o.n_halo_max = hmax;
    o.n_nbr_max = nmax;
    o.block = block;
    o.spt = spt;
    o.halo_src.assign(size_t(P) * std::max(hmax, 1), 0);
    o.n_halo.assign(P, 0);
    o.nbr.assign(size_t(P) * std::max(nmax, 1), -1);
    for (int w = 0; w < P; ++w) {
        o.n_halo[w] = int(halo[w].size());
        std::copy(halo[w].begin(), halo[w].end(), o.halo_src.begin() + size_t(w) * std::max(hmax, 1));
        int c = 0;
        for (int u = 0; u < P; ++u)
            if (link[w][u]) o.nbr[size_t(w) * std::max(nmax, 1) + c++] = u;
    }
    return true;
}

// Tables only the partitioned kernel can run: translation-invariant bond
// matrices and per-sublattice fields / onsite matrices.
bool partition_eligible(const csl_lattice* h) { return h->L.uniform && h->L.site_uniform; }

// Builds (or keeps) the device plan for P parts; P = 0 picks the fewest
// parts that fit.  False when none fits.
bool ensure_partition(csl_lattice* h, int P) {
    if (!partition_eligible(h)) return false;
    if (h->part.P && (P == 0 || P == h->part.P)) return true;
    PartHost o;
    bool ok = false;
    if (P > 0) {
        ok = plan_partition(h, P, o);
    } else {
        const size_t state = 3 * sizeof(double) * size_t(h->L.NP);
        const int p0 = int(std::max<size_t>(1, (state + h->lds_limit - 1) / h->lds_limit));
        for (int p = p0; p <= std::min(h->n_cu, h->L.n_cells) && !ok; ++p) ok = plan_partition(h, p, o);
    }
    if (!ok) return false;
    hipStream_t s = h->stream;
    PartPlan& d = h->part;
    d.P = o.P;
    d.c_per = o.c_per;
    d.cpad = o.cpad;
    d.n_loc = o.n_loc;
    d.n_halo_max = o.n_halo_max;
    d.n_nbr_max = o.n_nbr_max;
    d.block = o.block;
    d.spt = o.spt;
    d.dbuf = o.dbuf;
    d.col.upload(o.col.data(), o.col.size(), s);
    d.exported.upload(o.exported.data(), o.exported.size(), s);
    d.halo_src.upload(o.halo_src.data(), o.halo_src.size(), s);
    d.n_halo.upload(o.n_halo.data(), o.n_halo.size(), s);
    d.nbr.upload(o.nbr.data(), o.nbr.size(), s);
    hip_check(hipStreamSynchronize(s), "partition upload");
    return true;
}

// Runs the partitioned kernel over B trajectories (plan built); device
// outputs are left in h->rec_mags / h->rec_energy / h->bt_final.
void run_partitioned(csl_lattice* h, size_t B, const double* d_init, size_t init_stride, const double* d_tp,
                     const double* d_amp, const std::vector<double>& grid, double dt, size_t si, bool want_energy,
                     bool want_final, double* ms_out) {
    const PartPlan& pl = h->part;
    const size_t n_steps = grid.size() - 1;
    const size_t n_rec = records_for(n_steps, si);
    if (n_steps > size_t(std::numeric_limits<int>::max()) || n_rec > size_t(std::numeric_limits<int>::max()))
        fail(CSL_ERR_ARG, "too many steps");
    hipStream_t s = h->stream;
    csl::LatDev L = h->L;
    csl::PartArgs A{};
    A.B = int(B);
    A.init = d_init;
    A.init_stride = init_stride;
    A.t_pulse = d_tp;
    A.amp = d_amp;
    A.n_steps = int(n_steps);
    A.save_interval = int(si);
    A.n_rec = int(n_rec);
    A.dt = dt;
    if (h->rep_B) {
        if (B != h->rep_B) fail(CSL_ERR_ARG, "batch size differs from the per-replica Hamiltonians set");
        L.mats = h->rep_mats.p;
        L.field = h->rep_field.p;
        A.mats_stride = h->rep_mats_stride;
        A.field_stride = h->rep_field_stride;
    }
    A.P = pl.P;
    A.c_per = pl.c_per;
    A.cpad = pl.cpad;
    A.n_loc = pl.n_loc;
    A.n_halo_max = pl.n_halo_max;
    A.n_nbr_max = pl.n_nbr_max;
    A.dbuf = pl.dbuf;
    A.col = pl.col.p;
    A.halo_src = pl.halo_src.p;
    A.n_halo = pl.n_halo.p;
    A.exported = pl.exported.p;
    A.nbr = pl.nbr.p;
    int occ = 0;
    hip_check(csl::launch_partitioned(L, A, pl.block, pl.spt, s, &occ), "partitioned occupancy");
    const size_t co = size_t(std::max(occ, 0)) * size_t(h->n_cu);
    if (co < size_t(pl.P)) fail(CSL_ERR_UNSUPPORTED, "the parts of one trajectory do not fit the GPU at once");
    const int groups = int(std::min(B, co / size_t(pl.P)));
    const int grid_wg = groups * pl.P;
    A.groups = groups;
    A.xcd_map = (grid_wg % 8 == 0 && (grid_wg / 8) % pl.P == 0) ? 1 : 0;
    h->tgrid.upload(grid.data(), grid.size(), s);
    A.tgrid = h->tgrid.p;
    h->part_exch.ensure(size_t(groups) * 6 * size_t(L.NP));
    h->part_ver.ensure(size_t(grid_wg));
    h->ps_status.ensure(1);
    h->part_obs.ensure(B * n_rec * size_t(pl.P) * 10);
    h->rec_mags.ensure(n_rec * B * 9);
    if (want_energy) h->rec_energy.ensure(n_rec * B);
    if (want_final) h->bt_final.ensure(B * h->N * 3);
    A.exch = h->part_exch.p;
    A.ver = h->part_ver.p;
    A.status = h->ps_status.p;
    A.part_obs = h->part_obs.p;
    A.final_states = want_final ? h->bt_final.p : nullptr;
    A.want_energy = want_energy ? 1 : 0;
    hip_check(csl::launch_fill_u32(h->part_ver.p, 0u, size_t(grid_wg), s), "fill");
    hip_check(csl::launch_fill_u32(h->ps_status.p, 0u, 1, s), "fill");
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, s), "event record");
    hip_check(csl::launch_partitioned(L, A, pl.block, pl.spt, s, nullptr), "k_partitioned launch");
    hip_check(hipEventRecord(e1, s), "event record");
    hip_check(csl::launch_part_reduce(h->part_obs.p, pl.P, int(B), int(n_rec), 1.0 / double(h->N), h->rec_mags.p,
                                      want_energy ? h->rec_energy.p : nullptr, s),
              "k_part_reduce");
    hip_check(hipStreamSynchronize(s), "k_partitioned");
    unsigned st = 0;
    hip_check(hipMemcpy(&st, h->ps_status.p, sizeof(unsigned), hipMemcpyDeviceToHost), "D2H");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) fail(CSL_ERR_HIP, "a part of the partitioned launch waited too long for its neighbours' stage values");
    h->last_ms = ms;
    h->last_block = pl.block;
    h->last_spt = pl.spt;
    h->last_dbuf = pl.dbuf ? csl::kPersistDoubleBuffer | csl::kPersistUnrollStages : csl::kPersistUnrollStages;
    h->last_variant = csl::kVariantPartitioned;
    h->last_launches = 1;
    h->last_grid = grid_wg;
    h->last_parts = pl.P;
    h->last_groups = groups;
    if (ms_out) *ms_out = ms;
}

// Plain path: grid-wide stage kernels on SoA buffers, B trajectories in
// grid.y; X must already hold the SoA initial state(s).
void run_stages(csl_lattice* h, size_t B, const double* d_tp, const double* d_amp, bool drive_on,
                const std::vector<double>& grid, double dt, size_t si, bool want_energy,
                double* d_states_out /* [n_rec][B][N*3] or null */,
                const std::function<void(size_t)>& at_step = {} /* called with X = x_n */,
                bool use_graph = false) {
    h->last_variant = 0;
    const size_t n_steps = grid.size() - 1;
    const size_t n_rec = records_for(n_steps, si);
    hipStream_t s = h->stream;
    const csl::LatDev& L = h->L;
    const size_t plane = B * 3 * size_t(L.NP);
    h->ACC.ensure(plane);
    h->Y1.ensure(plane);
    h->Y2.ensure(plane);
    const int nblk = csl::stage_blocks(L);
    h->partials.ensure(B * size_t(nblk) * 10);
    h->rec_mags.ensure(n_rec * B * 9);
    if (want_energy) h->rec_energy.ensure(n_rec * B);
    h->tgrid.upload(grid.data(), grid.size(), s);
    if (drive_on) h->fac.ensure(std::max<size_t>(n_steps * 3 * B * 2, 1));
    const double inv_n = 1.0 / double(h->N);
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    size_t launches = 0;
    // The whole launch sequence of the integration (drive table, 4 stage
    // kernels per step, record reductions, final record).  With use_graph it
    // is captured once into a hipGraph and replayed while the key (shapes,
    // device buffers, lattice parameters) is unchanged; inputs are refreshed
    // in the same device buffers before every replay.
    auto issue = [&](bool timed) {
        if (drive_on)
            hip_check(csl::launch_drive_table(h->tgrid.p, int(n_steps), dt, int(B), d_tp, d_amp, L.width, L.freq,
                                              h->fac.p, s),
                      "drive table");
        if (timed) hip_check(hipEventRecord(e0, s), "event record");
        launches = 0;
    for (size_t n = 0; n < n_steps; ++n) {
        const bool rec = (n % si) == 0;
        const size_t ri = n / si;
        if (at_step) at_step(n);
        if (rec && d_states_out)
            hip_check(csl::launch_perm_to_aos(L, h->X.p, d_states_out + ri * B * h->N * 3, int(B), s),
                      "state record");
        for (int st = 0; st < 4; ++st) {
            csl::StageArgs A{};
            A.B = int(B);
            A.stage = st;
            A.x = h->X.p;
            A.acc = h->ACC.p;
            A.ycur = st == 0 ? h->X.p : (st == 2 ? h->Y2.p : h->Y1.p);
            A.ynext = st == 0 ? h->Y1.p : (st == 1 ? h->Y2.p : (st == 2 ? h->Y1.p : h->Y2.p));
            const int stt = st == 0 ? 0 : (st == 3 ? 2 : 1);
            A.fac = drive_on ? h->fac.p + ((n * 3 + stt) * B) * 2 : nullptr;
            A.dt = dt;
            A.rec = (st == 0 && rec) ? 1 : 0;
            A.partials = h->partials.p;
            hip_check(csl::launch_stage(L, A, s), "k_stage launch");
            ++launches;
        }
        if (rec)
            hip_check(csl::launch_reduce_obs(h->partials.p, nblk, int(B), int(ri), inv_n, h->rec_mags.p,
                                             want_energy ? h->rec_energy.p : nullptr, s),
                      "reduce");
    }
        if (timed) hip_check(hipEventRecord(e1, s), "event record");
        hip_check(csl::launch_observe(L, h->X.p, h->partials.p, int(B), s), "observe");
        hip_check(csl::launch_reduce_obs(h->partials.p, nblk, int(B), int(n_rec - 1), inv_n, h->rec_mags.p,
                                         want_energy ? h->rec_energy.p : nullptr, s),
                  "reduce");
        if (d_states_out)
            hip_check(csl::launch_perm_to_aos(L, h->X.p, d_states_out + (n_rec - 1) * B * h->N * 3, int(B), s),
                      "state record");
    };
    if (!use_graph) {
        issue(true);
    } else {
        if (at_step) fail(CSL_ERR_UNSUPPORTED, "per-step host hooks cannot be replayed from a graph");
        // key: everything the captured launches bake in (the lattice block is
        // copied from the handle, whose bytes are stable between calls)
        std::vector<unsigned char> key;
        auto put = [&](const void* p, size_t n) {
            const auto* c = static_cast<const unsigned char*>(p);
            key.insert(key.end(), c, c + n);
        };
        put(&h->L, sizeof(csl::LatDev));
        const size_t shape[4] = {n_steps, si, B, size_t((drive_on ? 1 : 0) | (want_energy ? 2 : 0))};
        put(shape, sizeof(shape));
        put(&dt, sizeof(dt));
        const void* ptrs[] = {h->X.p, h->ACC.p, h->Y1.p, h->Y2.p, h->partials.p, h->rec_mags.p,
                              want_energy ? h->rec_energy.p : nullptr, drive_on ? h->fac.p : nullptr,
                              h->tgrid.p, d_tp, d_amp, d_states_out};
        put(ptrs, sizeof(ptrs));
        if (!h->graph_exec || key != h->graph_key) {
            if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
            h->graph_exec = nullptr;
            hipGraph_t g = nullptr;
            hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
            try {
                issue(false);
            } catch (...) {
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hip_check(hipStreamEndCapture(s, &g), "end capture");
            const hipError_t ie = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hip_check(ie, "graph instantiate");
            h->graph_key = key;
        }
        hip_check(hipEventRecord(e0, s), "event record");
        hip_check(hipGraphLaunch(h->graph_exec, s), "graph launch");
        hip_check(hipEventRecord(e1, s), "event record");
        launches = 4 * n_steps;  // stage launches of the replayed sequence
    }
    hip_check(hipStreamSynchronize(s), "stage integration");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->last_ms = launches ? ms / double(launches) : 0.0;
    h->last_launches = launches;
    h->last_block = -1;  // grid-wide stage kernels
    h->last_spt = 0;
    h->last_dbuf = 0;
}

// Reference drive factor f = A exp(-((t - t_c) / 2w)^2) cos(w (t - t_c))
// (lattice_md.cpp:45-56), evaluated on the host for one stage time.
double host_drive_factor(double amp, double t, double tc, double width, double freq) {
    if (amp == 0.0) return 0.0;
    const double d = t - tc;
    const double a = d / (2.0 * width);
    return amp * std::exp(-(a * a)) * std::cos(freq * d);
}

// Method name of the reference GPU stepper -> stage table; an unknown name is
// reported on stderr and runs ssprk53, as gpu::ode::step does
// (integrator.cuh:454-458).
Tableau method_tableau(const char* method) {
    bool known = true;
    Tableau T = tableau_for(method, &known);
    if (!known)
        std::fprintf(stderr, "Warning: Unknown GPU integration method '%s', using ssprk53\n",
                     method ? method : "");
    return T;
}

// Generic fixed-step engine for the reference GPU stepper's methods other
// than rk4 (integrator.cuh:91-458), one trajectory in h->X (SoA): per stage
// one k_eval launch (the derivative of the stage input at t + c_s dt into
// K[s], drive factors from the host) and one k_rk_combine launch (the next
// stage input, or the result in place of x, from x and the K's of the same
// element).  Records as run_stages: M, E and optionally the state at every
// save_interval-th step start and after the last step.
void run_tableau(csl_lattice* h, const Tableau& T, const std::vector<double>& grid, double dt, size_t si,
                 bool want_energy, double* d_states_out) {
    const size_t n_steps = grid.size() - 1;
    const size_t n_rec = records_for(n_steps, si);
    hipStream_t s = h->stream;
    const csl::LatDev& L = h->L;
    const size_t n = 3 * size_t(L.NP);
    const int S = T.S;
    h->KS.ensure(size_t(S) * n);
    h->Y1.ensure(n);
    // padded slots of K are never written by k_eval; keep them zero
    hip_check(hipMemsetAsync(h->KS.p, 0, size_t(S) * n * sizeof(double), s), "memset");
    hip_check(hipMemsetAsync(h->Y1.p, 0, n * sizeof(double), s), "memset");
    const int nblk = csl::stage_blocks(L);
    h->partials.ensure(size_t(nblk) * 10);
    h->rec_mags.ensure(n_rec * 9);
    if (want_energy) h->rec_energy.ensure(n_rec);
    const bool drive_on = h->amp != 0.0;
    const double inv_n = 1.0 / double(h->N);
    // combination rows: stage inputs 1..S-1 and the result
    std::vector<csl::RkCombine> rows(static_cast<size_t>(S));
    for (int st = 0; st < S; ++st) {
        const Lin& l = st + 1 < S ? T.in[st + 1] : T.out;
        csl::RkCombine& c = rows[size_t(st)];
        c.g = l.g;
        c.m = lin_terms(l, st + 1, dt, c.idx, c.w);
    }
    auto observe = [&](size_t ri) {
        hip_check(csl::launch_observe(L, h->X.p, h->partials.p, 1, s), "observe");
        hip_check(csl::launch_reduce_obs(h->partials.p, nblk, 1, int(ri), inv_n, h->rec_mags.p,
                                         want_energy ? h->rec_energy.p : nullptr, s),
                  "reduce");
        if (d_states_out)
            hip_check(csl::launch_perm_to_aos(L, h->X.p, d_states_out + ri * h->N * 3, 1, s), "state record");
    };
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, s), "event record");
    for (size_t k = 0; k < n_steps; ++k) {
        if (k % si == 0) observe(k / si);
        const double t = grid[k];
        for (int st = 0; st < S; ++st) {
            const double ts = t + T.c[st] * dt;
            const double f1 = drive_on ? host_drive_factor(h->amp, ts, h->t1, h->width, h->freq) : 0.0;
            const double f2 = drive_on ? host_drive_factor(h->amp, ts, h->t2, h->width, h->freq) : 0.0;
            const double* in = st == 0 ? h->X.p : h->Y1.p;
            hip_check(csl::launch_eval(L, in, nullptr, h->KS.p + size_t(st) * n, f1, f2, drive_on ? 1 : 0, s),
                      "k_eval");
            double* out = st + 1 < S ? h->Y1.p : h->X.p;
            hip_check(csl::launch_rk_combine(out, h->X.p, h->KS.p, n, rows[size_t(st)], s), "k_rk_combine");
        }
    }
    hip_check(hipEventRecord(e1, s), "event record");
    observe(n_rec - 1);
    hip_check(hipStreamSynchronize(s), "stage integration");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->last_launches = n_steps * size_t(S) * 2;
    h->last_ms = h->last_launches ? ms / double(h->last_launches) : 0.0;
    h->last_block = -1;
    h->last_spt = 0;
    h->last_dbuf = 0;
}

}  // namespace

extern "C" {

const char* csl_last_error(void) { return g_err.c_str(); }

int csl_version(int* major, int* minor) {
    return guarded([&] {
        if (!major || !minor) fail(CSL_ERR_NULL, "null output");
        *major = 0;
        *minor = 1;
    });
}

int csl_device_count(int* count) {
    return guarded([&] {
        if (!count) fail(CSL_ERR_NULL, "null output");
        int c = 0;
        hip_check(hipGetDeviceCount(&c), "hipGetDeviceCount");
        *count = c;
    });
}

int csl_create_lattice(size_t N, size_t D, size_t n_atoms, size_t max_bi, const double* field,
                       const double* onsite, const double* bilinear, const size_t* partners,
                       const size_t* counts, csl_lattice** out) {
    return guarded([&] {
        if (!out || !field || !onsite || !counts || (max_bi && (!bilinear || !partners)))
            fail(CSL_ERR_NULL, "null table pointer");
        *out = nullptr;
        if (D != 3) fail(CSL_ERR_UNSUPPORTED, "spin_dim != 3 is handled by the mixed-lattice API");
        // padded reference layout -> CSR (lattice.h:5459-5479)
        std::vector<size_t> off(N + 1, 0), part;
        std::vector<double> J;
        for (size_t i = 0; i < N; ++i) {
            if (counts[i] > max_bi) fail(CSL_ERR_ARG, "bond count exceeds max_bilinear");
            off[i + 1] = off[i] + counts[i];
            for (size_t n = 0; n < counts[i]; ++n) {
                part.push_back(partners[i * max_bi + n]);
                J.insert(J.end(), bilinear + (i * max_bi + n) * D * D, bilinear + (i * max_bi + n + 1) * D * D);
            }
        }
        csl_lattice* h = new_handle();
        try {
            build_tables(h, N, D, n_atoms, field, onsite, off, part, J.data());
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int csl_create_lattice_csr(size_t N, size_t D, size_t n_atoms, const double* field,
                           const double* onsite, const size_t* offset, const size_t* partner,
                           const double* J, csl_lattice** out) {
    return guarded([&] {
        if (!out || !field || !onsite || !offset) fail(CSL_ERR_NULL, "null table pointer");
        *out = nullptr;
        std::vector<size_t> off(offset, offset + N + 1);
        if (off[N] && (!partner || !J)) fail(CSL_ERR_NULL, "null bond table");
        std::vector<size_t> part(partner, partner + off[N]);
        csl_lattice* h = new_handle();
        try {
            build_tables(h, N, D, n_atoms, field, onsite, off, part, J);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int csl_destroy_lattice(csl_lattice* h) {
    return guarded([&] {
        if (!h) fail(CSL_ERR_NULL, "null lattice handle");
        (void)hipSetDevice(h->device);
        (void)hipStreamSynchronize(h->stream);
        delete h;
    });
}

int csl_lattice_info(const csl_lattice* h, size_t* N, size_t* D, size_t* n_atoms, size_t* n_padded,
                     size_t* max_bonds, int* uniform) {
    return guarded([&] {
        if (!h) fail(CSL_ERR_NULL, "null lattice handle");
        if (N) *N = h->N;
        if (D) *D = h->D;
        if (n_atoms) *n_atoms = h->n_atoms;
        if (n_padded) *n_padded = size_t(h->L.NP);
        if (max_bonds) *max_bonds = size_t(h->L.zmax);
        if (uniform) *uniform = h->L.uniform;
    });
}

int csl_set_pulse(csl_lattice* h, const double* field_drive, double amp, double width, double freq,
                  double t_pulse_1, double t_pulse_2) {
    return guarded([&] {
        check_handle(h);
        if (!field_drive) fail(CSL_ERR_NULL, "null field_drive");
        check_finite(field_drive, 2 * h->n_atoms * 3, "field_drive");
        if (!std::isfinite(amp) || !std::isfinite(width) || !std::isfinite(freq) ||
            !std::isfinite(t_pulse_1) || !std::isfinite(t_pulse_2))
            fail(CSL_ERR_NONFINITE, "non-finite pulse parameter");
        h->h_fdrive.assign(field_drive, field_drive + 2 * h->n_atoms * 3);
        h->fdrive.upload(h->h_fdrive.data(), h->h_fdrive.size(), h->stream);
        hip_check(hipStreamSynchronize(h->stream), "pulse upload");
        h->amp = amp;
        h->width = width;
        h->freq = freq;
        h->t1 = t_pulse_1;
        h->t2 = t_pulse_2;
        h->L.width = width;
        h->L.freq = freq;
    });
}

int csl_set_damping(csl_lattice* h, double alpha) {
    return guarded([&] {
        check_handle(h);
        if (!std::isfinite(alpha) || alpha < 0.0) fail(CSL_ERR_ARG, "alpha must be finite and >= 0");
        h->L.alpha = alpha;
    });
}

int csl_set_frames(csl_lattice* h, const double* frames, const double* afm_signs) {
    return guarded([&] {
        check_handle(h);
        if (frames) {
            check_finite(frames, h->n_atoms * 9, "frames");
            h->h_frames.assign(frames, frames + h->n_atoms * 9);
        }
        if (afm_signs) {
            check_finite(afm_signs, h->n_atoms, "afm_signs");
            h->h_afm.assign(afm_signs, afm_signs + h->n_atoms);
        }
        h->frames.upload(h->h_frames.data(), h->h_frames.size(), h->stream);
        h->afm.upload(h->h_afm.data(), h->h_afm.size(), h->stream);
        hip_check(hipStreamSynchronize(h->stream), "frames upload");
    });
}

int csl_set_spins(csl_lattice* h, const double* spins, size_t n) {
    return guarded([&] {
        check_handle(h);
        if (!spins) fail(CSL_ERR_NULL, "null spins");
        if (n != h->N * h->D) fail(CSL_ERR_ARG, "spin array size must be lattice_size*spin_dim");
        check_finite(spins, n, "spins");
        h->state.upload(spins, n, h->stream);
        hip_check(hipStreamSynchronize(h->stream), "spin upload");
        h->has_state = true;
    });
}

int csl_get_spins(csl_lattice* h, double* spins, size_t n) {
    return guarded([&] {
        check_handle(h);
        if (!spins) fail(CSL_ERR_NULL, "null spins");
        if (!h->has_state) fail(CSL_ERR_NOSTATE, "spins were never set");
        if (n != h->N * h->D) fail(CSL_ERR_ARG, "spin array size must be lattice_size*spin_dim");
        copy_back(spins, h->state.p, n, h->stream);
        hip_check(hipStreamSynchronize(h->stream), "spin download");
    });
}

// Pulse-window segments of the pulse drivers (W3, build_pulse_segments,
// pulse_chunking.h:82-161): every window [c - window, c + window] is snapped
// to the grid T0 + k*T_step (nearest point) and clipped to [T0, Te], where
// Te is the last grid point <= T1; overlapping or touching windows merge;
// the stretches between them are free segments with the step hint
// min(max(T_step, free_dt_factor*T_step), length).  Output rows are
// (t0, t1, dt_init), contiguous from T0 to Te; empty when Te <= T0.
static std::vector<std::array<double, 3>> pulse_segments(double T0, double T1, const double* centers, size_t nc,
                                                         double window, double T_step, double free_dt_factor) {
    std::vector<std::array<double, 3>> out;
    if (!(T1 > T0) || !(T_step > 0.0)) return out;
    const double Te = T0 + std::floor((T1 - T0) / T_step + 1e-9) * T_step;
    if (!(Te > T0)) return out;
    auto on_grid = [&](double t) { return T0 + std::round((t - T0) / T_step) * T_step; };
    std::vector<std::pair<double, double>> win;
    for (size_t i = 0; i < nc; ++i) {
        const double lo = std::max(T0, on_grid(centers[i] - window));
        const double hi = std::min(Te, on_grid(centers[i] + window));
        if (lo < hi) win.emplace_back(lo, hi);
    }
    std::sort(win.begin(), win.end());
    std::vector<std::pair<double, double>> active;
    for (const auto& w : win) {
        if (!active.empty() && w.first <= active.back().second)
            active.back().second = std::max(active.back().second, w.second);
        else
            active.push_back(w);
    }
    const double free_hint = std::max(T_step, free_dt_factor * T_step);
    double at = T0;
    for (const auto& w : active) {
        if (at < w.first) out.push_back({at, w.first, std::min(free_hint, w.first - at)});
        out.push_back({w.first, w.second, T_step});
        at = w.second;
    }
    if (at < Te) out.push_back({at, Te, std::min(free_hint, Te - at)});
    return out;
}

// half-width of a pulse window in units of the width, and the free-segment
// step-hint factor (pulse_chunking.h:173,178)
constexpr double kPulseWindowWidths = 9.0;
constexpr double kFreeStepFactor = 20.0;

// observer calls of one integrate_const run over [s0, s1] with step dt
static size_t const_calls(double s0, double s1, double dt) {
    size_t k = 0;
    while ((s0 + double(k) * dt + dt) - s1 <= 2.220446049250313e-16) ++k;
    return k + 1;
}

int csl_pulse_segments(double T_start, double T_end, const double* centers, size_t n_centers, double window,
                       double T_step, double free_dt_factor, size_t max_out, double* out, size_t* n_out) {
    return guarded([&] {
        if (n_centers && !centers) fail(CSL_ERR_NULL, "null centers");
        if (!std::isfinite(T_start) || !std::isfinite(T_end) || !std::isfinite(window) || !std::isfinite(T_step))
            fail(CSL_ERR_ARG, "non-finite time argument");
        check_finite(centers, n_centers, "centers");
        const auto segs = pulse_segments(T_start, T_end, centers, n_centers, window, T_step, free_dt_factor);
        if (n_out) *n_out = segs.size();
        if (segs.size() > max_out) fail(CSL_ERR_CAPACITY, "max_out is smaller than the number of segments");
        if (segs.size() && !out) fail(CSL_ERR_NULL, "null out");
        for (size_t i = 0; i < segs.size(); ++i)
            for (int j = 0; j < 3; ++j) out[3 * i + j] = segs[i][j];
    });
}

int csl_count_records(double T0, double T1, double dt, size_t save_interval, int tgrid,
                      size_t* n_steps, size_t* n_rec) {
    return guarded([&] {
        if (save_interval == 0) fail(CSL_ERR_ARG, "save_interval must be >= 1");
        auto g = make_grid(T0, T1, dt, tgrid);
        if (n_steps) *n_steps = g.size() - 1;
        if (n_rec) *n_rec = records_for(g.size() - 1, save_interval);
    });
}

int csl_integrate(csl_lattice* h, double T0, double T1, double dt, size_t save_interval,
                  const char* method, int tgrid, int engine, size_t max_rec, double* times,
                  double* states, double* mags, double* energy, size_t* n_rec_out) {
    return guarded([&] {
        check_handle(h);
        if (!h->has_state) fail(CSL_ERR_NOSTATE, "spins were never set");
        if (save_interval == 0) fail(CSL_ERR_ARG, "save_interval must be >= 1");
        const bool rk4 = method_is_rk4(method);
        if (!rk4 && engine != 0 && engine != 1)
            fail(CSL_ERR_UNSUPPORTED, "methods other than \"rk4\" run on the stage engine (engine 0 or 1)");
        auto grid = make_grid(T0, T1, dt, tgrid);
        const size_t n_steps = grid.size() - 1;
        const size_t n_rec = records_for(n_steps, save_interval);
        if (n_rec_out) *n_rec_out = n_rec;
        if ((times || states || mags || energy) && max_rec < n_rec)
            fail(CSL_ERR_CAPACITY, "max_rec is smaller than the number of records");
        hipStream_t s = h->stream;
        const csl::LatDev& L = h->L;
        const bool drive_on = h->amp != 0.0;
        const double tp[2] = {h->t1, h->t2};
        const double am[2] = {h->amp, h->amp};
        h->bt_tp.upload(tp, 2, s);
        h->bt_amp.upload(am, 2, s);
        bool persistent = false;
        if (engine == 2) persistent = true;
        else if (engine == 0) persistent = rk4 && !states && !persistent_candidates(h).empty();
        else if (engine != 1 && engine != 3) fail(CSL_ERR_ARG, "engine must be 0, 1, 2 or 3");
        if (persistent && states) fail(CSL_ERR_ARG, "the persistent engine does not record full states");
        if (!rk4) {
            const Tableau T = method_tableau(method);
            h->X.ensure(3 * size_t(L.NP));
            hip_check(csl::launch_perm_to_soa(L, h->state.p, 0, h->X.p, 1, s), "perm");
            if (states) h->rec_states.ensure(n_rec * h->N * 3);
            run_tableau(h, T, grid, dt, save_interval, energy != nullptr, states ? h->rec_states.p : nullptr);
            hip_check(csl::launch_perm_to_aos(L, h->X.p, h->state.p, 1, s), "perm");
        } else if (persistent) {
            TuneVal tv = pick_persistent(h, 1, 0);
            run_persistent(h, 1, h->state.p, 0, h->bt_tp.p, h->bt_amp.p, grid, dt, save_interval,
                           energy != nullptr, true, tv, nullptr);
            hip_check(hipMemcpyAsync(h->state.p, h->bt_final.p, h->N * 3 * sizeof(double),
                                     hipMemcpyDeviceToDevice, s),
                      "final state");
        } else {
            h->X.ensure(3 * size_t(L.NP));
            hip_check(csl::launch_perm_to_soa(L, h->state.p, 0, h->X.p, 1, s), "perm");
            if (states) h->rec_states.ensure(n_rec * h->N * 3);
            run_stages(h, 1, h->bt_tp.p, h->bt_amp.p, drive_on, grid, dt, save_interval, energy != nullptr,
                       states ? h->rec_states.p : nullptr, {}, engine == 3);
            hip_check(csl::launch_perm_to_aos(L, h->X.p, h->state.p, 1, s), "perm");
        }
        copy_back(mags, h->rec_mags.p, n_rec * 9, s);
        copy_back(energy, h->rec_energy.p, n_rec, s);
        copy_back(states, h->rec_states.p, n_rec * h->N * 3, s);
        hip_check(hipStreamSynchronize(s), "integrate");
        if (times) {
            size_t r = 0;
            for (size_t n = 0; n < n_steps; n += save_interval) times[r++] = grid[n];
            times[r] = grid[n_steps];
        }
    });
}

int csl_step(csl_lattice* h, double t, double dt, const char* method) {
    return guarded([&] {
        check_handle(h);
        if (!h->has_state) fail(CSL_ERR_NOSTATE, "spins were never set");
        if (!(dt > 0.0) || !std::isfinite(t)) fail(CSL_ERR_ARG, "need finite t and dt > 0");
        hipStream_t s = h->stream;
        const csl::LatDev& L = h->L;
        std::vector<double> grid = {t, t + dt};
        if (!method_is_rk4(method)) {
            const Tableau T = method_tableau(method);
            h->X.ensure(3 * size_t(L.NP));
            hip_check(csl::launch_perm_to_soa(L, h->state.p, 0, h->X.p, 1, s), "perm");
            run_tableau(h, T, grid, dt, 1, false, nullptr);
            hip_check(csl::launch_perm_to_aos(L, h->X.p, h->state.p, 1, s), "perm");
            hip_check(hipStreamSynchronize(s), "step");
            return;
        }
        const double tp[2] = {h->t1, h->t2};
        const double am[2] = {h->amp, h->amp};
        h->bt_tp.upload(tp, 2, s);
        h->bt_amp.upload(am, 2, s);
        h->X.ensure(3 * size_t(L.NP));
        hip_check(csl::launch_perm_to_soa(L, h->state.p, 0, h->X.p, 1, s), "perm");
        run_stages(h, 1, h->bt_tp.p, h->bt_amp.p, h->amp != 0.0, grid, dt, 1, false, nullptr);
        hip_check(csl::launch_perm_to_aos(L, h->X.p, h->state.p, 1, s), "perm");
        hip_check(hipStreamSynchronize(s), "step");
    });
}

int csl_local_field(csl_lattice* h, double* H, size_t n) {
    return guarded([&] {
        check_handle(h);
        if (!h->has_state) fail(CSL_ERR_NOSTATE, "spins were never set");
        if (!H || n != h->N * 3) fail(CSL_ERR_ARG, "output must hold lattice_size*spin_dim doubles");
        hipStream_t s = h->stream;
        const csl::LatDev& L = h->L;
        h->X.ensure(3 * size_t(L.NP));
        h->aux.ensure(3 * size_t(L.NP));
        hip_check(hipMemsetAsync(h->aux.p, 0, 3 * size_t(L.NP) * sizeof(double), s), "memset");
        hip_check(csl::launch_perm_to_soa(L, h->state.p, 0, h->X.p, 1, s), "perm");
        hip_check(csl::launch_eval(L, h->X.p, h->aux.p, nullptr, 0.0, 0.0, 0, s), "eval");
        h->Y1.ensure(h->N * 3);
        hip_check(csl::launch_perm_to_aos(L, h->aux.p, h->Y1.p, 1, s), "perm");
        copy_back(H, h->Y1.p, n, s);
        hip_check(hipStreamSynchronize(s), "local field");
    });
}

int csl_rhs(csl_lattice* h, double t, double* dsdt, size_t n) {
    return guarded([&] {
        check_handle(h);
        if (!h->has_state) fail(CSL_ERR_NOSTATE, "spins were never set");
        if (!dsdt || n != h->N * 3) fail(CSL_ERR_ARG, "output must hold lattice_size*spin_dim doubles");
        hipStream_t s = h->stream;
        const csl::LatDev& L = h->L;
        // drive factors evaluated exactly as the reference host does
        // (lattice_md.cpp:45-56)
        double f1 = 0.0, f2 = 0.0;
        const bool drive = h->amp != 0.0;
        if (drive) {
            const double d1 = t - h->t1, d2 = t - h->t2;
            f1 = h->amp * std::exp(-std::pow(d1 / (2.0 * h->width), 2)) * std::cos(h->freq * d1);
            f2 = h->amp * std::exp(-std::pow(d2 / (2.0 * h->width), 2)) * std::cos(h->freq * d2);
        }
        h->X.ensure(3 * size_t(L.NP));
        h->aux.ensure(3 * size_t(L.NP));
        hip_check(hipMemsetAsync(h->aux.p, 0, 3 * size_t(L.NP) * sizeof(double), s), "memset");
        hip_check(csl::launch_perm_to_soa(L, h->state.p, 0, h->X.p, 1, s), "perm");
        hip_check(csl::launch_eval(L, h->X.p, nullptr, h->aux.p, f1, f2, drive ? 1 : 0, s), "eval");
        h->Y1.ensure(h->N * 3);
        hip_check(csl::launch_perm_to_aos(L, h->aux.p, h->Y1.p, 1, s), "perm");
        copy_back(dsdt, h->Y1.p, n, s);
        hip_check(hipStreamSynchronize(s), "rhs");
    });
}

int csl_energy(csl_lattice* h, double* E) {
    return guarded([&] {
        check_handle(h);
        if (!E) fail(CSL_ERR_NULL, "null output");
        if (!h->has_state) fail(CSL_ERR_NOSTATE, "spins were never set");
        hipStream_t s = h->stream;
        const csl::LatDev& L = h->L;
        h->X.ensure(3 * size_t(L.NP));
        const int nblk = csl::stage_blocks(L);
        h->partials.ensure(size_t(nblk) * 10);
        h->rec_mags.ensure(9);
        h->rec_energy.ensure(1);
        hip_check(csl::launch_perm_to_soa(L, h->state.p, 0, h->X.p, 1, s), "perm");
        hip_check(csl::launch_observe(L, h->X.p, h->partials.p, 1, s), "observe");
        hip_check(csl::launch_reduce_obs(h->partials.p, nblk, 1, 0, 1.0, h->rec_mags.p, h->rec_energy.p, s),
                  "reduce");
        copy_back(E, h->rec_energy.p, 1, s);
        hip_check(hipStreamSynchronize(s), "energy");
    });
}

int csl_normalize(csl_lattice* h, double spin_length) {
    return guarded([&] {
        check_handle(h);
        if (!h->has_state) fail(CSL_ERR_NOSTATE, "spins were never set");
        hip_check(csl::launch_normalize(h->state.p, int(h->N), spin_length, h->stream), "normalize");
        hip_check(hipStreamSynchronize(h->stream), "normalize");
    });
}

int csl_magnetization(csl_lattice* h, double* mag_local, double* mag_staggered) {
    return guarded([&] {
        check_handle(h);
        if (!mag_local || !mag_staggered) fail(CSL_ERR_NULL, "null output");
        if (!h->has_state) fail(CSL_ERR_NOSTATE, "spins were never set");
        // lattice_gpu.cu:729-760 downloads the state and sums on the host
        // with sign (i % 2 == 0 ? 1 : -1); same arithmetic here.
        std::vector<double> st(h->N * 3);
        copy_back(st.data(), h->state.p, st.size(), h->stream);
        hip_check(hipStreamSynchronize(h->stream), "download");
        double ml[3] = {0, 0, 0}, ms[3] = {0, 0, 0};
        for (size_t i = 0; i < h->N; ++i) {
            const double sg = (i % 2 == 0) ? 1.0 : -1.0;
            for (int d = 0; d < 3; ++d) {
                ml[d] += st[i * 3 + d];
                ms[d] += st[i * 3 + d] * sg;
            }
        }
        for (int d = 0; d < 3; ++d) {
            mag_local[d] = ml[d] / double(h->N);
            mag_staggered[d] = ms[d] / double(h->N);
        }
    });
}

int csl_integrate_batched(csl_lattice* h, size_t B, const double* init, int init_per_replica,
                          const double* t_pulse, const double* pulse_amp, double T0, double T1,
                          double dt, size_t save_interval, const char* method, int tgrid, int block,
                          size_t max_rec, double* times, double* mags, double* energy,
                          double* final_states, size_t* n_rec_out) {
    return guarded([&] {
        check_handle(h);
        if (B == 0) fail(CSL_ERR_ARG, "batch size must be >= 1");
        if (!init || !t_pulse) fail(CSL_ERR_NULL, "null init / t_pulse");
        if (!method_is_rk4(method)) fail(CSL_ERR_UNSUPPORTED, "batched mode implements \"rk4\" only");
        if (save_interval == 0) fail(CSL_ERR_ARG, "save_interval must be >= 1");
        if (B > 65535) fail(CSL_ERR_ARG, "at most 65535 trajectories per call");
        auto grid = make_grid(T0, T1, dt, tgrid);
        const size_t n_steps = grid.size() - 1;
        const size_t n_rec = records_for(n_steps, save_interval);
        if (n_rec_out) *n_rec_out = n_rec;
        if ((times || mags || energy) && max_rec < n_rec)
            fail(CSL_ERR_CAPACITY, "max_rec is smaller than the number of records");
        const size_t nd = h->N * 3;
        const size_t n_init = init_per_replica ? B * nd : nd;
        check_finite(init, n_init, "initial state");
        check_finite(t_pulse, 2 * B, "t_pulse");
        hipStream_t s = h->stream;
        h->bt_init.upload(init, n_init, s);
        h->bt_tp.upload(t_pulse, 2 * B, s);
        std::vector<double> am(2 * B, h->amp);
        if (pulse_amp) {
            check_finite(pulse_amp, 2 * B, "pulse_amp");
            am.assign(pulse_amp, pulse_amp + 2 * B);
        }
        h->bt_amp.upload(am.data(), am.size(), s);
        bool drive_on = false;
        for (double a : am) drive_on |= (a != 0.0);
        if (block < -1) fail(CSL_ERR_ARG, "block must be >= -1");
        auto cands = persistent_candidates(h);
        const bool part = block == 0 && (h->part_req > 0 || (h->part_req == 0 && cands.empty())) &&
                          ensure_partition(h, std::max(h->part_req, 0));
        if (part) {
            run_partitioned(h, B, h->bt_init.p, init_per_replica ? nd : 0, h->bt_tp.p, h->bt_amp.p, grid, dt,
                            save_interval, energy != nullptr, final_states != nullptr, nullptr);
            copy_back(final_states, h->bt_final.p, B * nd, s);
        } else if (!cands.empty() && block >= 0) {
            TuneVal tv = pick_persistent(h, B, block);
            run_persistent(h, B, h->bt_init.p, init_per_replica ? nd : 0, h->bt_tp.p, h->bt_amp.p, grid, dt,
                           save_interval, energy != nullptr, final_states != nullptr, tv, nullptr);
            copy_back(final_states, h->bt_final.p, B * nd, s);
        } else {
            if (h->rep_B) fail(CSL_ERR_UNSUPPORTED, "per-replica Hamiltonians need the persistent kernel");
            const csl::LatDev& L = h->L;
            h->X.ensure(B * 3 * size_t(L.NP));
            hip_check(csl::launch_perm_to_soa(L, h->bt_init.p, init_per_replica ? nd : 0, h->X.p, int(B), s),
                      "perm");
            run_stages(h, B, h->bt_tp.p, h->bt_amp.p, drive_on, grid, dt, save_interval, energy != nullptr,
                       nullptr, {}, h->stage_graph != 0);
            if (final_states) {
                h->bt_final.ensure(B * nd);
                hip_check(csl::launch_perm_to_aos(L, h->X.p, h->bt_final.p, int(B), s), "perm");
                copy_back(final_states, h->bt_final.p, B * nd, s);
            }
        }
        copy_back(mags, h->rec_mags.p, n_rec * B * 9, s);
        copy_back(energy, h->rec_energy.p, n_rec * B, s);
        hip_check(hipStreamSynchronize(s), "integrate_batched");
        if (times) {
            size_t r = 0;
            for (size_t n = 0; n < n_steps; n += save_interval) times[r++] = grid[n];
            times[r] = grid[n_steps];
        }
    });
}

int csl_integrate_batched_noalloc(csl_lattice* h, size_t B, const double* d_init, int init_per_replica,
                                  const double* d_t_pulse, const double* d_pulse_amp, double T0, double T1,
                                  double dt, size_t save_interval, int tgrid, double* d_mags, double* d_energy,
                                  double* d_final_states, size_t* n_rec_out) {
    return guarded([&] {
        check_handle(h);
        if (B == 0) fail(CSL_ERR_ARG, "batch size must be >= 1");
        if (B > 65535) fail(CSL_ERR_ARG, "at most 65535 trajectories per call");
        if (!d_init || !d_t_pulse || !d_pulse_amp) fail(CSL_ERR_NULL, "null init / t_pulse / pulse_amp");
        if (save_interval == 0) fail(CSL_ERR_ARG, "save_interval must be >= 1");
        if (persistent_candidates(h).empty())
            fail(CSL_ERR_UNSUPPORTED, "the lattice does not fit one workgroup (no allocation-free path)");
        auto grid = make_grid(T0, T1, dt, tgrid);
        const size_t n_steps = grid.size() - 1;
        if (n_rec_out) *n_rec_out = records_for(n_steps, save_interval);
        unsigned char* sp = scratch_ptr();
        const TuneVal tv = pick_persistent(h, B, 0);
        const ExtBufs ext{d_mags, d_energy, d_final_states, sp, size_t(CSL_NOALLOC_SCRATCH_BYTES)};
        run_persistent(h, B, d_init, init_per_replica ? h->N * 3 : 0, d_t_pulse, d_pulse_amp, grid, dt,
                       save_interval, d_energy != nullptr, d_final_states != nullptr, tv, nullptr, &ext);
    });
}

int csl_integrate_multi(csl_lattice* const* hs, size_t n_lat, const size_t* B_of, const double* init,
                        const double* t_pulse, const double* pulse_amp, double T0, double T1, double dt,
                        size_t save_interval, int tgrid, size_t max_rec, double* times, double* mags,
                        double* energy, size_t* n_rec_out) {
    return guarded([&] {
        if (!hs || !B_of || !init || !t_pulse) fail(CSL_ERR_NULL, "null handles / B_of / init / t_pulse");
        if (n_lat == 0) fail(CSL_ERR_ARG, "need at least one lattice");
        size_t B = 0, n_init = 0, maxNP = 0, rep = 0;
        int zmax = 0;
        for (size_t i = 0; i < n_lat; ++i) {
            check_handle(hs[i]);
            if (hs[i]->device != hs[0]->device) fail(CSL_ERR_ARG, "all lattices must live on one device");
            if (hs[i]->rep_B) fail(CSL_ERR_UNSUPPORTED, "per-replica Hamiltonians are not combined with lattices");
            if (hs[i]->L.uniform != hs[0]->L.uniform)
                fail(CSL_ERR_UNSUPPORTED, "the lattices must all have translation-invariant tables or none");
            if (B_of[i] == 0) fail(CSL_ERR_ARG, "every lattice needs at least one trajectory");
            B += B_of[i];
            n_init += B_of[i] * hs[i]->N * 3;
            if (size_t(hs[i]->L.NP) > maxNP) {
                maxNP = size_t(hs[i]->L.NP);
                rep = i;
            }
            zmax = std::max(zmax, hs[i]->L.zmax);
        }
        if (B > 65535) fail(CSL_ERR_ARG, "at most 65535 trajectories per call");
        if (save_interval == 0) fail(CSL_ERR_ARG, "save_interval must be >= 1");
        csl_lattice* h = hs[0];
        auto grid = make_grid(T0, T1, dt, tgrid);
        const size_t n_steps = grid.size() - 1;
        const size_t n_rec = records_for(n_steps, save_interval);
        if (n_rec_out) *n_rec_out = n_rec;
        if ((times || mags || energy) && max_rec < n_rec)
            fail(CSL_ERR_CAPACITY, "max_rec is smaller than the number of records");
        if (n_steps > size_t(std::numeric_limits<int>::max())) fail(CSL_ERR_ARG, "too many steps");
        check_finite(init, n_init, "initial states");
        check_finite(t_pulse, 2 * B, "t_pulse");
        // launch shape: the fewest sites per thread whose block holds the
        // largest lattice and whose LDS holds every lattice
        csl::LatDev Lrep = hs[rep]->L;
        Lrep.zmax = zmax;
        const int flags = csl::kPersistDoubleBuffer | csl::kPersistUnrollStages | csl::kPersistMulti;
        int block = 0, spt = 0;
        size_t lds = 0;
        for (int c : {1, 2, 3, 4, 6}) {
            int bl = int((maxNP + size_t(c) - 1) / size_t(c));
            bl = (bl + 63) / 64 * 64;
            if (bl > csl::persistent_block_limit(c)) continue;
            size_t l = 0;
            for (size_t i = 0; i < n_lat; ++i) l = std::max(l, csl::persistent_lds_bytes(hs[i]->L, bl, true));
            if (l > h->lds_limit) continue;
            block = bl;
            spt = c;
            lds = l;
            break;
        }
        if (!block) fail(CSL_ERR_UNSUPPORTED, "the largest lattice does not fit one workgroup");
        (void)lds;
        hipStream_t s = h->stream;
        std::vector<csl::LatDev> lats(n_lat);
        std::vector<int> lat_of(B);
        std::vector<size_t> off(B);
        std::vector<double> am(2 * B);
        for (size_t i = 0, b = 0, o = 0; i < n_lat; ++i)
            for (size_t k = 0; k < B_of[i]; ++k, ++b) {
                lats[i] = hs[i]->L;
                lat_of[b] = int(i);
                off[b] = o;
                o += hs[i]->N * 3;
                am[2 * b] = am[2 * b + 1] = hs[i]->amp;
            }
        if (pulse_amp) {
            check_finite(pulse_amp, 2 * B, "pulse_amp");
            am.assign(pulse_amp, pulse_amp + 2 * B);
        }
        // every lattice's tables must be on the device before the launch on
        // the first lattice's stream
        for (size_t i = 1; i < n_lat; ++i) hip_check(hipStreamSynchronize(hs[i]->stream), "lattice tables");
        h->bt_init.upload(init, n_init, s);
        h->bt_tp.upload(t_pulse, 2 * B, s);
        h->bt_amp.upload(am.data(), am.size(), s);
        h->ps_lats.upload(lats.data(), n_lat, s);
        h->ps_latof.upload(lat_of.data(), B, s);
        h->ps_initoff.upload(off.data(), B, s);
        h->tgrid.upload(grid.data(), grid.size(), s);
        h->rec_mags.ensure(n_rec * B * 9);
        if (energy) h->rec_energy.ensure(n_rec * B);
        csl::PersistArgs A{};
        A.B = int(B);
        A.init = h->bt_init.p;
        A.t_pulse = h->bt_tp.p;
        A.amp = h->bt_amp.p;
        A.tgrid = h->tgrid.p;
        A.n_steps = int(n_steps);
        A.save_interval = int(save_interval);
        A.n_rec = int(n_rec);
        A.dt = dt;
        A.mags = h->rec_mags.p;
        A.energy = energy ? h->rec_energy.p : nullptr;
        A.lats = h->ps_lats.p;
        A.lat_of = h->ps_latof.p;
        A.init_off = h->ps_initoff.p;
        const int occ = h->split_batches ? csl::persistent_occupancy(Lrep, block, spt, flags) : 0;
        const size_t G = size_t(std::max(occ, 0)) * size_t(h->n_cu);
        h->last_grid = int(B);
        if (G > 0 && B > G && n_steps > 0) {
            size_t max_p = 1;
            const std::vector<int> sched = split_schedule(B, G, n_steps, max_p);
            h->ps_sched.upload(sched.data(), sched.size(), s);
            h->ps_saved.ensure(G * 3 * maxNP);
            h->ps_flags.ensure(G);
            h->ps_status.ensure(1);
            hip_check(hipMemsetAsync(h->ps_flags.p, 0, G * sizeof(unsigned), s), "memset");
            hip_check(hipMemsetAsync(h->ps_status.p, 0, sizeof(unsigned), s), "memset");
            A.sched = h->ps_sched.p;
            A.max_pieces = int(max_p);
            A.grid = int(G);
            A.saved = h->ps_saved.p;
            A.flags = h->ps_flags.p;
            A.status = h->ps_status.p;
            h->last_grid = int(G);
        }
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        hip_check(hipEventRecord(e0, s), "event record");
        hip_check(csl::launch_persistent(Lrep, A, block, spt, flags, s), "k_persistent launch");
        hip_check(hipEventRecord(e1, s), "event record");
        hip_check(hipEventSynchronize(e1), "k_persistent");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        h->last_ms = ms;
        h->last_block = block;
        h->last_spt = spt;
        h->last_dbuf = flags;
        h->last_launches = 1;
        if (A.sched) {
            unsigned st = 0;
            hip_check(hipMemcpy(&st, h->ps_status.p, sizeof(unsigned), hipMemcpyDeviceToHost), "D2H");
            if (st) fail(CSL_ERR_HIP, "a workgroup of the split batch waited too long for a trajectory's state");
        }
        copy_back(mags, h->rec_mags.p, n_rec * B * 9, s);
        copy_back(energy, h->rec_energy.p, n_rec * B, s);
        hip_check(hipStreamSynchronize(s), "integrate_multi");
        if (times) {
            size_t r = 0;
            for (size_t n = 0; n < n_steps; n += save_interval) times[r++] = grid[n];
            times[r] = grid[n_steps];
        }
    });
}

int csl_integrate_dopri5(csl_lattice* h, size_t B, const double* init, int init_per_replica,
                         const double* t_pulse, const double* pulse_amp, double T0, double T1, double dt,
                         int control, double abs_tol, double rel_tol, size_t save_interval, int tgrid,
                         const double* segments, const size_t* n_segments, size_t max_segments,
                         size_t max_attempts, size_t max_rec, double* times, double* mags, double* energy,
                         double* final_states, size_t* n_rec_out, size_t* n_acc_out, size_t* n_rej_out) {
    return guarded([&] {
        check_handle(h);
        if (B == 0) fail(CSL_ERR_ARG, "batch size must be >= 1");
        if (!init) fail(CSL_ERR_NULL, "null init");
        if (B > 65535) fail(CSL_ERR_ARG, "at most 65535 trajectories per call");
        if (control < 0 || control > 2) fail(CSL_ERR_ARG, "control must be 0, 1 or 2");
        if (save_interval == 0 || save_interval > size_t(std::numeric_limits<int>::max()))
            fail(CSL_ERR_ARG, "save_interval must be >= 1");
        if (!(dt > 0.0) || !std::isfinite(dt) || !std::isfinite(T0) || !std::isfinite(T1))
            fail(CSL_ERR_ARG, "need finite T0, T1 and dt > 0");
        if (control != 0 && (!(abs_tol >= 0.0) || !(rel_tol >= 0.0) || !(abs_tol + rel_tol > 0.0)))
            fail(CSL_ERR_ARG, "tolerances must be >= 0 and not both 0");
        if (max_rec == 0 || max_rec > size_t(std::numeric_limits<int>::max()))
            fail(CSL_ERR_ARG, "max_rec must be in [1, 2^31)");
        const int block = csl::dopri5_block(h->L);
        if (block == 0 || csl::dopri5_lds_bytes(h->L, block) > h->lds_limit)
            fail(CSL_ERR_UNSUPPORTED, "the Dormand-Prince kernel holds at most 2048 sites per trajectory");
        std::vector<double> grid;
        if (control == 0) {
            grid = make_grid(T0, T1, dt, tgrid);
            if (grid.size() > size_t(std::numeric_limits<int>::max())) fail(CSL_ERR_ARG, "too many grid points");
        }
        std::vector<double> sg;
        std::vector<int> nsg;
        if (segments) {
            if (control != 2) fail(CSL_ERR_ARG, "segments apply to control 2");
            if (!n_segments) fail(CSL_ERR_NULL, "null n_segments");
            if (max_segments == 0 || max_segments > 4096) fail(CSL_ERR_ARG, "max_segments must be in [1, 4096]");
            check_finite(segments, B * max_segments * 2, "segments");
            sg.assign(segments, segments + B * max_segments * 2);
            nsg.resize(B);
            for (size_t b = 0; b < B; ++b) {
                if (n_segments[b] == 0 || n_segments[b] > max_segments)
                    fail(CSL_ERR_ARG, "every trajectory needs 1..max_segments segments");
                nsg[b] = int(n_segments[b]);
                for (size_t j = 0; j < n_segments[b]; ++j)
                    if (!(sg[(b * max_segments + j) * 2 + 1] > sg[(b * max_segments + j) * 2]))
                        fail(CSL_ERR_ARG, "a segment must end after it starts");
            }
        }
        const size_t nd = h->N * 3;
        const size_t n_init = init_per_replica ? B * nd : nd;
        check_finite(init, n_init, "initial state");
        hipStream_t s = h->stream;
        h->bt_init.upload(init, n_init, s);
        std::vector<double> tp(2 * B, 0.0), am(2 * B, h->amp);
        if (t_pulse) {
            check_finite(t_pulse, 2 * B, "t_pulse");
            tp.assign(t_pulse, t_pulse + 2 * B);
        }
        if (pulse_amp) {
            check_finite(pulse_amp, 2 * B, "pulse_amp");
            am.assign(pulse_amp, pulse_amp + 2 * B);
        }
        h->bt_tp.upload(tp.data(), tp.size(), s);
        h->bt_amp.upload(am.data(), am.size(), s);
        if (!grid.empty()) h->tgrid.upload(grid.data(), grid.size(), s);
        if (!sg.empty()) {
            h->dp_segs.upload(sg.data(), sg.size(), s);
            h->dp_nseg.upload(nsg.data(), nsg.size(), s);
        }
        h->dp_times.ensure(B * max_rec);
        h->rec_mags.ensure(B * max_rec * 9);
        if (energy) h->rec_energy.ensure(B * max_rec);
        if (final_states) h->bt_final.ensure(B * nd);
        h->dp_counters.ensure(B * 4);
        csl::Dopri5Args A{};
        A.B = int(B);
        A.init = h->bt_init.p;
        A.init_stride = init_per_replica ? nd : 0;
        A.t_pulse = h->bt_tp.p;
        A.amp = h->bt_amp.p;
        A.control = control;
        A.tgrid = grid.empty() ? nullptr : h->tgrid.p;
        A.n_grid = int(grid.size());
        A.segs = sg.empty() ? nullptr : h->dp_segs.p;
        A.n_segs = sg.empty() ? nullptr : h->dp_nseg.p;
        A.max_seg = int(max_segments);
        A.T0 = T0;
        A.T1 = T1;
        A.dt = dt;
        A.abs_tol = abs_tol;
        A.rel_tol = rel_tol;
        A.save_interval = int(save_interval);
        A.max_rec = int(max_rec);
        A.max_attempts = max_attempts ? (long long)std::min<size_t>(max_attempts, size_t(1) << 62) : 50000000LL;
        A.times = h->dp_times.p;
        A.mags = h->rec_mags.p;
        A.energy = energy ? h->rec_energy.p : nullptr;
        A.final_states = final_states ? h->bt_final.p : nullptr;
        A.counters = h->dp_counters.p;
        csl::LatDev L = h->L;
        if (h->rep_B) {
            if (B != h->rep_B) fail(CSL_ERR_ARG, "batch size differs from the per-replica Hamiltonians set");
            L.mats = h->rep_mats.p;
            L.field = h->rep_field.p;
            A.mats_stride = h->rep_mats_stride;
            A.field_stride = h->rep_field_stride;
        }
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        hip_check(hipEventRecord(e0, s), "event record");
        hip_check(csl::launch_dopri5(L, A, s), "k_dopri5 launch");
        hip_check(hipEventRecord(e1, s), "event record");
        hip_check(hipEventSynchronize(e1), "k_dopri5");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        h->last_ms = ms;
        h->last_block = block;
        h->last_spt = csl::dopri5_sites_per_thread(h->L);
        h->last_dbuf = 0;
        h->last_launches = 1;
        std::vector<long long> cnt(B * 4);
        hip_check(hipMemcpyAsync(cnt.data(), h->dp_counters.p, cnt.size() * sizeof(long long),
                                 hipMemcpyDeviceToHost, s), "D2H");
        copy_back(times, h->dp_times.p, B * max_rec, s);
        copy_back(mags, h->rec_mags.p, B * max_rec * 9, s);
        copy_back(energy, h->rec_energy.p, B * max_rec, s);
        copy_back(final_states, h->bt_final.p, B * nd, s);
        hip_check(hipStreamSynchronize(s), "integrate_dopri5");
        bool over = false;
        for (size_t b = 0; b < B; ++b) {
            if (n_rec_out) n_rec_out[b] = size_t(cnt[b * 4 + 0]);
            if (n_acc_out) n_acc_out[b] = size_t(cnt[b * 4 + 1]);
            if (n_rej_out) n_rej_out[b] = size_t(cnt[b * 4 + 2]);
            over |= size_t(cnt[b * 4 + 0]) > max_rec;
        }
        for (size_t b = 0; b < B; ++b) {
            const long long st = cnt[b * 4 + 3];
            if (st == csl::kDopri5StepFailed)
                fail(CSL_ERR_STEP, "trajectory " + std::to_string(b) + ": step size control failed (500 rejections)");
            if (st == csl::kDopri5AttemptCap)
                fail(CSL_ERR_STEP, "trajectory " + std::to_string(b) + ": more than max_attempts attempts");
            if (st == csl::kDopri5BadStep)
                fail(CSL_ERR_STEP, "trajectory " + std::to_string(b) + ": step size no longer advances time");
        }
        if (over) fail(CSL_ERR_CAPACITY, "max_rec is smaller than the number of records of some trajectory");
    });
}

int csl_pump_probe(csl_lattice* h, const double* ground, const double* field_in, double amp,
                   double width, double freq, double tau_start, double tau_end, double tau_step,
                   double T0, double T1, double T_step, size_t max_rec, size_t max_tau,
                   size_t* n_rec_out, size_t* n_tau_out, double* times, double* M0, double* M1,
                   double* M01, double* tau_values) {
    return csl_pump_probe_ex(h, ground, field_in, amp, width, freq, tau_start, tau_end, tau_step, T0, T1, T_step,
                             0, 0.0, 1, "rk4", 0.0, 0.0, max_rec, max_tau, n_rec_out, n_tau_out, times, M0, M1,
                             M01, tau_values, nullptr, nullptr);
}

int csl_pump_probe_ex(csl_lattice* h, const double* ground, const double* field_in, double amp,
                      double width, double freq, double tau_start, double tau_end, double tau_step,
                      double T0, double T1, double T_step, int reuse_m0_for_m1, double stationarity_tol,
                      int pulse_window_chunking, const char* method, double abs_tol, double rel_tol,
                      size_t max_rec, size_t max_tau, size_t* n_rec_out, size_t* n_tau_out, double* times,
                      double* M0, double* M1, double* M01, double* tau_values, int* used_w1, double* max_dsdt) {
    return guarded([&] {
        check_handle(h);
        const bool dp5 = method && std::strcmp(method, "dopri5") == 0;
        if (!dp5 && !method_is_rk4(method)) fail(CSL_ERR_UNSUPPORTED, "pump-probe implements \"rk4\" and \"dopri5\"");
        if (!ground || !field_in) fail(CSL_ERR_NULL, "null ground state / field_in");
        if (!(tau_step != 0.0) || !std::isfinite(tau_step)) fail(CSL_ERR_ARG, "tau_step must be non-zero");
        if (reuse_m0_for_m1 && !(stationarity_tol >= 0.0)) fail(CSL_ERR_ARG, "stationarity_tol must be >= 0");
        // tau grid (lattice_md.cpp:768; tau_start + i*tau_step as in the W2 loop :872)
        const double q = std::abs((tau_end - tau_start) / tau_step);
        if (!std::isfinite(q) || q > 1e6) fail(CSL_ERR_ARG, "tau grid too large");
        const size_t n_tau = size_t(q) + 1;
        if (n_tau_out) *n_tau_out = n_tau;
        if (n_tau > max_tau) fail(CSL_ERR_CAPACITY, "max_tau is smaller than the number of delays");
        const size_t A = h->n_atoms;
        std::vector<double> zero(2 * A * 3, 0.0);
        // W1 guard (lattice_md.cpp:729-736, max_dSdt_norm_no_drive :236-280):
        // L_inf norm of dS/dt at the ground state with the drive off.
        bool w1 = false;
        if (reuse_m0_for_m1) {
            const size_t nd = h->N * 3;
            int rc = csl_set_spins(h, ground, nd);
            if (rc == CSL_OK) rc = csl_set_pulse(h, zero.data(), 0.0, width, freq, 0.0, 0.0);
            std::vector<double> d(nd);
            if (rc == CSL_OK) rc = csl_rhs(h, 0.0, d.data(), nd);
            if (rc != CSL_OK) fail(rc, g_err);
            double mx = 0.0;
            for (double v : d) mx = std::max(mx, std::abs(v));
            if (max_dsdt) *max_dsdt = mx;
            w1 = mx <= stationarity_tol;
        }
        if (used_w1) *used_w1 = w1 ? 1 : 0;
        // Lattice::set_pulse: field_drive[p] = R_atom * field_in (lattice_md.cpp:215-216)
        std::vector<double> fd(2 * A * 3, 0.0);
        for (size_t a = 0; a < A; ++a)
            for (int mu = 0; mu < 3; ++mu) {
                double v = 0.0;
                for (int nu = 0; nu < 3; ++nu) v += h->h_frames[a * 9 + mu * 3 + nu] * field_in[a * 3 + nu];
                fd[a * 3 + mu] = v;
                fd[(A + a) * 3 + mu] = v;
            }
        const int rc = csl_set_pulse(h, fd.data(), amp, width, freq, 0.0, 0.0);
        if (rc != CSL_OK) fail(rc, g_err);
        // replicas: M0, [M1(tau) unless W1], M01(tau)
        const size_t n1 = w1 ? 0 : n_tau;
        const size_t B = 1 + n1 + n_tau;
        std::vector<double> tp(2 * B, 0.0), am(2 * B, 0.0), taus(n_tau);
        for (size_t i = 0; i < n_tau; ++i) taus[i] = tau_start + double(i) * tau_step;
        am[0] = amp;  // M0: pulse at 0 only
        for (size_t i = 0; i < n_tau; ++i) {
            if (!w1) {
                tp[2 * (1 + i)] = taus[i];  // M1(tau): pulse at tau only
                am[2 * (1 + i)] = amp;
            }
            tp[2 * (1 + n1 + i) + 1] = taus[i];  // M01(tau): pulses at 0 and tau
            am[2 * (1 + n1 + i)] = amp;
            am[2 * (1 + n1 + i) + 1] = amp;
        }
        // CPU observer gate (lattice_md.cpp:532-554) over one trajectory's records
        auto gate = [&](const double* t, size_t n) {
            std::vector<size_t> k;
            double last = T0;
            for (size_t r = 0; r < n; ++r)
                if (t[r] - last >= T_step - 1e-10 || t[r] >= T1 - 1e-10) {
                    k.push_back(r);
                    last = t[r];
                }
            return k;
        };
        std::vector<double> tt, mg;  // gated records: times [nk], mags [nk][B][9]
        std::vector<size_t> keep;
        int rc2 = CSL_OK;
        if (dp5) {
            // integrate_const with the controlled Dormand-Prince stepper
            // (single_pulse_drive / double_pulse_drive, lattice_md.cpp:562-575),
            // one workgroup and step-size sequence per replica; with W3 chunking
            // one integrate_const call per pulse-window segment of the replica
            // (windows of 9 widths around its pulse centres)
            std::vector<double> sg;
            std::vector<size_t> nsg(B, 1);
            size_t max_seg = 1, cap = 0;
            bool empty = false;
            if (pulse_window_chunking) {
                std::vector<std::vector<std::array<double, 3>>> per(B);
                for (size_t b = 0; b < B; ++b) {
                    // pulse centres: M0 {0}, M1(tau) {tau}, M01(tau) {0, tau}
                    const double c[2] = {tp[2 * b], tp[2 * b + 1]};
                    const size_t nc = b < 1 + n1 ? 1 : 2;
                    per[b] = pulse_segments(T0, T1, c, nc, kPulseWindowWidths * width, T_step, kFreeStepFactor);
                    empty |= per[b].empty();
                    max_seg = std::max(max_seg, per[b].size());
                }
                if (!empty) {
                    sg.assign(B * max_seg * 2, 0.0);
                    for (size_t b = 0; b < B; ++b) {
                        nsg[b] = per[b].size();
                        size_t calls = 0;
                        for (size_t j = 0; j < per[b].size(); ++j) {
                            sg[(b * max_seg + j) * 2] = per[b][j][0];
                            sg[(b * max_seg + j) * 2 + 1] = per[b][j][1];
                            calls += const_calls(per[b][j][0], per[b][j][1], T_step);
                        }
                        cap = std::max(cap, calls);
                    }
                }
            } else {
                cap = const_calls(T0, T1, T_step);
            }
            if (!empty) {
                std::vector<double> td(B * cap), md(B * cap * 9);
                std::vector<size_t> nrb(B);
                rc2 = csl_integrate_dopri5(h, B, ground, 0, tp.data(), am.data(), T0, T1, T_step, 2, abs_tol,
                                           rel_tol, 1, CSL_TGRID_BOOST, sg.empty() ? nullptr : sg.data(),
                                           sg.empty() ? nullptr : nsg.data(), max_seg, 0, cap, td.data(),
                                           md.data(), nullptr, nullptr, nrb.data(), nullptr, nullptr);
                if (rc2 == CSL_OK) {
                    // each replica through its own observer gate; the reference's
                    // segments keep every trajectory on the same sample count
                    std::vector<std::vector<size_t>> kb(B);
                    for (size_t b = 0; b < B; ++b) kb[b] = gate(&td[b * cap], nrb[b]);
                    const size_t nk = kb[0].size();
                    for (size_t b = 1; b < B; ++b)
                        if (kb[b].size() != nk) fail(CSL_ERR_ARG, "replicas differ in their sample counts");
                    tt.resize(nk);
                    mg.resize(nk * B * 9);
                    for (size_t j = 0; j < nk; ++j) {
                        tt[j] = td[kb[0][j]];
                        for (size_t b = 0; b < B; ++b)
                            std::memcpy(&mg[(j * B + b) * 9], &md[(b * cap + kb[b][j]) * 9], 9 * sizeof(double));
                    }
                    keep.resize(nk);
                    for (size_t j = 0; j < nk; ++j) keep[j] = j;
                }
            }
        } else {
            // RK4: chunked segments share the step and the grid, so the run is
            // the unchunked one up to the rounding of the grid times
            auto g = make_grid(T0, T1, T_step, CSL_TGRID_BOOST);
            const size_t n_rec = records_for(g.size() - 1, 1);
            tt.resize(n_rec);
            mg.resize(n_rec * B * 9);
            size_t nr = 0;
            rc2 = csl_integrate_batched(h, B, ground, 0, tp.data(), am.data(), T0, T1, T_step, 1, "rk4",
                                        CSL_TGRID_BOOST, 0, n_rec, tt.data(), mg.data(), nullptr, nullptr, &nr);
            if (rc2 == CSL_OK) keep = gate(tt.data(), nr);
        }
        // reset the pulse as the reference drivers do (lattice_md.cpp:578-583)
        (void)csl_set_pulse(h, zero.data(), 0.0, width, freq, 0.0, 0.0);
        if (rc2 != CSL_OK) fail(rc2, g_err);
        const size_t nk = keep.size();
        if (n_rec_out) *n_rec_out = nk;
        if (nk > max_rec) fail(CSL_ERR_CAPACITY, "max_rec is smaller than the number of records");
        for (size_t j = 0; j < nk; ++j) {
            const size_t r = keep[j];
            if (times) times[j] = tt[r];
            if (M0) std::memcpy(M0 + j * 9, &mg[(r * B + 0) * 9], 9 * sizeof(double));
            for (size_t i = 0; i < n_tau; ++i) {
                if (M1 && !w1)
                    std::memcpy(M1 + (i * max_rec + j) * 9, &mg[(r * B + 1 + i) * 9], 9 * sizeof(double));
                if (M01)
                    std::memcpy(M01 + (i * max_rec + j) * 9, &mg[(r * B + 1 + n1 + i) * 9], 9 * sizeof(double));
            }
        }
        if (w1 && M1 && nk > 0) {
            // synthesize_M1_from_M0 (lattice_md.cpp:283-317): M1(t_i) = M_ground
            // before tau + T_start_M0, else M0 at index lround((t_i - tau - T_start_M0) / T_step)
            // clamped to the record range.  M_ground (lattice_md.cpp:737-742) is
            // {magnetization_local_antiferro, magnetization_local,
            // magnetization_global} of the ground state: alternating sign (-1)^i
            // over the site index, plain mean, and mean of R_atom S (lattice.h:
            // 4430-4460, 4885-4902), a 9-value reduction of the host input.
            double Mg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (size_t i = 0; i < h->N; ++i) {
                const double* S = ground + i * 3;
                const double sgn = (i & 1) ? -1.0 : 1.0;
                const double* R = &h->h_frames[(i % A) * 9];
                for (int mu = 0; mu < 3; ++mu) {
                    Mg[mu] += sgn * S[mu];
                    Mg[3 + mu] += S[mu];
                    Mg[6 + mu] += R[mu * 3 + 0] * S[0] + R[mu * 3 + 1] * S[1] + R[mu * 3 + 2] * S[2];
                }
            }
            for (double& v : Mg) v /= double(h->N);
            const double t_first = tt[keep[0]];
            for (size_t i = 0; i < n_tau; ++i) {
                const double thr = taus[i] + t_first;
                for (size_t j = 0; j < nk; ++j) {
                    const double ti = tt[keep[j]];
                    const double* src = Mg;
                    if (!(ti < thr)) {
                        long long idx = std::llround((ti - taus[i] - t_first) / T_step);
                        if (idx < 0) idx = 0;
                        if (idx >= (long long)nk) idx = (long long)nk - 1;
                        src = &mg[(keep[size_t(idx)] * B + 0) * 9];
                    }
                    std::memcpy(M1 + (i * max_rec + j) * 9, src, 9 * sizeof(double));
                }
            }
        }
        if (tau_values) std::memcpy(tau_values, taus.data(), n_tau * sizeof(double));
    });
}

int csl_autotune_batched(csl_lattice* h, size_t B, size_t n_probe_steps, int* best_block,
                         double* best_ms) {
    return guarded([&] {
        check_handle(h);
        if (B == 0 || n_probe_steps == 0) fail(CSL_ERR_ARG, "need B >= 1 and n_probe_steps >= 1");
        auto cands = persistent_candidates(h);
        if (cands.empty()) fail(CSL_ERR_UNSUPPORTED, "lattice does not fit the persistent kernel");
        const size_t nd = h->N * 3;
        hipStream_t s = h->stream;
        // probe state: unit spins along a fixed direction
        std::vector<double> init(nd);
        for (size_t i = 0; i < h->N; ++i) {
            init[3 * i] = 0.6;
            init[3 * i + 1] = 0.0;
            init[3 * i + 2] = 0.8;
        }
        h->bt_init.upload(init.data(), nd, s);
        std::vector<double> tp(2 * B, 0.0), am(2 * B, h->amp);
        h->bt_tp.upload(tp.data(), tp.size(), s);
        h->bt_amp.upload(am.data(), am.size(), s);
        std::vector<double> grid(n_probe_steps + 1);
        for (size_t k = 0; k <= n_probe_steps; ++k) grid[k] = 0.01 * double(k);
        double best = 1e300;
        TuneVal bv = cands.front();
        // probe with the handle's shared tables (per-replica tables, if any,
        // have their own batch size)
        struct RepGuard {
            csl_lattice* h;
            size_t saved;
            ~RepGuard() { h->rep_B = saved; }
        } guard{h, h->rep_B};
        h->rep_B = 0;
        for (auto& c : cands) {
            double ms = 0.0;
            run_persistent(h, B, h->bt_init.p, 0, h->bt_tp.p, h->bt_amp.p, grid, 0.01, 1, true, false, c, &ms);
            run_persistent(h, B, h->bt_init.p, 0, h->bt_tp.p, h->bt_amp.p, grid, 0.01, 1, true, false, c, &ms);
            if (ms < best) {
                best = ms;
                bv = c;
            }
        }
        h->tune[batch_class(B)] = bv;
        if (best_block) *best_block = bv.block;
        if (best_ms) *best_ms = best;
    });
}

int csl_set_replica_hamiltonians(csl_lattice* h, size_t B, const double* J, const double* field) {
    return guarded([&] {
        check_handle(h);
        if (B == 0) {  // clear
            h->rep_B = 0;
            return;
        }
        if (!J || !field) fail(CSL_ERR_NULL, "null replica tables");
        const csl::LatDev& L = h->L;
        const size_t N = h->N, A = h->n_atoms, nnz = h->h_off[N];
        check_finite(J, B * nnz * 9, "replica bond matrices");
        check_finite(field, B * N * 3, "replica fields");
        const auto& off = h->h_off;
        const auto& slot = h->h_slot_of;
        const size_t ms = L.uniform ? A * size_t(L.zmax) * 9 : size_t(L.zmax) * L.NP * 9;
        const size_t fs = L.site_uniform ? A * 3 : size_t(L.NP) * 3;
        std::vector<double> mats(B * ms, 0.0), fld(B * fs, 0.0);
        std::vector<char> seen(L.uniform ? A * size_t(L.zmax) : 0);
        for (size_t b = 0; b < B; ++b) {
            const double* Jb = J + b * nnz * 9;
            const double* Fb = field + b * N * 3;
            double* mb = &mats[b * ms];
            double* fb = &fld[b * fs];
            std::fill(seen.begin(), seen.end(), 0);
            for (size_t i = 0; i < N; ++i) {
                const size_t atom = i % A, cell = i / A;
                const size_t p = atom * L.npad + cell;
                for (size_t k = off[i]; k < off[i + 1]; ++k) {
                    const size_t q = slot[k];
                    if (L.uniform) {
                        double* dst = mb + (atom * L.zmax + q) * 9;
                        if (!seen[atom * L.zmax + q]) {
                            std::memcpy(dst, Jb + k * 9, 9 * sizeof(double));
                            seen[atom * L.zmax + q] = 1;
                        } else if (std::memcmp(dst, Jb + k * 9, 9 * sizeof(double)) != 0) {
                            fail(CSL_ERR_ARG, "replica bond matrices are not translation invariant like the handle's");
                        }
                    } else {
                        std::memcpy(mb + (q * L.NP + p) * 9, Jb + k * 9, 9 * sizeof(double));
                    }
                }
                if (L.site_uniform) {
                    if (cell == 0) std::memcpy(fb + atom * 3, Fb + i * 3, 3 * sizeof(double));
                    else if (std::memcmp(fb + atom * 3, Fb + i * 3, 3 * sizeof(double)) != 0)
                        fail(CSL_ERR_ARG, "replica fields are not translation invariant like the handle's");
                } else {
                    std::memcpy(fb + p * 3, Fb + i * 3, 3 * sizeof(double));
                }
            }
        }
        h->rep_mats.upload(mats.data(), mats.size(), h->stream);
        h->rep_field.upload(fld.data(), fld.size(), h->stream);
        hip_check(hipStreamSynchronize(h->stream), "replica upload");
        h->rep_B = B;
        h->rep_mats_stride = ms;
        h->rep_field_stride = fs;
    });
}

int csl_adjoint_gradient(csl_lattice* h, size_t B, const double* init, int init_per_replica,
                         const double* t_pulse, const double* pulse_amp, double T0, double T1, double dt,
                         int tgrid, const double* target, size_t n_target, const double* weights,
                         size_t n_params, const double* dJ, const double* dB, size_t ckpt_interval,
                         double* loss, double* grad, double* records, double* grad_init) {
    return guarded([&] {
        check_handle(h);
        if (B == 0 || B > 65535) fail(CSL_ERR_ARG, "batch size must be 1 .. 65535");
        if (!init || !t_pulse || !target || !weights || !loss || !grad) fail(CSL_ERR_NULL, "null argument");
        if (n_params > 8) fail(CSL_ERR_ARG, "at most 8 parameters");
        if (n_params && (!dJ || !dB)) fail(CSL_ERR_NULL, "null parameter tangents");
        csl::LatDev L = h->L;
        if (!L.uniform) fail(CSL_ERR_UNSUPPORTED, "the adjoint needs translation-invariant bond tables");
        if (L.zmax > 6) fail(CSL_ERR_UNSUPPORTED, "the adjoint handles up to 6 bonds per site");
        if (csl::adjoint_sites_per_thread(L) == 0)
            fail(CSL_ERR_UNSUPPORTED, "lattice too large for the one-workgroup adjoint kernel");
        if (h->rep_B && h->rep_B != B) fail(CSL_ERR_ARG, "replica Hamiltonians were set for another batch size");
        auto grid = make_grid(T0, T1, dt, tgrid);
        const size_t n_steps = grid.size() - 1;
        if (n_steps == 0) fail(CSL_ERR_ARG, "empty time grid");
        if (n_steps > size_t(std::numeric_limits<int>::max() / 4)) fail(CSL_ERR_ARG, "too many steps");
        const size_t n_rec = n_steps + 1;
        if (n_target != n_rec) fail(CSL_ERR_CAPACITY, "target must hold one record per step plus the final one");
        const size_t nd = h->N * 3;
        check_finite(init, init_per_replica ? B * nd : nd, "initial state");
        check_finite(t_pulse, 2 * B, "t_pulse");
        check_finite(target, n_rec * 9, "target");
        check_finite(weights, 9, "weights");
        const size_t A_ = h->n_atoms, N = h->N, nnz = h->h_off[N];
        if (n_params) {
            check_finite(dJ, n_params * nnz * 9, "dJ");
            check_finite(dB, n_params * N * 3, "dB");
        }
        // parameter tangents -> per (sublattice, slot) / per sublattice
        std::vector<double> dm(std::max<size_t>(n_params, 1) * A_ * L.zmax * 9, 0.0),
            dfl(std::max<size_t>(n_params, 1) * A_ * 3, 0.0);
        for (size_t q = 0; q < n_params; ++q) {
            std::vector<char> seen(A_ * L.zmax, 0);
            for (size_t i = 0; i < N; ++i) {
                const size_t atom = i % A_;
                for (size_t k = h->h_off[i]; k < h->h_off[i + 1]; ++k) {
                    double* dst = &dm[((q * A_ + atom) * L.zmax + h->h_slot_of[k]) * 9];
                    const double* src = dJ + (q * nnz + k) * 9;
                    if (!seen[atom * L.zmax + h->h_slot_of[k]]) {
                        std::memcpy(dst, src, 9 * sizeof(double));
                        seen[atom * L.zmax + h->h_slot_of[k]] = 1;
                    } else if (std::memcmp(dst, src, 9 * sizeof(double)) != 0) {
                        fail(CSL_ERR_ARG, "dJ is not translation invariant like the handle's tables");
                    }
                }
                const double* fb = dB + (q * N + i) * 3;
                double* fd = &dfl[(q * A_ + atom) * 3];
                if (i < A_) std::memcpy(fd, fb, 3 * sizeof(double));
                else if (std::memcmp(fd, fb, 3 * sizeof(double)) != 0)
                    fail(CSL_ERR_ARG, "dB is not translation invariant");
            }
        }
        hipStream_t s = h->stream;
        DBuf<double> d_dm, d_df, d_tgt, d_rec, d_loss, d_grad, d_gi;
        d_dm.upload(dm.data(), dm.size(), s);
        d_df.upload(dfl.data(), dfl.size(), s);
        d_tgt.upload(target, n_rec * 9, s);
        h->bt_init.upload(init, init_per_replica ? B * nd : nd, s);
        h->bt_tp.upload(t_pulse, 2 * B, s);
        std::vector<double> am(2 * B, h->amp);
        if (pulse_amp) {
            check_finite(pulse_amp, 2 * B, "pulse_amp");
            am.assign(pulse_amp, pulse_amp + 2 * B);
        }
        h->bt_amp.upload(am.data(), am.size(), s);
        bool drive_on = false;
        for (double a : am) drive_on |= (a != 0.0);
        h->tgrid.upload(grid.data(), grid.size(), s);
        if (drive_on) {
            h->fac.ensure(n_steps * 3 * B * 2);
            hip_check(csl::launch_drive_table(h->tgrid.p, int(n_steps), dt, int(B), h->bt_tp.p, h->bt_amp.p, L.width,
                                              L.freq, h->fac.p, s),
                      "drive table");
        }
        // Stage inputs of every step in device memory when they fit (and no
        // checkpoint spacing is asked for); otherwise checkpoints every C
        // steps and one segment of stage inputs, re-integrated per segment.
        const size_t SZ = size_t(3) * L.NP;
        const size_t per_seg_step = 2 * B * 4 * SZ * sizeof(double);  // Y and U
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const size_t budget = std::min<size_t>(free_b / 2, size_t(48) << 30);
        const bool full = ckpt_interval == 0 && per_seg_step * n_steps <= budget;
        const size_t C = full ? n_steps
                              : (ckpt_interval ? std::min(ckpt_interval, n_steps)
                                               : size_t(std::ceil(std::sqrt(double(n_steps)))));
        const size_t n_seg = (n_steps + C - 1) / C;
        // kept on the handle between calls (a fit calls this every iteration)
        DBuf<double>&d_Y = h->adj_Y, &d_U = h->adj_U, &d_lam = h->adj_lam, &d_lossn = h->adj_lossn,
                     &d_part = h->adj_part, &d_ckpt = h->adj_ckpt;
        d_Y.ensure(B * C * 4 * SZ);
        d_U.ensure(B * C * 4 * SZ);
        if (!full) d_ckpt.ensure(B * n_seg * SZ);
        d_lam.ensure(B * SZ);
        d_rec.ensure(n_rec * B * 9);
        d_lossn.ensure(B * n_rec);
        d_part.ensure(B * n_steps * 4 * std::max<size_t>(n_params, 1));
        d_loss.ensure(B);
        d_grad.ensure(B * std::max<size_t>(n_params, 1));
        if (grad_init) d_gi.ensure(B * nd);
        csl::AdjArgs Aa{};
        Aa.B = int(B);
        Aa.n_steps = int(n_steps);
        Aa.P = int(n_params);
        Aa.dt = dt;
        Aa.init = h->bt_init.p;
        Aa.init_stride = init_per_replica ? nd : 0;
        Aa.fac = drive_on ? h->fac.p : nullptr;
        Aa.target = d_tgt.p;
        for (int j = 0; j < 9; ++j) Aa.w[j] = weights[j];
        Aa.dmats = d_dm.p;
        Aa.dfield = d_df.p;
        Aa.seg_len = int(C);
        Aa.n_ckpt = int(n_seg);
        Aa.ckpt = d_ckpt.p;
        Aa.Y = d_Y.p;
        Aa.U = d_U.p;
        Aa.lam = d_lam.p;
        Aa.rec = d_rec.p;
        Aa.loss_n = d_lossn.p;
        Aa.part = d_part.p;
        Aa.grad_init = grad_init ? d_gi.p : nullptr;
        Aa.loss = d_loss.p;
        Aa.grad = d_grad.p;
        if (h->rep_B) {
            L.mats = h->rep_mats.p;
            L.field = h->rep_field.p;
            Aa.mats_stride = h->rep_mats_stride;
            Aa.field_stride = h->rep_field_stride;
        }
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        hip_check(hipEventRecord(e0, s), "event record");
        size_t launches = 0;
        {   // forward: records, and the stage inputs (full) or the checkpoints
            csl::AdjArgs F = Aa;
            F.n0 = 0;
            F.n1 = int(n_steps);
            F.ckpt_src = -1;
            // with every stage input stored, the per-step records come from
            // them on the whole GPU (same arithmetic, same bits)
            F.record = full ? 2 : 1;
            F.ckpt_interval = full ? 0 : int(C);
            F.store_stages = full ? 1 : 0;
            F.y0 = 0;
            hip_check(csl::launch_adjoint_forward(L, F, s), "k_adj_fwd launch");
            ++launches;
            if (full) {
                hip_check(csl::launch_adjoint_records(L, F, s), "k_adj_rec launch");
                ++launches;
            }
        }
        if (full) {
            // the backward sweep in segments; the parameter terms of a
            // finished segment run on the second stream (on the CUs the
            // sweep leaves idle) while the sweep continues
            if (!h->stream2) hip_check(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking), "stream");
            const size_t n_part = std::min<size_t>(16, n_steps);
            const size_t len = (n_steps + n_part - 1) / n_part;
            std::vector<hipEvent_t> ev;
            for (size_t j = n_part; j-- > 0;) {
                csl::AdjArgs G = Aa;
                G.n0 = int(std::min(n_steps, j * len));
                G.n1 = int(std::min(n_steps, (j + 1) * len));
                if (G.n0 >= G.n1) continue;
                G.y0 = 0;
                G.first = G.n1 == int(n_steps) ? 1 : 0;
                hip_check(csl::launch_adjoint_backward(L, G, s), "k_adj_bwd launch");
                hipEvent_t e;
                hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
                ev.push_back(e);
                hip_check(hipEventRecord(e, s), "event record");
                hip_check(hipStreamWaitEvent(h->stream2, e, 0), "stream wait");
                // few workgroups (one per idle CU for the whole batch), so
                // that the next segment's sweep still finds room on the GPU;
                // after the last segment the whole GPU is free
                const int bpt = j == 0 ? std::max(1, 8 * h->n_cu / int(B))
                                       : std::max(1, int((h->n_cu > int(B) ? h->n_cu - int(B) : 1) / int(B)));
                hip_check(csl::launch_adjoint_params(L, G, bpt, h->stream2), "k_adj_par launch");
                launches += 2;
            }
            hipEvent_t done;
            hip_check(hipEventCreateWithFlags(&done, hipEventDisableTiming), "event");
            ev.push_back(done);
            hip_check(hipEventRecord(done, h->stream2), "event record");
            hip_check(hipStreamWaitEvent(s, done, 0), "stream wait");
            hip_check(csl::launch_adjoint_final(Aa, s), "k_adj_final launch");
            hip_check(hipEventRecord(e1, s), "event record");
            hip_check(hipEventSynchronize(e1), "adjoint");
            for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        } else {
            for (size_t j = n_seg; j-- > 0;) {
                csl::AdjArgs G = Aa;
                G.n0 = int(j * C);
                G.n1 = int(std::min(n_steps, (j + 1) * C));
                G.y0 = G.n0;
                G.ckpt_src = int(j);
                G.record = 0;
                G.ckpt_interval = 0;
                G.store_stages = 1;
                hip_check(csl::launch_adjoint_forward(L, G, s), "k_adj_fwd launch");
                G.first = j == n_seg - 1 ? 1 : 0;
                hip_check(csl::launch_adjoint_backward(L, G, s), "k_adj_bwd launch");
                hip_check(csl::launch_adjoint_params(L, G, 2048, s), "k_adj_par launch");
                launches += 3;
            }
            hip_check(csl::launch_adjoint_final(Aa, s), "k_adj_final launch");
            hip_check(hipEventRecord(e1, s), "event record");
        }
        ++launches;
        copy_back(loss, d_loss.p, B, s);
        copy_back(grad, d_grad.p, B * n_params, s);
        copy_back(records, d_rec.p, n_rec * B * 9, s);
        copy_back(grad_init, d_gi.p, B * nd, s);
        hip_check(hipStreamSynchronize(s), "adjoint");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        h->last_ms = ms;
        h->last_launches = launches;
    });
}

int csl_quench(csl_lattice* h, size_t B, double* states, size_t max_sweeps, double rel_tol, double spin_length,
               size_t* sweeps_done) {
    return guarded([&] {
        check_handle(h);
        if (B == 0 || B > 65535) fail(CSL_ERR_ARG, "batch size must be 1 .. 65535");
        if (!states) fail(CSL_ERR_NULL, "null states");
        if (!(rel_tol >= 0.0) || !(spin_length > 0.0)) fail(CSL_ERR_ARG, "need rel_tol >= 0 and spin_length > 0");
        const size_t N = h->N, nd = N * 3;
        check_finite(states, B * nd, "states");
        const csl::LatDev& L = h->L;
        // greedy colouring in site order: smallest colour unused by a neighbour
        std::vector<int> color(N, -1);
        int n_colors = 0;
        for (size_t i = 0; i < N; ++i) {
            std::vector<char> used(size_t(n_colors) + 1, 0);
            for (size_t k = h->h_off[i]; k < h->h_off[i + 1]; ++k) {
                const int c = color[h->h_par[k]];
                if (c >= 0) used[size_t(c)] = 1;
            }
            int c = 0;
            while (used[size_t(c)]) ++c;
            color[i] = c;
            n_colors = std::max(n_colors, c + 1);
        }
        std::vector<int> dcol(size_t(L.NP), -1);
        for (size_t i = 0; i < N; ++i) dcol[(i % h->n_atoms) * L.npad + i / h->n_atoms] = color[i];
        hipStream_t s = h->stream;
        DBuf<int> d_col, d_frozen, d_sweeps;
        DBuf<double> d_e, d_ep;
        d_col.upload(dcol.data(), dcol.size(), s);
        std::vector<int> zeros(B, 0);
        d_frozen.upload(zeros.data(), B, s);
        d_sweeps.upload(zeros.data(), B, s);
        d_e.ensure(B);
        d_ep.ensure(B);
        h->bt_init.upload(states, B * nd, s);
        h->X.ensure(B * 3 * size_t(L.NP));
        hip_check(hipMemsetAsync(h->X.p, 0, B * 3 * size_t(L.NP) * sizeof(double), s), "memset");
        hip_check(csl::launch_perm_to_soa(L, h->bt_init.p, nd, h->X.p, int(B), s), "perm");
        const int nblk = csl::stage_blocks(L);
        h->partials.ensure(B * size_t(nblk) * 10);
        auto energy = [&](double* out) {
            hip_check(csl::launch_observe(L, h->X.p, h->partials.p, int(B), s), "observe");
            hip_check(csl::launch_reduce_obs(h->partials.p, nblk, int(B), 0, 1.0 / double(N), nullptr, out, s),
                      "reduce");
        };
        energy(d_ep.p);
        std::vector<int> fz(B, 0);
        for (size_t sw = 0; sw < max_sweeps; ++sw) {
            for (int c = 0; c < n_colors; ++c)
                hip_check(csl::launch_quench_class(L, h->X.p, d_col.p, c, d_frozen.p, spin_length, int(B), s),
                          "quench sweep");
            energy(d_e.p);
            hip_check(csl::launch_quench_check(d_e.p, d_ep.p, d_frozen.p, d_sweeps.p, int(B), int(sw), rel_tol, s),
                      "quench check");
            if ((sw + 1) % 32 == 0 || sw + 1 == max_sweeps) {
                hip_check(hipMemcpyAsync(fz.data(), d_frozen.p, B * sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
                hip_check(hipStreamSynchronize(s), "quench");
                if (std::all_of(fz.begin(), fz.end(), [](int v) { return v != 0; })) break;
            }
        }
        h->bt_final.ensure(B * nd);
        hip_check(csl::launch_perm_to_aos(L, h->X.p, h->bt_final.p, int(B), s), "perm");
        copy_back(states, h->bt_final.p, B * nd, s);
        std::vector<int> sd(B, 0);
        hip_check(hipMemcpyAsync(sd.data(), d_sweeps.p, B * sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipStreamSynchronize(s), "quench");
        if (sweeps_done)
            for (size_t b = 0; b < B; ++b) sweeps_done[b] = size_t(sd[b]);
    });
}

int csl_structure_factor(csl_lattice* h, size_t B, const double* init, int init_per_replica, double T0,
                         double T1, double dt, int tgrid, size_t sample_interval, size_t L1, size_t L2,
                         size_t L3, const double* positions, size_t max_t, double* sqw, size_t* n_t_out,
                         double* mags) {
    return guarded([&] {
        check_handle(h);
        if (B == 0 || B > 65535) fail(CSL_ERR_ARG, "batch size must be 1 .. 65535");
        if (!init || !positions || !sqw) fail(CSL_ERR_NULL, "null argument");
        if (sample_interval == 0) fail(CSL_ERR_ARG, "sample_interval must be >= 1");
        const csl::LatDev& L = h->L;
        if (L1 * L2 * L3 != size_t(L.n_cells)) fail(CSL_ERR_ARG, "L1*L2*L3 must equal the number of unit cells");
        check_finite(positions, h->n_atoms * 3, "positions");
        auto grid = make_grid(T0, T1, dt, tgrid);
        const size_t n_steps = grid.size() - 1;
        const size_t n_t = n_steps / sample_interval;
        if (n_t_out) *n_t_out = n_t;
        if (n_t < 2) fail(CSL_ERR_ARG, "fewer than two time samples");
        const size_t Q = L1 * L2 * L3;
        if (max_t < n_t) fail(CSL_ERR_CAPACITY, "max_t is smaller than the number of time samples");
        const size_t nd = h->N * 3;
        check_finite(init, init_per_replica ? B * nd : nd, "initial state");
        hipStream_t s = h->stream;
        h->bt_init.upload(init, init_per_replica ? B * nd : nd, s);
        h->X.ensure(B * 3 * size_t(L.NP));
        hip_check(hipMemsetAsync(h->X.p, 0, B * 3 * size_t(L.NP) * sizeof(double), s), "memset");
        hip_check(csl::launch_perm_to_soa(L, h->bt_init.p, init_per_replica ? nd : 0, h->X.p, int(B), s), "perm");
        DBuf<double> pos, cin, cf, sqt, out;
        pos.upload(positions, h->n_atoms * 3, s);
        const size_t nbatch3 = B * h->n_atoms * 3;
        cin.ensure(2 * nbatch3 * Q);
        cf.ensure(2 * nbatch3 * Q);
        const size_t R = B * 3 * Q;
        sqt.ensure(2 * n_t * R);
        out.ensure(Q * n_t);
        hipfftHandle p3 = 0, p1 = 0;
        auto fft_check = [](hipfftResult r, const char* what) {
            if (r != HIPFFT_SUCCESS) fail(CSL_ERR_HIP, std::string(what) + " failed (hipFFT status " + std::to_string(int(r)) + ")");
        };
        struct PlanGuard {
            hipfftHandle& a;
            hipfftHandle& b;
            ~PlanGuard() {
                if (a) hipfftDestroy(a);
                if (b) hipfftDestroy(b);
            }
        } guard{p3, p1};
        int dims3[3] = {int(L1), int(L2), int(L3)};
        fft_check(hipfftCreate(&p3), "hipfftCreate");
        size_t ws = 0;
        fft_check(hipfftMakePlanMany(p3, 3, dims3, nullptr, 1, int(Q), nullptr, 1, int(Q), HIPFFT_Z2Z, int(nbatch3), &ws),
                  "3-D plan");
        fft_check(hipfftSetStream(p3, s), "hipfftSetStream");
        int dims1[1] = {int(n_t)};
        fft_check(hipfftCreate(&p1), "hipfftCreate");
        fft_check(hipfftMakePlanMany(p1, 1, dims1, dims1, int(R), 1, dims1, int(R), 1, HIPFFT_Z2Z, int(R), &ws),
                  "time plan");
        fft_check(hipfftSetStream(p1, s), "hipfftSetStream");
        auto* cin2 = reinterpret_cast<double2*>(cin.p);
        auto* cf2 = reinterpret_cast<double2*>(cf.p);
        auto* sqt2 = reinterpret_cast<double2*>(sqt.p);
        auto sample = [&](size_t n) {
            if (n % sample_interval || n / sample_interval >= n_t) return;
            const size_t ti = n / sample_interval;
            hip_check(csl::launch_sq_fill(L, h->X.p, int(B), cin2, s), "k_sq_fill");
            fft_check(hipfftExecZ2Z(p3, reinterpret_cast<hipfftDoubleComplex*>(cin2),
                                    reinterpret_cast<hipfftDoubleComplex*>(cf2), HIPFFT_FORWARD),
                      "3-D FFT");
            hip_check(csl::launch_sq_combine(cf2, int(B), int(h->n_atoms), int(L1), int(L2), int(L3), pos.p,
                                             sqt2 + ti * R, s),
                      "k_sq_combine");
        };
        std::vector<double> zero(2 * B, 0.0);
        h->bt_tp.upload(zero.data(), zero.size(), s);
        h->bt_amp.upload(zero.data(), zero.size(), s);
        if (h->rep_B) fail(CSL_ERR_UNSUPPORTED, "per-replica Hamiltonians need the persistent kernel");
        run_stages(h, B, h->bt_tp.p, h->bt_amp.p, false, grid, dt, 1, false, nullptr, sample);
        fft_check(hipfftExecZ2Z(p1, reinterpret_cast<hipfftDoubleComplex*>(sqt2),
                                reinterpret_cast<hipfftDoubleComplex*>(sqt2), HIPFFT_FORWARD),
                  "time FFT");
        hip_check(csl::launch_sq_power(sqt2, int(n_t), int(B), Q, 1.0 / (double(B) * double(h->N) * double(n_t)), out.p, s),
                  "k_sq_power");
        copy_back(sqw, out.p, Q * n_t, s);
        copy_back(mags, h->rec_mags.p, (n_steps + 1) * B * 9, s);
        hip_check(hipStreamSynchronize(s), "structure factor");
    });
}

int csl_set_launch_config(csl_lattice* h, size_t B, int block, int sites_per_thread, int double_buffer) {
    return guarded([&] {
        check_handle(h);
        for (const auto& c : persistent_candidates(h))
            if (c.block == block && c.spt == sites_per_thread && c.flags == double_buffer) {
                h->tune[batch_class(B)] = c;
                return;
            }
        fail(CSL_ERR_ARG, "launch shape does not fit this lattice");
    });
}

int csl_persistent_candidates(csl_lattice* h, size_t max_out, int* block, int* sites_per_thread, int* flags,
                              int* variant, size_t* n_out) {
    return guarded([&] {
        check_handle(h);
        const auto c = persistent_candidates(h);
        if (n_out) *n_out = c.size();
        if (c.size() > max_out) fail(CSL_ERR_CAPACITY, "max_out is smaller than the number of candidates");
        for (size_t i = 0; i < c.size(); ++i) {
            if (block) block[i] = c[i].block;
            if (sites_per_thread) sites_per_thread[i] = c[i].spt;
            if (flags) flags[i] = c[i].flags;
            if (variant) variant[i] = csl::persistent_variant(h->L, c[i].block, c[i].spt, c[i].flags);
        }
    });
}

int csl_last_launch_variant(csl_lattice* h, int* variant) {
    return guarded([&] {
        if (!h) fail(CSL_ERR_NULL, "null lattice handle");
        if (variant) *variant = h->last_variant;
    });
}

int csl_split_schedule(size_t B, size_t G, size_t n_steps, size_t max_rows, int* rows, size_t* pieces_per_wg) {
    return guarded([&] {
        if (G == 0 || B <= G || n_steps == 0) fail(CSL_ERR_ARG, "a split needs B > G >= 1 and n_steps >= 1");
        if (B > (size_t(1) << 31) / std::max<size_t>(n_steps, 1)) fail(CSL_ERR_ARG, "too many trajectory-steps");
        size_t max_p = 1;
        const auto r = split_schedule(B, G, n_steps, max_p);
        if (pieces_per_wg) *pieces_per_wg = max_p;
        if (G * max_p > max_rows) fail(CSL_ERR_CAPACITY, "max_rows is smaller than G * pieces_per_wg");
        if (!rows) fail(CSL_ERR_NULL, "null rows");
        std::memcpy(rows, r.data(), r.size() * sizeof(int));
    });
}

int csl_set_stage_graph(csl_lattice* h, int enable) {
    return guarded([&] {
        check_handle(h);
        h->stage_graph = enable ? 1 : 0;
    });
}

int csl_set_batch_split(csl_lattice* h, int enable) {
    return guarded([&] {
        check_handle(h);
        h->split_batches = enable ? 1 : 0;
    });
}

int csl_set_partition(csl_lattice* h, int parts) {
    return guarded([&] {
        check_handle(h);
        if (parts < -1) fail(CSL_ERR_ARG, "parts must be >= -1");
        if (parts != h->part_req) h->part.P = 0;  // rebuild the plan for the new request
        if (parts > 0) {
            if (!partition_eligible(h))
                fail(CSL_ERR_UNSUPPORTED, "the partitioned kernel needs translation-invariant tables");
            if (!ensure_partition(h, parts))
                fail(CSL_ERR_UNSUPPORTED, "the lattice does not split into " + std::to_string(parts) +
                                              " parts that fit one workgroup each");
        }
        h->part_req = parts;
    });
}

int csl_last_partition(csl_lattice* h, int* parts, int* groups) {
    return guarded([&] {
        check_handle(h);
        const bool was = (h->last_variant & csl::kVariantPartitioned) != 0;
        if (parts) *parts = was ? h->last_parts : 0;
        if (groups) *groups = was ? h->last_groups : 0;
    });
}

int csl_last_grid(csl_lattice* h, int* workgroups) {
    return guarded([&] {
        check_handle(h);
        if (workgroups) *workgroups = h->last_grid;
    });
}

int csl_last_launch_config(csl_lattice* h, int* block, int* sites_per_thread, int* double_buffer) {
    return guarded([&] {
        if (!h) fail(CSL_ERR_NULL, "null lattice handle");
        if (block) *block = h->last_block;
        if (sites_per_thread) *sites_per_thread = h->last_spt;
        if (double_buffer) *double_buffer = h->last_dbuf;
    });
}

int csl_last_kernel_ms(csl_lattice* h, double* ms, size_t* launches) {
    return guarded([&] {
        if (!h) fail(CSL_ERR_NULL, "null lattice handle");
        if (ms) *ms = h->last_ms;
        if (launches) *launches = h->last_launches;
    });
}

}  // extern "C"
